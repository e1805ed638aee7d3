"""Independent Hamiltonians for tests: second quantisation on bitstrings.

Used as the mathematical oracle for energies and matrix elements: H is built
from the operator definition (hopping / one- and two-body terms with
Jordan-Wigner signs in ascending spin-orbital order, odd = beta), without any
of the engine's or the CPU oracle's Slater-Condon code.
"""
import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def _popcount(x):
    return bin(x).count("1")


def hubbard_dets(nsites, nalpha, nbeta):
    out = []
    for a in itertools.combinations(range(1, nsites + 1), nalpha):
        for b in itertools.combinations(range(1, nsites + 1), nbeta):
            bits = 0
            for s in a:
                bits |= 1 << (2 * s - 1)      # alpha = orbital 2s -> bit 2s-1
            for s in b:
                bits |= 1 << (2 * s - 2)      # beta  = orbital 2s-1 -> bit 2s-2
            out.append(bits)
    return sorted(out)


def hubbard_hamiltonian(nsites, nbr, u, bhub, dets):
    """H = bhub * sum_<ij>,s c+_is c_js + U sum_i n_i,up n_i,dn over the given dets."""
    index = {d: i for i, d in enumerate(dets)}
    rows, cols, vals = [], [], []
    for i, d in enumerate(dets):
        nd = _popcount(d & (d >> 1) & 0x5555555555555555)
        rows.append(i); cols.append(i); vals.append(u * nd)
        for s in range(1, nsites + 1):
            for n in nbr[s - 1]:
                if n == 0:
                    continue
                for spin in (0, 1):        # 0 beta, 1 alpha
                    p = 2 * s - 2 + spin   # bit of orbital on site s
                    q = 2 * n - 2 + spin   # bit on neighbour n
                    # c+_p c_q |d>
                    if not (d >> q) & 1 or (d >> p) & 1:
                        continue
                    lo, hi = min(p, q), max(p, q)
                    between = _popcount(d & (((1 << hi) - 1) ^ ((1 << (lo + 1)) - 1)))
                    nd2 = (d & ~(1 << q)) | (1 << p)
                    rows.append(index[nd2]); cols.append(i)
                    vals.append(bhub * (-1.0) ** between)
    n = len(dets)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def ground_state(h):
    if h.shape[0] <= 400:
        w = np.linalg.eigvalsh(h.toarray())
        return float(w[0])
    w = sla.eigsh(h, k=1, which="SA", tol=1e-12)[0]
    return float(w[0])


def apply_excitation_op(d, ops):
    """Apply a product of creation/annihilation operators (rightmost first).
    ops: list of (orbital_bit, is_creation). Returns (sign, new_bits) or (0, None)."""
    sign = 1
    for bit, cre in reversed(ops):
        occ = (d >> bit) & 1
        if cre == bool(occ):
            return 0, None
        sign *= (-1) ** _popcount(d & ((1 << bit) - 1))
        d ^= 1 << bit
    return sign, d


def molecular_hamiltonian_dense(h1, eri, ecore, dets):
    """<I|H|J> from H = sum_pq h_pq a+_p a_q + 1/2 sum_pqrs (pq|rs) a+_p a+_r a_s a_q
    over spin orbitals (spatial integrals, spin conserved per electron)."""
    n = h1.shape[0]
    nb = 2 * n
    index = {d: i for i, d in enumerate(dets)}
    H = np.zeros((len(dets), len(dets)))
    for j, d in enumerate(dets):
        H[j, j] += ecore
        for P in range(nb):
            for Q in range(nb):
                if P % 2 != Q % 2:
                    continue
                v = h1[P // 2, Q // 2]
                if v == 0.0:
                    continue
                s, nd = apply_excitation_op(d, [(P, True), (Q, False)])
                if s:
                    H[index[nd], j] += v * s
        for P, Q, R, S in itertools.product(range(nb), repeat=4):
            if P % 2 != Q % 2 or R % 2 != S % 2:
                continue
            v = eri[P // 2, Q // 2, R // 2, S // 2]
            if v == 0.0:
                continue
            s, nd = apply_excitation_op(d, [(P, True), (R, True), (S, False), (Q, False)])
            if s:
                H[index[nd], j] += 0.5 * v * s
    return H


def all_dets(nbasis, nel, nalpha=None):
    out = []
    for occ in itertools.combinations(range(nbasis), nel):
        bits = 0
        for b in occ:
            bits |= 1 << b
        if nalpha is not None and _popcount(bits & 0xAAAAAAAAAAAAAAAA) != nalpha:
            continue
        out.append(bits)
    return out
