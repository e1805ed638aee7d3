"""Host-side system set-up: lattices, FCIDUMP integrals, determinants, hashing.

These are the host responsibilities NECI keeps outside the iteration
(SysInit/IntInit/DetInit, NECICore.F90:292-365): reading integrals, building
the lattice, choosing the reference, RandomHash2 and pSingles.
"""
import os

import numpy as np

# --------------------------------------------------------------- bit layout


def nwords(nbasis):
    """NIfD + 1 = int(nBasis/64) + 1 words per determinant (BitReps.F90:174)."""
    return nbasis // 64 + 1


def encode_det(orbs, nbasis):
    """EncodeBitDet: orbital k -> word (k-1)//64, bit (k-1)%64 (macros.h:16,74-77)."""
    w = np.zeros(nwords(nbasis), dtype=np.uint64)
    for o in orbs:
        w[(o - 1) >> 6] |= np.uint64(1) << np.uint64((o - 1) & 63)
    return w.view(np.int64)


def decode_det(words):
    """Occupied orbitals (1-based, ascending) of one determinant given as its
    W words, in any shape that holds exactly W words ((W,) or (1, W))."""
    u = np.ascontiguousarray(words).view(np.uint64).ravel()
    out = []
    for wi, x in enumerate(u):
        x = int(x)
        while x:
            b = (x & -x).bit_length() - 1
            out.append(64 * wi + b + 1)
            x &= x - 1
    return out


def umat_ind(i, j, k, l):
    """UMatInd_base (UMatCache.F90:257-303), real hermitian, 1-based."""
    a = i * (i - 1) // 2 + k if i > k else k * (k - 1) // 2 + i
    b = j * (j - 1) // 2 + l if j > l else l * (l - 1) // 2 + j
    return a * (a - 1) // 2 + b if a > b else b * (b - 1) // 2 + a


def umat_len(nspat):
    npair = nspat * (nspat + 1) // 2
    return npair * (npair + 1) // 2


# --------------------------------------------------------------- Hubbard


class HubbardChain:
    """Real-space Hubbard chain (init_real_space_hubbard, real_space_hubbard.F90:140-283).

    TMAT2D holds bhub between same-spin orbitals of neighbouring sites
    (real_space_hubbard.F90:984-992); the diagonal is U times the number of
    doubly occupied sites (:2303-2321).
    """

    def __init__(self, nsites, nel, u, bhub=-1.0, periodic=True):
        self.nsites = nsites
        self.nel = nel
        self.nbasis = 2 * nsites
        self.uhub = float(u)
        self.bhub = float(bhub)
        self.periodic = periodic
        nbr = np.zeros((nsites, 2), dtype=np.int32)
        for s in range(1, nsites + 1):
            cand = []
            if s > 1:
                cand.append(s - 1)
            elif periodic and nsites > 2:
                cand.append(nsites)
            if s < nsites:
                cand.append(s + 1)
            elif periodic and nsites > 2:
                cand.append(1)
            cand = list(dict.fromkeys(cand))
            nbr[s - 1, :len(cand)] = cand
        self.nbr = nbr

    def neel_state(self):
        """create_neel_state_chain (lattice_models_utils.F90:1300-1308)."""
        st = list(range(1, 2 * self.nel, 2))
        for i in range(1, self.nel, 2):
            st[i] += 1
        return st

    def tmat(self):
        t = np.zeros((self.nbasis, self.nbasis))
        for s in range(1, self.nsites + 1):
            for n in self.nbr[s - 1]:
                if n:
                    t[2 * s - 2, 2 * n - 2] = self.bhub
                    t[2 * s - 1, 2 * n - 1] = self.bhub
        return t


# ------------------------------------------------------------ molecular


class MolecularSystem:
    """Spin-restricted molecular Hamiltonian as READFCIINT leaves it.

    tmat: TMAT2D over spin orbitals (nbasis x nbasis);
    umat: UMAT packed by UMatInd over spatial orbitals (UMAT(1) at index 0);
    ecore: ECORE.  Chemist (pq|rs) = UMAT(UMatInd(p, r, q, s)) (readint.F90:889).
    """

    def __init__(self, nspat, nel, h1, eri, ecore=0.0, ms2=0):
        self.nspat = nspat
        self.nel = nel
        self.nbasis = 2 * nspat
        self.ms2 = ms2
        self.h1 = np.asarray(h1, dtype=np.float64)
        self.eri = eri  # dense chemist (pq|rs) or None
        self.ecore = float(ecore)
        self.tmat = np.zeros((self.nbasis, self.nbasis))
        for p in range(nspat):
            for q in range(nspat):
                self.tmat[2 * p, 2 * q] = self.h1[p, q]
                self.tmat[2 * p + 1, 2 * q + 1] = self.h1[p, q]
        self.umat = None
        if eri is not None:
            self.umat = pack_umat(eri)

    def tmat_colmajor(self):
        return np.asfortranarray(self.tmat).ravel(order="F")

    def aufbau_reference(self):
        """HF determinant: lowest orbitals by diagonal one-electron energy, nel/2 per spin."""
        order = np.argsort(np.diag(self.h1), kind="stable")
        nb = (self.nel - self.ms2) // 2
        na = self.nel - nb
        orbs = [2 * (p + 1) - 1 for p in order[:nb]] + [2 * (p + 1) for p in order[:na]]
        return sorted(orbs)


def pack_umat(eri):
    """UMAT(UMatInd(i,k,j,l)) = (ij|kl) for a dense 8-fold symmetric chemist array."""
    n = eri.shape[0]
    u = np.zeros(umat_len(n))
    for i in range(1, n + 1):
        for j in range(1, i + 1):
            for k in range(1, n + 1):
                for l in range(1, k + 1):
                    if i * (i - 1) // 2 + j < k * (k - 1) // 2 + l:
                        continue
                    u[umat_ind(i, k, j, l) - 1] = eri[i - 1, j - 1, k - 1, l - 1]
    return u


def synthetic_molecule(nspat, nel, rank, seed):
    """Synthetic C1 Hamiltonian: h_pq symmetric with diagonal = sorted U[-20,-0.5]
    and off-diagonal N(0, 0.05); (pq|rs) = sum_P L^P_pq L^P_rs with L^P symmetric,
    L^P_pq ~ N(0,1)*0.3*exp(-|p-q|/4).  numpy PCG64 stream from seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    h = rng.normal(0.0, 0.05, size=(nspat, nspat))
    h = np.triu(h, 1)
    h = h + h.T
    h[np.diag_indices(nspat)] = np.sort(rng.uniform(-20.0, -0.5, size=nspat))
    idx = np.arange(nspat)
    damp = 0.3 * np.exp(-np.abs(idx[:, None] - idx[None, :]) / 4.0)
    a = rng.normal(0.0, 1.0, size=(rank, nspat, nspat)) * damp[None]
    lp = 0.5 * (a + a.transpose(0, 2, 1))
    eri = np.einsum("Ppq,Prs->pqrs", lp, lp, optimize=True)
    return MolecularSystem(nspat, nel, h, eri, ecore=0.0)


def freeze_core(mol, nfrozen=1, orbitals=None):
    """The FREEZE keyword's core part (FREEZE n 0 freezes n = 2*nfrozen spin
    orbitals; IntFREEZEBASIS, Integrals_neci.F90:1083-1140 for ECORE and
    :1228-1340 for TMAT): the frozen spatial orbitals are doubly
    occupied in every determinant and leave the basis.  Their mean field moves
    into the one-electron integrals and the core energy:
      h'_pq = h_pq + sum_c [2 (pq|cc) - (pc|cq)],
      E'    = E_core + sum_c 2 h_cc + sum_cd [2 (cc|dd) - (cd|dc)],
    so every Hamiltonian element between determinants with the frozen
    orbitals doubly occupied is unchanged.  The frozen orbitals are the
    nfrozen lowest by h_pp (the reference orders by orbital energy; the two
    agree for core orbitals), or the 0-based spatial indices in orbitals.
    Returns a MolecularSystem of nspat - nfrozen orbitals and
    nel - 2 nfrozen electrons (orbsym kept for the remaining orbitals)."""
    n = mol.nspat
    if orbitals is None:
        core = [int(c) for c in np.argsort(np.diag(mol.h1), kind="stable")[:nfrozen]]
    else:
        core = sorted(int(c) for c in orbitals)
    if len(set(core)) != len(core) or any(c < 0 or c >= n for c in core) or 2 * len(core) > mol.nel:
        raise ValueError(f"cannot freeze orbitals {core} of {n} with {mol.nel} electrons")
    act = [p for p in range(n) if p not in core]

    def eri(p, q, r, s):  # chemist (pq|rs), 0-based
        if mol.eri is not None:
            return mol.eri[p, q, r, s]
        return mol.umat[umat_ind(p + 1, r + 1, q + 1, s + 1) - 1]

    e = mol.ecore
    for c in core:
        e += 2.0 * mol.h1[c, c]
        for d in core:
            e += 2.0 * eri(c, c, d, d) - eri(c, d, d, c)
    na = len(act)
    h = np.zeros((na, na))
    for a, p in enumerate(act):
        for b, q in enumerate(act):
            v = mol.h1[p, q]
            for c in core:
                v += 2.0 * eri(p, q, c, c) - eri(p, c, c, q)
            h[a, b] = v
    u = np.zeros(umat_len(na))
    for i in range(1, na + 1):
        for j in range(1, i + 1):
            for k in range(1, na + 1):
                for l in range(1, k + 1):
                    if i * (i - 1) // 2 + j < k * (k - 1) // 2 + l:
                        continue
                    u[umat_ind(i, k, j, l) - 1] = eri(act[i - 1], act[j - 1], act[k - 1], act[l - 1])
    out = MolecularSystem(na, mol.nel - 2 * len(core), h, None, ecore=e, ms2=mol.ms2)
    out.umat = u
    if getattr(mol, "orbsym", None) is not None:
        out.orbsym = np.asarray(mol.orbsym)[act]
        out.isym = getattr(mol, "isym", 1)
    return out


# --------------------------------------------------------------- FCIDUMP


def write_fcidump(path, mol, thresh=1e-12):
    """Molpro/NECI FCIDUMP: &FCI namelist then 'Z I J K L' lines (chemist order)."""
    n = mol.nspat
    with open(path, "w") as f:
        f.write(f" &FCI NORB={n},NELEC={mol.nel},MS2={mol.ms2},\n")
        f.write("  ORBSYM=" + ",".join(["1"] * n) + ",\n")
        f.write("  ISYM=1,\n &END\n")
        eri = mol.eri
        for i in range(1, n + 1):
            for j in range(1, i + 1):
                for k in range(1, n + 1):
                    for l in range(1, k + 1):
                        if i * (i - 1) // 2 + j < k * (k - 1) // 2 + l:
                            continue
                        z = eri[i - 1, j - 1, k - 1, l - 1]
                        if abs(z) > thresh:
                            f.write(f" {z: .16e} {i:4d} {j:4d} {k:4d} {l:4d}\n")
        for i in range(1, n + 1):
            for j in range(1, i + 1):
                z = mol.h1[i - 1, j - 1]
                if abs(z) > thresh:
                    f.write(f" {z: .16e} {i:4d} {j:4d} {0:4d} {0:4d}\n")
        f.write(f" {mol.ecore: .16e} {0:4d} {0:4d} {0:4d} {0:4d}\n")


def read_fcidump(path, umat_eps=1e-8):
    """READFCIINT (readint.F90:585-950) through the engine library's reader
    (ngpu_fcidump_header + ngpu_read_fcidump): spin-restricted real integrals
    as the reference stores them, UMAT(UMatInd(I,K,J,L)) = Z, TMAT2D(2I-s, 2J-s) = Z
    for both spins, ECORE from the 0 0 0 0 line, |Z| < UMatEps skipped.  Returns
    a MolecularSystem with tmat/umat filled, plus .orbsym, .isym and .n_two
    (the reference's NonZeroInt)."""
    import ctypes
    from ._lib import check, load_library, ptr
    lib = load_library()
    bpath = os.fsencode(path)
    norb, nelec, ms2, isym = (ctypes.c_int32() for _ in range(4))
    check(lib.ngpu_fcidump_header(bpath, ctypes.byref(norb), ctypes.byref(nelec), ctypes.byref(ms2),
                                  ctypes.byref(isym), None, 0), "ngpu_fcidump_header")
    n = norb.value
    orbsym = np.zeros(n, dtype=np.int32)
    check(lib.ngpu_fcidump_header(bpath, None, None, None, None, ptr(orbsym), n), "ngpu_fcidump_header")
    nbasis = 2 * n
    tmat = np.zeros((nbasis, nbasis))
    umat = np.zeros(umat_len(n))
    ecore = ctypes.c_double()
    n_two = ctypes.c_int64()
    check(lib.ngpu_read_fcidump(bpath, umat_eps, ptr(tmat), nbasis, ptr(umat), len(umat), ctypes.byref(ecore),
                                ctypes.byref(n_two)), "ngpu_read_fcidump")
    h1 = tmat[0::2, 0::2].copy()
    mol = MolecularSystem(n, nelec.value, h1, None, ecore=ecore.value, ms2=ms2.value)
    mol.tmat = tmat
    mol.umat = umat
    mol.orbsym = orbsym
    mol.isym = isym.value
    mol.n_two = n_two.value
    return mol


# ------------------------------------------------------------- hashing etc.


def random_hash2(nbasis, seed):
    """RandomHash2 (fcimc_initialisation.fpp:890-915): distinct integers
    INT(nBasis*r*1000)+1 drawn by rejection; r from numpy PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    seen = set()
    while len(out) < nbasis:
        v = int(nbasis * rng.random() * 1000) + 1
        if v in seen:
            continue
        seen.add(v)
        out.append(v)
    return np.array(out, dtype=np.int32)


def find_walker_hash(orbs, rh, nbuckets):
    """FindWalkerHash (hash.F90:23-37) in Python integers with int64 wrap."""
    h = 0
    for i, o in enumerate(orbs, start=1):
        h = (1099511628211 * h + int(rh[o - 1]) * i) & 0xFFFFFFFFFFFFFFFF
    if h >= 1 << 63:
        h -= 1 << 64
    m = abs(h) % nbuckets
    return m + 1


def count_ref_excitations(ref_orbs, nbasis):
    """Spin-allowed singles and doubles from the reference (C1 symmetry)."""
    occ = set(ref_orbs)
    virt = [o for o in range(1, nbasis + 1) if o not in occ]
    nsing = sum(1 for i in occ for a in virt if (i & 1) == (a & 1))
    oa = [o for o in occ if o % 2 == 0]
    ob = [o for o in occ if o % 2 == 1]
    va = [o for o in virt if o % 2 == 0]
    vb = [o for o in virt if o % 2 == 1]

    def c2(x):
        return x * (x - 1) // 2

    ndoub = c2(len(oa)) * c2(len(va)) + c2(len(ob)) * c2(len(vb)) + len(oa) * len(ob) * len(va) * len(vb)
    return nsing, ndoub


def psingles_from_reference(ref_orbs, nbasis):
    """CalcApproxpDoubles (fcimc_initialisation.fpp:3470-3570): pSingles =
    nSingles / (nSingles + nDoubles) counted from the reference."""
    ns, nd = count_ref_excitations(ref_orbs, nbasis)
    if ns == 0 or nd == 0:
        return 0.05
    return ns / float(ns + nd)
