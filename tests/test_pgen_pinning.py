"""Generation probabilities of the oracle's excitation generators, pinned
independently of the oracle's own code (CPU only).

For every excitation a generator can produce from a set of determinants of
small systems, the oracle's calc_pgen must equal a closed form evaluated here
in numpy, to 1e-12 relative, and the pgen summed over all valid excitations
must equal 1 - p_null, where p_null is the probability of the generator's
null outcomes, evaluated separately along its decision tree:

* PCHB doubles (GAS_doubles_PCHB_spinorb_gen_exc and get_pgen,
  gasci_pchb_doubles_spinorb_fastweighted.fpp:131-268): pgen = pDoubles *
  p(ij) * |<ij||ab>| / sum_{a'b'} |<ij||a'b'>|, the sum over every
  spin-conserving pair a' < b' outside {i, j}, with
  - UNIF-UNIF particles (pick_biased_elecs, excit_gens_int_weighted.F90:722):
    p(ij) = pParallel / (n_aa + n_bb) for a parallel pair, else
    (1 - pParallel) / (n_a n_b);
  - UNIF-FAST particles (draw_PC_FastWeightedParticles_t,
    gasci_pchb_doubles_select_particles.fpp:515-560): p(ij) =
    (q(j|i) + q(i|j)) / nel with q(j|i) = w_ij / sum_j' w_ij' and
    w_ij = sum_{ab} |<ij||ab>| (weights over all orbitals j');
* heat-bath singles (gen_single_4ind_ex / pgen_single_4ind,
  excit_gens_int_weighted.F90:442-491): pgen = pSingles / nel *
  |H_ia| / sum_b |H_ib| over the unoccupied orbitals b of i's spin;
* the uniform generator (gen_rand_excit, C1, no forbidden orbitals):
  singles pSingles / (nel u_s), parallel doubles 2 pDoubles / (P u_s (u_s - 1)),
  opposite-spin doubles pDoubles (1/u_a + 1/u_b) / (P (n_virt)) with
  P = nel (nel - 1) / 2;
* the Hubbard generator (gen_excit_rs_hubbard, real_space_hubbard.F90:1938):
  pgen = 1 / (nel * empty same-spin neighbours).

The matrix elements come from second quantisation on bitstrings
(tests/ed_reference.py's operator algebra), not from Slater-Condon rules.
"""
import itertools

import numpy as np
import pytest

import ed_reference as ed
import oracle_ctypes as oc
from common import molecule
from neci_stable_amd import HubbardChain, Spec, encode_det

RTOL = 1e-12


def bits_of(orbs):
    b = 0
    for o in orbs:
        b |= 1 << (o - 1)
    return b


def sq_element(h1, eri, orbs_i, orbs_j):
    """<I|H|J> by second quantisation: H = sum h_pq a+_p a_q +
    1/2 sum (pq|rs) a+_p a+_r a_s a_q over spin orbitals (bit o-1 = orbital o,
    odd orbitals beta), restricted to the operator strings that can map J
    onto I (creations inside I, annihilations inside J)."""
    bi, bj = bits_of(orbs_i), bits_of(orbs_j)
    oi = [o - 1 for o in orbs_i]
    oj = [o - 1 for o in orbs_j]
    v = 0.0
    for P in oi:
        for Q in oj:
            if P % 2 != Q % 2:
                continue
            s, nd = ed.apply_excitation_op(bj, [(P, True), (Q, False)])
            if s and nd == bi:
                v += s * h1[P // 2, Q // 2]
    for P, R in itertools.product(oi, repeat=2):
        for Q, S in itertools.product(oj, repeat=2):
            if P % 2 != Q % 2 or R % 2 != S % 2:
                continue
            s, nd = ed.apply_excitation_op(bj, [(P, True), (R, True), (S, False), (Q, False)])
            if s and nd == bi:
                v += 0.5 * s * eri[P // 2, Q // 2, R // 2, S // 2]
    return v


def pair_weight(mol, i, j, a, b):
    """|<ij||ab>|: the two-body element between the two-electron
    determinants {i, j} and {a, b} (determinant independent for doubles)."""
    if len({i, j, a, b}) < 4 or (i % 2) + (j % 2) != (a % 2) + (b % 2):
        return 0.0
    zero = np.zeros_like(mol.h1)
    return abs(sq_element(zero, mol.eri, sorted([a, b]), sorted([i, j])))


def pchb_tables(mol, nb):
    """W[(i, j)] = {(a, b): |<ij||ab>|} for i < j, a < b, and the pair sums."""
    W = {}
    for i, j in itertools.combinations(range(1, nb + 1), 2):
        row = {}
        for a, b in itertools.combinations(range(1, nb + 1), 2):
            w = pair_weight(mol, i, j, a, b)
            if w > 0.0:
                row[(a, b)] = w
        W[(i, j)] = row
    return W


def dets_with(nb, nel, nalpha, n, seed):
    """n distinct determinants of nel electrons (nalpha of them alpha = even
    orbitals), the aufbau one first."""
    rng = np.random.default_rng(seed)
    alpha = list(range(2, nb + 1, 2))
    beta = list(range(1, nb + 1, 2))
    out = [sorted(alpha[:nalpha] + beta[:nel - nalpha])]
    while len(out) < n:
        d = sorted([int(x) for x in rng.choice(alpha, nalpha, replace=False)] +
                   [int(x) for x in rng.choice(beta, nel - nalpha, replace=False)])
        if d not in out:
            out.append(d)
    return out


def excitations(orbs, nb):
    """Every spin-conserving single and double of orbs: (ic, ex4)."""
    occ = set(orbs)
    virt = [o for o in range(1, nb + 1) if o not in occ]
    out = []
    for i in orbs:
        for a in virt:
            if i % 2 == a % 2:
                out.append((1, [i, a, 0, 0]))
    for i, j in itertools.combinations(orbs, 2):
        for a, b in itertools.combinations(virt, 2):
            if (i % 2) + (j % 2) == (a % 2) + (b % 2):
                out.append((2, [i, j, a, b]))
    return out


def heatbath_single(mol, orbs, nb, i, a):
    """(|H_ia|, sum_b |H_ib|) over the unoccupied b of i's spin."""
    occ = set(orbs)
    tot = 0.0
    hia = 0.0
    for b in range(1, nb + 1):
        if b in occ or b % 2 != i % 2:
            continue
        tgt = sorted((occ - {i}) | {b})
        h = abs(sq_element(mol.h1, mol.eri, tgt, orbs))
        tot += h
        if b == a:
            hia = h
    return hia, tot


@pytest.mark.parametrize("particles", [0, 1])
@pytest.mark.parametrize("nspat,nel,nalpha", [(5, 4, 2), (6, 5, 3)])
def test_pchb_and_heatbath_pgen_closed_forms(particles, nspat, nel, nalpha):
    mol = molecule(nspat, nel, rank=6, seed=11)
    nb = 2 * nspat
    ps, ppar = 0.3, 0.4
    spec = Spec(mol, init_walkers=100, particles=particles, psingles=ps, pparallel=ppar)
    o = oc.Oracle(spec, 0.01)
    W = pchb_tables(mol, nb)
    wsum = {k: sum(v.values()) for k, v in W.items()}
    # weights of the FAST particle selection: w_ij over all orbitals j
    wij = np.zeros((nb + 1, nb + 1))
    for (i, j), s in wsum.items():
        wij[i, j] = wij[j, i] = s
    for orbs in dets_with(nb, nel, nalpha, 6, seed=nspat):
        det = encode_det(orbs, nb)
        occ = set(orbs)
        na = sum(1 for x in orbs if x % 2 == 0)
        nbe = nel - na
        npar = na * (na - 1) // 2 + nbe * (nbe - 1) // 2

        def p_pair(i, j):
            if particles == 0:
                return ppar / npar if i % 2 == j % 2 else (1.0 - ppar) / (na * nbe)
            q_ji = wij[i, j] / wij[i].sum() if wij[i].sum() > 0 else 0.0
            q_ij = wij[j, i] / wij[j].sum() if wij[j].sum() > 0 else 0.0
            return (q_ji + q_ij) / nel

        total = 0.0
        for ic, ex in excitations(orbs, nb):
            pg = o.calc_pgen(det, ic, np.array(ex, dtype=np.int32))
            if ic == 1:
                hia, tot = heatbath_single(mol, orbs, nb, ex[0], ex[1])
                expect = ps / nel * hia / tot if tot > 0 else 0.0
            else:
                i, j, a, b = ex
                w = W[(i, j)].get((a, b), 0.0)
                expect = (1.0 - ps) * p_pair(i, j) * (w / wsum[(i, j)] if wsum[(i, j)] > 0 else 0.0)
            assert abs(pg - expect) <= RTOL * max(expect, 1e-300), (orbs, ic, ex, pg, expect)
            total += pg
        # null outcomes along the generator's decision tree
        p_null = 0.0
        for i in orbs:  # singles: an electron with no connected target
            if heatbath_single(mol, orbs, nb, i, 0)[1] == 0.0:
                p_null += ps / nel
        if particles == 0:
            pairs = [(i, j) for i, j in itertools.combinations(orbs, 2)]
            if npar == 0:
                p_null += (1.0 - ps) * ppar
            if na * nbe == 0:
                p_null += (1.0 - ps) * (1.0 - ppar)
            pp = {(i, j): p_pair(i, j) for i, j in pairs if (i % 2 == j % 2 and npar) or
                  (i % 2 != j % 2 and na * nbe)}
        else:
            # first electron uniform, second from q(.|i) over all orbitals:
            # null when it lands on an empty orbital
            p_ok = sum(wij[i, j] / wij[i].sum() for i in orbs for j in orbs
                       if j != i and wij[i].sum() > 0) / nel
            p_null += (1.0 - ps) * (1.0 - p_ok)
            pp = {(i, j): p_pair(i, j) for i, j in itertools.combinations(orbs, 2)}
        for (i, j), p in pp.items():  # holes: a pair of weight zero, or an occupied hole
            if wsum[(i, j)] == 0.0:
                p_null += (1.0 - ps) * p
                continue
            occ_w = sum(w for (a, b), w in W[(i, j)].items() if a in occ or b in occ)
            p_null += (1.0 - ps) * p * occ_w / wsum[(i, j)]
        assert total == pytest.approx(1.0 - p_null, rel=0, abs=1e-12), (orbs, total, p_null)


@pytest.mark.parametrize("nspat,nel,nalpha", [(5, 4, 2), (6, 6, 3), (7, 5, 2)])
def test_uniform_generator_pgen_closed_form(nspat, nel, nalpha):
    mol = molecule(nspat, nel, rank=6, seed=11)
    nb = 2 * nspat
    ps = 0.3
    o = oc.Oracle(Spec(mol, init_walkers=100, excitgen="uniform", psingles=ps), 0.01)
    P = nel * (nel - 1) // 2
    for orbs in dets_with(nb, nel, nalpha, 5, seed=nspat + 1):
        det = encode_det(orbs, nb)
        na = sum(1 for x in orbs if x % 2 == 0)
        u = {0: nspat - na, 1: nspat - (nel - na)}  # unoccupied alpha (even), beta (odd)
        assert min(u.values()) >= 2  # no forbidden orbitals for these determinants
        total = 0.0
        for ic, ex in excitations(orbs, nb):
            pg = o.calc_pgen(det, ic, np.array(ex, dtype=np.int32))
            if ic == 1:
                expect = ps / (nel * u[ex[0] % 2])
            elif ex[0] % 2 == ex[1] % 2:
                us = u[ex[0] % 2]
                expect = 2.0 * (1.0 - ps) / (P * us * (us - 1))
            else:
                expect = (1.0 - ps) * (1.0 / u[0] + 1.0 / u[1]) / (P * (nb - nel))
            assert abs(pg - expect) <= RTOL * expect, (orbs, ex, pg, expect)
            total += pg
        assert total == pytest.approx(1.0, rel=0, abs=1e-12)  # this generator has no null outcome here


def test_hubbard_generator_pgen_closed_form():
    chain = HubbardChain(6, 6, 4.0, bhub=-1.0)
    o = oc.Oracle(Spec(chain, init_walkers=100), 0.01)
    nb = 12
    for bits in ed.hubbard_dets(6, 3, 3)[::7]:
        orbs = [b + 1 for b in range(nb) if (bits >> b) & 1]
        det = encode_det(orbs, nb)
        occ = set(orbs)
        total, p_null = 0.0, 0.0
        for i in orbs:
            site = (i + 1) // 2
            empty = [2 * n - (i % 2) for n in chain.nbr[site - 1] if n and (2 * n - (i % 2)) not in occ]
            if not empty:
                p_null += 1.0 / 6
            for a in empty:
                pg = o.calc_pgen(det, 1, np.array([i, a, 0, 0], dtype=np.int32))
                expect = 1.0 / (6 * len(empty))
                assert abs(pg - expect) <= RTOL * expect
                total += pg
        assert total == pytest.approx(1.0 - p_null, rel=0, abs=1e-12)


@pytest.mark.parametrize("nspat,nel,nalpha", [(5, 4, 2), (6, 5, 3), (7, 6, 3)])
def test_heatbath_single_element_is_the_exact_element(nspat, nel, nalpha):
    """The reference-relative weight a heat-bath single is drawn with and
    spawns with (the oracle's hb_weight, which the device restates term for
    term) equals |<D_i^a|H|D>| from second quantisation to 1e-12, for every
    single excitation of determinants up to several excitations away from the
    reference."""
    mol = molecule(nspat, nel, rank=6, seed=13)
    nb = 2 * nspat
    spec = Spec(mol, init_walkers=100, particles=0, psingles=0.2)
    o = oc.Oracle(spec, 0.01)
    n = 0
    for orbs in dets_with(nb, nel, nalpha, 8, seed=nspat):
        det = encode_det(orbs, nb)
        occ = set(orbs)
        for i in orbs:
            for a in range(1, nb + 1):
                if a in occ or a % 2 != i % 2:
                    continue
                w = o.hb_weight(det, i, a)
                h = abs(sq_element(mol.h1, mol.eri, sorted((occ - {i}) | {a}), orbs))
                assert abs(abs(w) - h) <= 1e-12 * max(h, 1.0), (orbs, i, a, w, h)
                n += 1
    assert n > 50
    o.close()
