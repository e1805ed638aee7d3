"""Pins the CPU oracle (CPU only, no GPU): each oracle function is checked
against something that fixes it independently of its own code: published
known-answer vectors, the reference's own unit-test values, brute-force
second quantisation, exhaustive enumeration, or exact diagonalisation.
"""
import itertools

import numpy as np
import pytest

import ed_reference as ed
import oracle_ctypes as oc
from common import hubbard_spec, mol_spec, molecule
from neci_stable_amd import HubbardChain, Spec, decode_det, encode_det
from neci_stable_amd.systems import find_walker_hash, umat_ind, umat_len


# --------------------------------------------------------------- Philox


@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    """Random123 philox4x32_10 known-answer vectors (kat_vectors)."""
    assert oc.philox(ctr, key) == expect


# -------------------------------------------------------------- hashing


def test_find_walker_hash_against_integer_restatement():
    """hash.F90:23-37 restated with Python big integers and explicit int64 wrap."""
    rng = np.random.default_rng(1)
    for nbasis, nel in [(20, 10), (40, 10), (80, 26), (126, 30)]:
        rh = rng.permutation(np.arange(1, 1000 * nbasis + 1))[:nbasis].astype(np.int32)
        for L in [7, 1000, 700000, 2 ** 27]:
            for _ in range(200):
                orbs = sorted(rng.choice(np.arange(1, nbasis + 1), nel, replace=False).tolist())
                assert oc.find_walker_hash(orbs, rh, L) == find_walker_hash(orbs, rh, L)
    # a hand-checked value: h = 1099511628211*(1*3) + 2*5 = 3298534884643 ; mod 1000 = 643
    assert find_walker_hash([1, 2], np.array([3, 5], dtype=np.int32), 1000) == 644
    assert oc.find_walker_hash([1, 2], np.array([3, 5], dtype=np.int32), 1000) == 644


def test_find_walker_hash_negative_wrap():
    """Products that wrap int64 negative take |mod(h, L)| + 1 with a truncating mod."""
    rh = np.array([40000, 39999, 39998, 39997, 39996, 39995, 39994, 39993], dtype=np.int32)
    orbs = [1, 2, 3, 4, 5, 6, 7, 8]
    h = 0
    for i, o in enumerate(orbs, 1):
        h = (1099511628211 * h + int(rh[o - 1]) * i) % (1 << 64)
    hs = h - (1 << 64) if h >= 1 << 63 else h
    assert hs < 0
    L = 999983
    m = abs(hs) % L  # Fortran mod truncates toward zero: |mod(h, L)| = |h| mod L
    assert oc.find_walker_hash(orbs, rh, L) == m + 1


# ------------------------------------------------------------ integrals


def test_umat_ind_is_a_bijection_on_8fold_classes():
    n = 7
    seen = {}
    for i, j, k, l in itertools.product(range(1, n + 1), repeat=4):
        idx = umat_ind(i, j, k, l)
        assert idx == oc.lib().o_umat_ind(i, j, k, l)
        # <ij|kl> = <kj|il> = <il|kj> = <jk... : the 8 permutations of physical notation
        for p in [(k, j, i, l), (i, l, k, j), (k, l, i, j), (j, i, l, k), (l, i, j, k), (j, k, l, i),
                  (l, k, j, i)]:
            assert umat_ind(*p) == idx
        cls = tuple(sorted([tuple(sorted((i, k))), tuple(sorted((j, l)))]))
        seen.setdefault(idx, set()).add(cls)
    assert all(len(v) == 1 for v in seen.values())
    assert sorted(seen) == list(range(1, umat_len(n) + 1))


# --------------------------------------------------------- excitations


def _perm_parity(seq):
    seq = list(seq)
    par = 0
    for i in range(len(seq)):
        for j in range(i + 1, len(seq)):
            if seq[i] > seq[j]:
                par ^= 1
    return par


@pytest.mark.parametrize("nb,nel", [(8, 3), (10, 5), (12, 4)])
def test_make_single_and_double_exhaustive(nb, nel):
    """nJ must be the sorted result and tParity the sign of the permutation that
    sorts nI with sources replaced in place by the targets (the fermionic sign)."""
    for nI in itertools.combinations(range(1, nb + 1), nel):
        virt = [o for o in range(1, nb + 1) if o not in nI]
        for e in range(1, nel + 1):
            for t in virt:
                nJ, ex, par = oc.make_single(nI, e, t)
                rep = list(nI)
                rep[e - 1] = t
                assert list(nJ) == sorted(rep)
                assert par == _perm_parity(rep)
        for e1, e2 in itertools.combinations(range(1, nel + 1), 2):
            for t1, t2 in itertools.combinations(virt, 2):
                nJ, ex, par = oc.make_double(nI, e1, e2, t1, t2)
                rep = list(nI)
                rep[e1 - 1] = t1
                rep[e2 - 1] = t2
                assert list(nJ) == sorted(rep)
                assert par == _perm_parity(rep)


# ------------------------------------------------------ matrix elements


def test_slater_condon_against_second_quantisation():
    """Every H_IJ of a 5-orbital (10 spin orbital) molecule from the oracle's
    Slater-Condon rules with packed UMAT equals the brute-force second-quantised
    matrix built from the dense (pq|rs) tensor."""
    mol = molecule(5, 4, rank=6, seed=11)
    spec = Spec(mol, init_walkers=100)
    orc = oc.Oracle(spec, 0.01)
    dets = ed.all_dets(10, 4, nalpha=2)
    H = ed.molecular_hamiltonian_dense(mol.h1, mol.eri, mol.ecore, dets)
    orbs = [decode_det(np.array([d], dtype=np.int64)) for d in dets]
    for a in range(len(dets)):
        for b in range(len(dets)):
            v = orc.helement(orbs[a], orbs[b])
            assert abs(v - H[a, b]) <= 1e-11 * max(1.0, abs(H[a, b])), (orbs[a], orbs[b], v, H[a, b])
    # hermiticity and the diagonal through its own routine
    for a in range(len(dets)):
        assert abs(orc.diag(orbs[a]) - H[a, a]) <= 1e-11 * max(1.0, abs(H[a, a]))


def _hub_oracle(nsites, nel, u, bhub, periodic):
    sysm = HubbardChain(nsites, nel, u, bhub=bhub, periodic=periodic)
    spec = Spec(sysm, init_walkers=100, ref_orbs=list(range(1, nel + 1)))
    return oc.Oracle(spec, 0.01)


def test_hubbard_helements_reference_unit_test_values():
    """get_helement_test (unit_tests/real_space_hubbard/test_real_space_hubbard.F90:2465-2595)."""
    o = _hub_oracle(2, 2, 0.0, 1.0, periodic=False)
    assert o.helement([1, 2], [1, 2]) == 0.0
    o = _hub_oracle(2, 2, 1.0, 1.0, periodic=False)
    assert o.helement([1, 2], [1, 2]) == 1.0
    assert o.helement([1, 2], [1, 4]) == 1.0
    assert o.helement([1, 2], [2, 3]) == -1.0
    assert o.helement([1, 2], [2, 4]) == 0.0
    o = _hub_oracle(3, 2, 0.0, 1.0, periodic=True)
    assert o.helement([1, 3], [1, 3]) == 0.0
    assert o.helement([1, 3], [1, 5]) == 1.0
    assert o.helement([1, 3], [3, 5]) == -1.0
    o = _hub_oracle(3, 4, 1.0, 1.0, periodic=True)
    assert o.helement([1, 2, 3, 4], [1, 2, 3, 4]) == 2.0
    assert o.helement([1, 2, 3, 4], [1, 2, 3, 6]) == 1.0
    assert o.helement([1, 2, 3, 4], [1, 3, 4, 6]) == 1.0
    assert o.helement([1, 2, 3, 4], [2, 3, 4, 5]) == -1.0


def test_hubbard_helements_against_second_quantisation():
    sysm = HubbardChain(6, 6, 3.0, bhub=-1.0)
    spec = Spec(sysm, init_walkers=100)
    o = oc.Oracle(spec, 0.01)
    dets = ed.hubbard_dets(6, 3, 3)
    H = ed.hubbard_hamiltonian(6, sysm.nbr, 3.0, -1.0, dets).toarray()
    orbs = [decode_det(np.array([d], dtype=np.int64)) for d in dets]
    for a in range(0, len(dets), 3):
        for b in range(len(dets)):
            assert o.helement(orbs[a], orbs[b]) == H[b, a]


# ---------------------------------------------------------- alias tables


@pytest.mark.parametrize("seed", range(5))
def test_alias_table_reproduces_weights(seed):
    """test_aliasTables (unit_tests/sampler/test_aliasTables.F90:135,170): the
    distribution encoded by (bias, alias) equals w / sum(w); probs == w / sum(w)."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 300))
    w = rng.random(n) * (rng.random(n) < 0.7)
    if w.sum() == 0:
        w[0] = 1.0
    bias, alias, prob = oc.alias_build(w)
    implied = bias / n
    for j in range(n):
        implied[alias[j]] += (1.0 - bias[j]) / n
    np.testing.assert_allclose(implied, w / w.sum(), atol=1e-13)
    ssum = 0.0
    for x in w:  # Fortran sum(): sequential
        ssum += float(x)
    np.testing.assert_array_equal(prob, w / ssum)


# ---------------------------------------------------- generator statistics


@pytest.mark.parametrize("particles", [0, 1])
def test_generator_pgen_consistency_and_frequencies(particles):
    """run_excit_gen_tester (src/unit_test_helpers.F90:438-638): the returned pgen
    equals calc_pgen; over N draws every connected det appears and its frequency
    matches pgen (15% relative criterion of pchb_excitgen_test_helper.F90:105-111)."""
    mol = molecule(5, 4, rank=6, seed=11)
    spec = Spec(mol, init_walkers=100, particles=particles, psingles=0.3)
    o = oc.Oracle(spec, 0.01)
    ref = o.ref
    rng = np.random.default_rng(4)
    n = 300000
    counts = {}
    pgens = {}
    draws = rng.random((n, 6))
    for i in range(n):
        r = o.gen_excit(ref, draws[i])
        if r is None:
            continue
        dj, ic, ex, tp, pg = r
        key = tuple(dj)
        counts[key] = counts.get(key, 0) + 1
        if key in pgens:
            assert pgens[key] == pg
        else:
            pgens[key] = pg
            assert abs(o.calc_pgen(ref, ic, ex) - pg) <= 1e-14
    # connected dets with nonzero H must all be generated
    occ = spec.ref_orbs
    for d in ed.all_dets(10, 4, nalpha=2):
        orbs = decode_det(np.array([d], dtype=np.int64))
        ic = len(set(occ) - set(orbs))
        if ic in (1, 2) and abs(o.helement(occ, orbs)) > 1e-10:
            assert tuple(encode_det(orbs, 10)) in counts, orbs
    for key, c in counts.items():
        if pgens[key] * n > 400:
            assert abs(c / n - pgens[key]) / pgens[key] < 0.15
    total = sum(pgens.values())
    assert total <= 1.0 + 1e-12


# ----------------------------------------------------- end-to-end energy


def test_oracle_fciqmc_energy_vs_exact_diagonalisation():
    """Plain FCIQMC on a 6-site Hubbard ring: shift within 3 sigma of ED."""
    from neci_stable_amd.fcimc import FciMCPar, OracleCycleAdapter, reblock
    sysm = HubbardChain(6, 6, 4.0, bhub=-1.0)
    spec = Spec(sysm, init_walkers=2000, seed=99, trunc_initiator=False)
    o = oc.Oracle(spec, 0.02)
    o.add_walkers(o.ref[None, :], np.array([[10.0]]), [1])
    drv = FciMCPar(OracleCycleAdapter(o, 1), tau=0.02, target_walkers=2000)
    drv.run(6000)
    sh = np.array([h[1][0] for h in drv.shift_hist])
    mean, err = reblock(sh[len(sh) // 4:])
    e_ed = ed.ground_state(ed.hubbard_hamiltonian(6, sysm.nbr, 4.0, -1.0, ed.hubbard_dets(6, 3, 3)))
    assert abs(mean + o.hii - e_ed) < 3 * err + 1e-3, (mean, err, e_ed)


def test_target_shiftdamp_holds_population_and_energy():
    """TARGET-SHIFTDAMP (fcimc_iter_utilities.F90:1147-1154): the shift varies
    from the start and its extra term ln(N/N_target)*SftDamp2 pulls the
    population to TOTALWALKERS; the shift still averages to the ED energy."""
    from neci_stable_amd.fcimc import FciMCPar, OracleCycleAdapter, reblock
    sysm = HubbardChain(6, 6, 4.0, bhub=-1.0)
    spec = Spec(sysm, init_walkers=3000, seed=7, trunc_initiator=False)
    o = oc.Oracle(spec, 0.02)
    o.add_walkers(o.ref[None, :], np.array([[10.0]]), [1])
    drv = FciMCPar(OracleCycleAdapter(o, 1), tau=0.02, target_walkers=3000, shift_damp=0.3,
                   target_shiftdamp=True)
    assert not drv.single_part.any()
    drv.run(6000)
    pops = np.array([r["tot_parts"][0] for r in drv.rows])
    assert abs(pops[-300:].mean() / 3000.0 - 1.0) < 0.1, pops[-300:].mean()
    sh = np.array([r["shift"][0] for r in drv.rows])
    mean, err = reblock(sh[len(sh) // 4:])
    e_ed = ed.ground_state(ed.hubbard_hamiltonian(6, sysm.nbr, 4.0, -1.0, ed.hubbard_dets(6, 3, 3)))
    assert abs(mean + o.hii - e_ed) < 3 * err + 1e-3, (mean, err, e_ed)


# ------------------------------------------------ C ABI of the engine


def test_engine_library_exports_every_declared_symbol():
    import ctypes
    import re
    import os
    from neci_stable_amd import _lib
    hdr = open(os.path.join(os.path.dirname(_lib.__file__), "..", "include", "neci_gpu.h")).read()
    names = set(re.findall(r"^int\s+(ngpu_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 20
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert names == set(_lib.EXPORTED)


def test_engine_refuses_to_run_without_gpu_or_library(monkeypatch, tmp_path):
    from neci_stable_amd import _lib
    with pytest.raises(RuntimeError):
        _lib.load_library(str(tmp_path / "missing.so"))


# ------------------------------------------- merge + annihilation (CPU)
def _random_hubbard_dets(spec, n, rng):
    nb, na = spec.nbasis, spec.nel // 2
    out = np.zeros((n, spec.W), dtype=np.int64)
    for i in range(n):
        orbs = list(rng.choice(np.arange(2, nb + 1, 2), na, replace=False)) + \
            list(rng.choice(np.arange(1, nb + 1, 2), spec.nel - na, replace=False))
        out[i] = encode_det(sorted(int(o) for o in orbs), nb)
    return out


@pytest.mark.parametrize("nrep", [1, 2])
def test_merge_and_annihilation_against_groupby(nrep):
    """CompressSpawnedList + AnnihilateSpawnedParts on an empty walker list,
    checked against a numpy group-by of the spawned list: merged sums, the
    annihilation count (sum|s| - |sum s| per det and replica, the telescoping
    form of Annihilation.F90:590-602), initiator flags where the coherent rule
    does not depend on the order of the run, and the abort of non-initiator
    spawns onto unoccupied dets (Annihilation.F90:1199-1239)."""
    spec = hubbard_spec(init_walkers=4000, n_replicas=nrep)
    orc = oc.Oracle(spec, 0.01)
    rng = np.random.default_rng(5 + nrep)
    pool = np.unique(_random_hubbard_dets(spec, 200, rng), axis=0)
    n = 3000
    sdet = pool[rng.integers(0, len(pool), n)]
    sval = rng.integers(-3, 4, n).astype(float)
    sval[sval == 0] = 2.0
    rep = rng.integers(0, nrep, n).astype(np.uint32)
    init = rng.random(n) < 0.3
    sinfo = rep | (init.astype(np.uint32) << 8)
    st, md, ms, mf = orc.annihilate(3, sdet, sval, sinfo)
    assert st["rc"] == 0
    # group-by over (det, replica)
    keys, inv = np.unique(sdet, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    nk = len(keys)
    sums = np.zeros((nk, nrep))
    abss = np.zeros((nk, nrep))
    cnt = np.zeros((nk, nrep), dtype=int)
    pos = np.zeros((nk, nrep), dtype=int)
    neg = np.zeros((nk, nrep), dtype=int)
    anyinit = np.zeros((nk, nrep), dtype=bool)
    for i in range(n):
        k, r = inv[i], rep[i]
        sums[k, r] += sval[i]
        abss[k, r] += abs(sval[i])
        cnt[k, r] += 1
        pos[k, r] += sval[i] > 0
        neg[k, r] += sval[i] < 0
        anyinit[k, r] |= bool(init[i])
    keep = np.abs(sums).sum(axis=1) > 0
    # merged list == group-by of nonzero totals
    got = {tuple(d): (s, f) for d, s, f in zip(md, ms, mf)}
    assert len(got) == int(keep.sum())
    for k in np.nonzero(keep)[0]:
        s, f = got[tuple(keys[k])]
        np.testing.assert_array_equal(s, sums[k])
        for r in range(nrep):
            flag = (int(f) >> r) & 1
            if anyinit[k, r]:
                assert flag == 1
            elif cnt[k, r] == 1:
                assert flag == 0
            elif cnt[k, r] >= 2 and (pos[k, r] == 0 or neg[k, r] == 0):
                assert flag == 1   # same-sign run: every step sees two nonzero amplitudes
    np.testing.assert_allclose(st["nannihil"][:nrep], (abss - np.abs(sums)).sum(axis=0), rtol=0, atol=0)
    # no resident walkers: merged dets whose replica is not flagged initiator are aborted
    aborted = np.zeros(nrep)
    for k in np.nonzero(keep)[0]:
        s, f = got[tuple(keys[k])]
        for r in range(nrep):
            if not (int(f) >> r) & 1:
                aborted[r] += abs(s[r])
    np.testing.assert_array_equal(st["naborted"][:nrep], aborted)


def test_fp32_spawn_storage_rounds_each_spawn():
    """FP32 spawned amplitudes (real coefficients): one iteration from the
    same start with FP64 and with FP32 spawn storage makes the same spawning
    decisions (the rounding comes after them), so both runs hold the same
    determinants; each amplitude then differs by FP32 rounding only, some do
    differ, and NoBorn (the sum of |spawn|) differs by at most half an FP32
    ulp per spawn."""
    res = {}
    for fp32 in (False, True):
        spec = mol_spec(8, 6, init_walkers=4000, all_real_coeff=True, seed=19, spawn_fp32=fp32)
        o = oc.Oracle(spec, 0.02)
        o.add_walkers(o.ref[None, :], np.full((1, 1), 3000.0), [1])
        st = o.iterate(1, 0.3)
        assert st["rc"] == 0
        d, sg, _, _, _ = o.walkers()
        order = np.lexsort(d.T[::-1])
        res[fp32] = (st, d[order], sg[order, 0])
    (a64, d64, s64), (a32, d32, s32) = res[False], res[True]
    np.testing.assert_array_equal(d64, d32)
    assert a64["n_spawned"] == a32["n_spawned"] > 100
    assert np.all(np.abs(s32 - s64) <= 1e-6 * (np.abs(s64) + 1.0))
    assert np.count_nonzero(s32 != s64) > 10
    nb64, nb32 = a64["nborn"][0], a32["nborn"][0]
    assert nb32 != nb64 and abs(nb32 - nb64) <= 2.0 ** -24 * nb64


# ------------------------------------------------ semistochastic core space
@pytest.mark.parametrize("fp32", [False, True])
def test_core_hamiltonian_and_projection_against_dense(fp32):
    """The oracle's core Hamiltonian (CSR, diagonal H_ii - Hii) against the
    second-quantised matrix of the same determinants, and one deterministic
    projection y = tau*(S*x - (H - Hii)*x) (semi_stoch_procs.F90:105-241)
    against a dense numpy product."""
    sysm = HubbardChain(6, 6, 3.0, bhub=-1.0)
    spec = Spec(sysm, init_walkers=1000, n_replicas=2)
    o = oc.Oracle(spec, 0.01)
    dets = ed.hubbard_dets(6, 3, 3)
    H = ed.hubbard_hamiltonian(6, sysm.nbr, 3.0, -1.0, dets).toarray()
    rng = np.random.default_rng(4)
    pick = np.sort(rng.choice(len(dets), 120, replace=False))
    core = np.array([encode_det(decode_det(np.array([dets[k]], dtype=np.int64)), 12) for k in pick])
    nnz = o.set_core(core, fp32=fp32)
    rp, cols, vals = o.core_csr()
    assert nnz == len(cols)
    hc = H[np.ix_(pick, pick)] - o.hii * np.eye(len(pick))
    dense = np.zeros_like(hc)
    for i in range(len(pick)):
        assert np.all(np.diff(cols[rp[i]:rp[i + 1]]) > 0)    # ascending columns
        dense[i, cols[rp[i]:rp[i + 1]]] = vals[rp[i]:rp[i + 1]]
    expect = hc.astype(np.float32).astype(np.float64) if fp32 else hc
    np.testing.assert_array_equal(dense, expect)
    x = rng.normal(size=(len(pick), 2))
    shift = np.array([0.3, -0.2])
    y = o.core_project(x, shift, 0.01)
    ref = 0.01 * (x * shift[None, :] - dense @ x)
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-14)
    # the core dets are resident, flagged deterministic and initiator in every replica
    d, s, f, _, _ = o.walkers()
    have = {tuple(r) for r in d}
    assert all(tuple(c) in have for c in core.reshape(-1, 1))
    fl = {tuple(r): int(x) for r, x in zip(d, f)}
    assert all(fl[tuple(c)] & (1 << 16) and fl[tuple(c)] & 3 == 3 for c in core.reshape(-1, 1))


# ------------------------------------------------- uniform molecular generator
def _all_excitations(orbs, nbasis):
    """Every spin-conserving single and double excitation of a determinant:
    (ic, ex=[src.., tgt..] sorted, child orbitals)."""
    occ = set(orbs)
    virt = [o for o in range(1, nbasis + 1) if o not in occ]
    out = []
    for a in orbs:
        for t in virt:
            if (a & 1) == (t & 1):
                out.append((1, [a, t, 0, 0], sorted((occ - {a}) | {t})))
    for i, a in enumerate(orbs):
        for b in orbs[i + 1:]:
            for j, t in enumerate(virt):
                for u in virt[j + 1:]:
                    if sorted([a & 1, b & 1]) == sorted([t & 1, u & 1]):
                        out.append((2, [a, b, t, u], sorted((occ - {a, b}) | {t, u})))
    return out


@pytest.mark.parametrize("nspat,nel,which", [(5, 4, "ref"), (5, 4, "random"), (5, 8, "ref"), (6, 7, "random")])
def test_uniform_generator_pgen_and_frequencies(nspat, nel, which):
    """gen_rand_excit restated for C1 systems: the returned pgen equals
    calc_pgen_symrandexcit2 bit for bit, pgen summed over every spin-allowed
    excitation is 1 (or below 1 where the reference's forbidden-orbital rules
    exclude electron pairs), and sampled frequencies match pgen within the 15%
    criterion of run_excit_gen_tester (pchb_excitgen_test_helper.F90:105-111).
    The cases cover the rejection loops, the enumerating branch (<= 3 allowed
    orbitals) and forbidden orbitals (one unoccupied orbital per spin)."""
    mol = molecule(nspat, nel, rank=6, seed=11)
    spec = Spec(mol, init_walkers=100, excitgen="uniform", psingles=0.3)
    o = oc.Oracle(spec, 0.01)
    rng = np.random.default_rng(7)
    nb = 2 * nspat
    if which == "ref":
        orbs = list(spec.ref_orbs)
    else:
        na = sum(1 for x in spec.ref_orbs if x % 2 == 0)
        orbs = sorted(list(rng.choice(np.arange(2, nb + 1, 2), na, replace=False)) +
                      list(rng.choice(np.arange(1, nb + 1, 2), nel - na, replace=False)))
    det = encode_det([int(x) for x in orbs], nb)
    exc = _all_excitations([int(x) for x in orbs], nb)
    pg_enum = {}
    for ic, ex, child in exc:
        pg_enum[tuple(encode_det(child, nb))] = o.calc_pgen(det, ic, ex)
    total = sum(pg_enum.values())
    assert total <= 1.0 + 1e-12
    n = 200000
    draws = rng.random((n, 6))
    counts = {}
    for i in range(n):
        r, fail = o.gen_excit_ctx(det, draws[i], 3, 0, i)
        assert not fail
        if r is None:
            continue
        dj, ic, ex, tp, pg = r
        key = tuple(dj)
        assert key in pg_enum
        assert pg == pg_enum[key]
        counts[key] = counts.get(key, 0) + 1
    for key, pgv in pg_enum.items():
        if pgv * n > 400:
            assert abs(counts.get(key, 0) / n - pgv) / pgv < 0.15, (key, counts.get(key, 0) / n, pgv)
    generated = sum(counts.values()) / n
    assert abs(generated - total) < 0.01
