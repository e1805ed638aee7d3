"""GPU parity: every engine entry point against the CPU oracle on identical inputs.

Integer/index work (hash buckets, excitation choice, walker sets, integer
walker signs, integer-valued statistics) must agree bit for bit; FP64 sums
whose order differs between the two (the projected-energy numerator, which the
GPU accumulates with atomics) within 1e-12 relative.
"""
import numpy as np
import pytest

from common import (hubbard_spec, internal_to_neci_flags, mol_spec, neci_to_internal_flags,
                    sort_walkers)
from neci_stable_amd.systems import decode_det, encode_det

pytestmark = pytest.mark.gpu

EXACT_FIELDS = ["nborn", "ndied", "nannihil", "naborted", "nremoved", "tot_parts", "norm_sq",
                "hf_cyc", "noat_doubs", "spawn_from_sing", "inst_no_at_hf", "no_init_dets",
                "no_init_walk", "no_noninit_dets", "no_noninit_walk"]


def engine(spec):
    from neci_stable_amd.engine import Engine
    return Engine(spec)


def oracle(spec, tau):
    from oracle_ctypes import Oracle
    return Oracle(spec, tau)


def random_dets(spec, n, rng):
    out = np.zeros((n, spec.W), dtype=np.int64)
    nb = spec.nbasis
    na = sum(1 for o in spec.ref_orbs if o % 2 == 0)
    nbeta = spec.nel - na
    alpha = np.arange(2, nb + 1, 2)
    beta = np.arange(1, nb + 1, 2)
    for i in range(n):
        orbs = list(rng.choice(alpha, na, replace=False)) + list(rng.choice(beta, nbeta, replace=False))
        out[i] = encode_det(sorted(int(o) for o in orbs), nb)
    return out


def hopping_neighbourhood(spec, ref, depth):
    """Dets reachable from ref by up to `depth` nearest-neighbour hops (Hubbard)."""
    nbr = spec.system.nbr
    seen = {tuple(ref)}
    front = [np.asarray(ref)]
    for _ in range(depth):
        nxt = []
        for d in front:
            orbs = decode_det(d.reshape(1, -1))
            occ = set(orbs)
            for o in orbs:
                site = (o + 1) // 2
                for t in nbr[site - 1]:
                    if t == 0:
                        continue
                    to = 2 * t - 1 if o % 2 == 1 else 2 * t
                    if to in occ:
                        continue
                    nd = encode_det(sorted((occ - {o}) | {to}), spec.nbasis)
                    if tuple(nd) not in seen:
                        seen.add(tuple(nd))
                        nxt.append(nd)
        front = nxt
    return np.array(sorted(seen), dtype=np.int64)


def assert_stats(a, b, S, exact=True):
    for f in EXACT_FIELDS:
        np.testing.assert_array_equal(a[f][:S], b[f][:S], err_msg=f)
    np.testing.assert_allclose(a["enum_cyc"][:S], b["enum_cyc"][:S], rtol=1e-12, atol=1e-12)
    for f in ["n_valid_excits", "n_invalid_excits", "tot_walkers", "n_attempts", "n_spawned",
              "n_unique", "n_new"]:
        assert a[f] == b[f], (f, a[f], b[f])
    np.testing.assert_array_equal(a["bloom_count"], b["bloom_count"])
    np.testing.assert_array_equal(a["bloom_size"], b["bloom_size"])
    assert a["max_cyc_spawn"] == b["max_cyc_spawn"]


def assert_same_walkers(eng, orc):
    d1, s1, f1, dh1, oh1 = eng.download()
    d2, s2, f2, dh2, oh2 = orc.walkers()
    f1 = neci_to_internal_flags(f1, eng.S)
    d1, s1, f1, dh1, oh1 = sort_walkers(d1, s1, f1, dh1, oh1)
    d2, s2, f2, dh2, oh2 = sort_walkers(d2, s2, f2, dh2, oh2)
    assert d1.shape == d2.shape, (d1.shape, d2.shape)
    np.testing.assert_array_equal(d1, d2)
    np.testing.assert_array_equal(s1, s2)
    np.testing.assert_array_equal(f1 & 0xFFFF, f2 & 0xFFFF)
    np.testing.assert_array_equal(dh1, dh2)
    np.testing.assert_array_equal(oh1, oh2)
    return len(d1)


SPECS = {
    "hubbard10": lambda **k: hubbard_spec(10, 10, 4.0, **k),
    "mol8o6e_unif_fast": lambda **k: mol_spec(8, 6, **k),
    "mol8o6e_unif_unif": lambda **k: mol_spec(8, 6, particles=0, **k),
    "mol33o4e_w2": lambda **k: mol_spec(33, 4, rank=4, mseed=3, **k),
    # C(nbasis, nel) > 2^32: the spawn sort runs on the determinant words
    # themselves instead of 32-bit colex ranks
    "hubbard18": lambda **k: hubbard_spec(18, 18, 4.0, **k),
    "mol33o8e_w2": lambda **k: mol_spec(33, 8, rank=4, mseed=3, **k),
    # C(60, 14) ~ 5e13: the spawn sort runs on 46-bit colex ranks (6 byte passes, not 8)
    "mol30o14e_rank64": lambda **k: mol_spec(30, 14, rank=8, mseed=5, **k),
    # uniform random excitations (gen_rand_excit) with rejection loops
    "mol8o6e_uniform": lambda **k: mol_spec(8, 6, excitgen="uniform", psingles=0.2, **k),
    "mol33o4e_w2_uniform": lambda **k: mol_spec(33, 4, rank=4, mseed=3, excitgen="uniform", **k),
    # plain FCIQMC: no initiator approximation (TRUNCINITIATOR off)
    "hubbard10_plain": lambda **k: hubbard_spec(10, 10, 4.0, trunc_initiator=False, **k),
    "mol8o6e_plain": lambda **k: mol_spec(8, 6, particles=0, trunc_initiator=False, **k),
    # weighted PCHB samplers: UNIF-FULL / FULL-FULL particles, fully weighted
    # holes (the reference's choice when nBasis <= 4 nEl)
    "mol8o6e_unif_full": lambda **k: mol_spec(8, 6, particles=2, holes=1, **k),
    "mol8o6e_full_full": lambda **k: mol_spec(8, 6, particles=3, holes=1, **k),
    "mol8o6e_full_fast": lambda **k: mol_spec(8, 6, particles=3, holes=0, **k),
    "mol33o4e_w2_unif_full": lambda **k: mol_spec(33, 4, rank=4, mseed=3, particles=2, holes=1, **k),
}


@pytest.mark.parametrize("name", list(SPECS))
def test_hash_and_helements(name):
    spec = SPECS[name]()
    eng = engine(spec)
    orc = oracle(spec, 0.01)
    rng = np.random.default_rng(5)
    dets = random_dets(spec, 400, rng)
    buckets = eng.find_walker_hash(dets)
    from oracle_ctypes import find_walker_hash
    for i in range(len(dets)):
        assert buckets[i] == find_walker_hash(decode_det(dets[i]), spec.random_hash, spec.n_hash_buckets)
    # H_ij between pairs at excitation levels 0..3
    pairs_i = []
    pairs_j = []
    for i in range(len(dets)):
        d = decode_det(dets[i])
        for k in range(4):
            nd = list(d)
            occ = set(nd)
            for _ in range(min(k, 2)):
                src = nd[rng.integers(len(nd))]
                virt = [o for o in range(1, spec.nbasis + 1) if o not in occ and (o & 1) == (src & 1)]
                if not virt:
                    continue
                tgt = int(rng.choice(virt))
                nd.remove(src)
                nd.append(tgt)
                occ = set(nd)
            pairs_i.append(dets[i])
            pairs_j.append(encode_det(sorted(nd), spec.nbasis))
    pi = np.array(pairs_i)
    pj = np.array(pairs_j)
    hel = eng.helement(pi, pj)
    for q in range(len(pi)):
        ref = orc.helement(decode_det(pi[q]), decode_det(pj[q]))
        assert hel[q] == ref, (q, hel[q], ref)
    diag, off = eng.diag(dets)
    for q in range(len(dets)):
        assert diag[q] == orc.diag(decode_det(dets[q]))
        assert off[q] == orc.offdiag(dets[q])
    assert eng.hii == orc.hii


@pytest.mark.parametrize("name", list(SPECS))
def test_generate_excitation_and_pgen(name):
    spec = SPECS[name]()
    eng = engine(spec)
    orc = oracle(spec, 0.01)
    rng = np.random.default_rng(11)
    dets = random_dets(spec, 2000, rng)
    draws = rng.random((2000, 6))
    dj, ic, ex, par, pg = eng.generate_excitations(dets, draws)
    nvalid = 0
    for i in range(len(dets)):
        r = orc.gen_excit(dets[i], draws[i])
        if r is None:
            assert ic[i] == 0, i
            continue
        nvalid += 1
        djo, ico, exo, tpo, pgo = r
        assert ic[i] == ico
        np.testing.assert_array_equal(dj[i], djo)
        assert par[i] == tpo
        assert pg[i] == pgo
        exs = ex[i][:2 * ico]
        np.testing.assert_array_equal(exs, exo[:2 * ico])
    assert nvalid > 100
    ok = ic > 0
    pgen_re = eng.calc_pgen(dets[ok], ic[ok], ex[ok])
    np.testing.assert_allclose(pgen_re, pg[ok], rtol=1e-12)


def test_attempt_die_and_create():
    spec = hubbard_spec()
    eng = engine(spec)
    rng = np.random.default_rng(2)
    n = 5000
    kii = rng.uniform(-1, 30, n)
    sg = rng.integers(-50, 50, n).astype(float)
    u = rng.random(n)
    nd = eng.attempt_die(kii, sg, u, 0.01, 1.5)
    for i in range(n):
        rat = 0.01 * (kii[i] - 1.5) * abs(sg[i])
        ip = float(int(rat))
        frac = rat - ip
        exp = ip + (np.sign(frac) if abs(frac) > u[i] else 0.0)
        assert nd[i] == exp
    hel = rng.normal(0, 1, n)
    pg = rng.uniform(1e-4, 1, n)
    child = eng.attempt_create(hel, pg, sg, u, 0.01)
    for i in range(n):
        x = -0.01 * hel[i] * (1.0 if sg[i] >= 0 else -1.0) / pg[i]
        ip = float(int(x))
        frac = x - ip
        exp = ip + (np.sign(x) if (abs(frac) >= 1e-12 and abs(frac) > u[i]) else 0.0)
        assert child[i] == exp


@pytest.mark.parametrize("name", ["hubbard10", "mol8o6e_unif_fast", "mol33o4e_w2", "hubbard18", "mol33o8e_w2",
                                  "mol30o14e_rank64"])
def test_annihilate_explicit_list(name):
    spec = SPECS[name](init_walkers=3000)
    eng = engine(spec)
    orc = oracle(spec, 0.01)
    rng = np.random.default_rng(21)
    # resident walkers
    dets = np.unique(random_dets(spec, 300, rng), axis=0)
    sg = rng.integers(-6, 7, (len(dets), 1)).astype(float)
    sg[sg == 0] = 1.0
    fl = (np.abs(sg[:, 0]) > 3).astype(np.uint32)
    eng.upload(eng.rows(dets, sg, internal_to_neci_flags(fl, 1)))
    orc.add_walkers(dets, sg, fl)
    # spawned list: repeats, cancellations, spawns on resident and new dets
    pool = np.concatenate([dets[:150], random_dets(spec, 150, rng)])
    idx = rng.integers(0, len(pool), 2500)
    sdet = pool[idx]
    sval = rng.integers(-3, 4, 2500).astype(float)
    sval[sval == 0] = 1.0
    sinfo = (rng.random(2500) < 0.5).astype(np.uint32) << 8
    a, md, ms, mf = eng.annihilate(7, sdet, sval, sinfo)
    b, md2, ms2, mf2 = orc.annihilate(7, sdet, sval, sinfo)
    assert b["rc"] == 0
    m1 = sort_walkers(md, ms, mf)
    m2 = sort_walkers(md2, ms2, mf2)
    for x, y in zip(m1, m2):
        np.testing.assert_array_equal(x, y)
    assert_stats(a, b, 1)
    assert_same_walkers(eng, orc)


@pytest.mark.parametrize("name", ["hubbard10", "mol33o4e_w2", "mol33o8e_w2", "mol30o14e_rank64"])
def test_annihilate_explicit_list_fp32(name):
    """An explicit list with real amplitudes under FP32 spawn storage: the
    engine rounds each input amplitude to FP32, so the oracle, given the same
    values rounded by numpy, merges and annihilates to identical results (the
    list's order fixes the order of each merge sum)."""
    spec = SPECS[name](init_walkers=3000, all_real_coeff=True, spawn_fp32=True)
    eng = engine(spec)
    orc = oracle(spec, 0.01)
    rng = np.random.default_rng(23)
    dets = np.unique(random_dets(spec, 300, rng), axis=0)
    sg = np.round(rng.normal(0, 4, (len(dets), 1)), 3)
    sg[sg == 0] = 1.0
    fl = (np.abs(sg[:, 0]) > 3).astype(np.uint32)
    eng.upload(eng.rows(dets, sg, internal_to_neci_flags(fl, 1)))
    orc.add_walkers(dets, sg, fl)
    pool = np.concatenate([dets[:150], random_dets(spec, 150, rng)])
    sdet = pool[rng.integers(0, len(pool), 2500)]
    sval = rng.normal(0, 1.5, 2500) + 0.1
    sinfo = (rng.random(2500) < 0.5).astype(np.uint32) << 8
    a, md, ms, mf = eng.annihilate(7, sdet, sval, sinfo)
    b, md2, ms2, mf2 = orc.annihilate(7, sdet, sval.astype(np.float32).astype(np.float64), sinfo)
    assert b["rc"] == 0
    m1 = sort_walkers(md, ms, mf)
    m2 = sort_walkers(md2, ms2, mf2)
    for x, y in zip(m1, m2):
        np.testing.assert_array_equal(x, y)
    assert_same_walkers(eng, orc)
    # and the FP64 list differs: the rounding took place
    b64, md3, ms3, mf3 = oracle(SPECS[name](init_walkers=3000, all_real_coeff=True), 0.01).annihilate(
        7, sdet, sval, sinfo)
    assert not np.array_equal(sort_walkers(md3, ms3, mf3)[1], m2[1])


@pytest.mark.parametrize("name,nrep,real", [("hubbard10", 1, False), ("hubbard10", 2, False),
                                            ("mol8o6e_unif_fast", 1, False),
                                            ("mol8o6e_unif_unif", 1, False),
                                            ("mol33o4e_w2", 1, False), ("hubbard18", 1, False),
                                            ("mol33o8e_w2", 1, False), ("hubbard10", 1, True),
                                            ("mol8o6e_unif_fast", 1, True), ("mol8o6e_uniform", 1, False),
                                            ("mol8o6e_uniform", 2, False), ("mol33o4e_w2_uniform", 1, False),
                                            ("hubbard10_plain", 1, False), ("hubbard10_plain", 2, False),
                                            ("mol8o6e_plain", 1, False), ("mol8o6e_unif_full", 1, False),
                                            ("mol8o6e_full_full", 1, False), ("mol8o6e_full_fast", 2, False),
                                            ("mol33o4e_w2_unif_full", 1, False), ("mol8o6e_unif_full", 1, True),
                                            ("mol30o14e_rank64", 1, False), ("mol30o14e_rank64", 2, True)])
def test_iterations_match_oracle(name, nrep, real):
    """Full iterations (spawn, merge, death, annihilation, stats) from the same
    walker set: identical walker sets after every iteration."""
    run_iterations(name, nrep, real)


@pytest.mark.parametrize("name,nrep", [("hubbard10", 1), ("hubbard10", 2), ("mol8o6e_unif_fast", 1),
                                       ("mol33o4e_w2", 1), ("mol8o6e_uniform", 1), ("hubbard18", 1),
                                       ("mol30o14e_rank64", 1)])
def test_iterations_fp32_spawns_match_oracle(name, nrep):
    """Spawned amplitudes stored in FP32 with real coefficients: the engine and
    the oracle round each spawn to FP32 at the same point, so walker signs and
    statistics agree to 1e-10 (an FP64 spawn on one side would differ by
    ~1e-8 relative)."""
    run_iterations(name, nrep, True, spawn_fp32=True)


def run_iterations(name, nrep, real, spawn_fp32=False):
    # the 8-electron 2-word molecule grows ~1.4x per iteration at tau 0.02
    tau = 0.01 if name.startswith("hub") else (0.005 if name == "mol33o8e_w2" else
                                               0.002 if name.startswith("mol30o14e") else 0.02)
    spec = SPECS[name](init_walkers=4000, n_replicas=nrep, all_real_coeff=real, seed=77, spawn_fp32=spawn_fp32)
    eng = engine(spec)
    orc = oracle(spec, tau)
    n0 = 30.0 if name.startswith("hub") else 400.0
    eng.start_single_part(n0)
    orc.add_walkers(orc.ref[None, :], np.full((1, nrep), n0), [(1 << nrep) - 1])
    shift = np.linspace(0.2, 0.4, nrep)
    for it in range(1, 41):
        a = eng.iterate(it, tau, shift)
        b = orc.iterate(it, shift)
        assert b["rc"] == 0
        if real:
            for f in EXACT_FIELDS + ["enum_cyc"]:
                np.testing.assert_allclose(a[f][:nrep], b[f][:nrep], rtol=1e-10, atol=1e-9, err_msg=f)
        else:
            assert_stats(a, b, nrep)
    if real:
        d1, s1, _, _, _ = eng.download()
        d2, s2, _, _, _ = orc.walkers()
        d1, s1 = sort_walkers(d1, s1)
        d2, s2 = sort_walkers(d2, s2)
        np.testing.assert_array_equal(d1, d2)
        np.testing.assert_allclose(s1, s2, rtol=1e-10, atol=1e-10)
    else:
        n = assert_same_walkers(eng, orc)
        assert n > 50


def test_config1_molecule_matches_oracle():
    """The 10e/20o synthetic molecule (seed 7, Cholesky rank 60): PCHB doubles
    with UNIF-UNIF particle selection and heat-bath singles (psingles 0.1),
    initiator threshold 3, tau 0.005.  The engine grows the population to
    about 1e5 walkers (TARGET-SHIFTDAMP); from that walker set the engine
    and the oracle run 20 more iterations with identical statistics and
    identical walker sets."""
    from neci_stable_amd import Spec
    from neci_stable_amd.fcimc import FciMCPar
    from common import molecule
    tau = 0.005
    spec = Spec(molecule(20, 10, 60, 7), init_walkers=1e5, seed=1234, particles=0, psingles=0.1,
                pparallel=0.5, memory_fac_spawn=8.0)
    eng = engine(spec)
    eng.start_single_part(100.0)
    drv = FciMCPar(eng, tau=tau, target_walkers=1e5, steps_shift=10, shift_damp=0.5, target_shiftdamp=True)
    while drv.rows == [] or drv.rows[-1]["tot_parts"][0] < 0.9e5:
        drv.run(10)
        assert drv.iter < 3000
    dets, sg, fl, _, _ = eng.download()
    print(f"config-1 molecule: {sg[:, 0].__abs__().sum():.0f} walkers on {len(dets)} determinants "
          f"after {drv.iter} iterations", flush=True)
    orc = oracle(spec, tau)
    orc.add_walkers_bulk(dets, sg, neci_to_internal_flags(fl, 1))
    shift = drv.shift.copy()
    for it in range(drv.iter + 1, drv.iter + 21):
        a = eng.iterate(it, tau, shift)
        b = orc.iterate(it, shift)
        assert b["rc"] == 0
        assert_stats(a, b, 1)
    assert assert_same_walkers(eng, orc) > 5e4
    eng.close()


def test_hash_rebuild_keeps_parity():
    """A walker hash small enough that tombstones trigger in-place rebuilds:
    the walker set must still follow the oracle exactly, rebuilds do happen,
    and after every iteration the live-entry count equals the number of
    determinants in the table, with the tombstones within the entries.
    110 buckets = 440 entries: growth at shift 0.6 fills ~62 % of them (new
    determinants mostly reuse the tombstones that removals leave); a shift of
    -30 then makes deaths outrun births, so tombstones pass 10 % of the
    entries while live plus tombstone entries stay above 60 %."""
    tau = 0.01
    spec = SPECS["hubbard10"](init_walkers=4000, seed=5, n_hash_buckets=110)
    eng = engine(spec)
    orc = oracle(spec, tau)
    eng.start_single_part(40.0)
    orc.add_walkers(orc.ref[None, :], np.full((1, 1), 40.0), [1])
    for it in range(1, 41):
        shift = 0.6 if it <= 31 else -30.0
        a = eng.iterate(it, tau, shift)
        b = orc.iterate(it, shift)
        assert b["rc"] == 0
        assert_stats(a, b, 1)
        hs = eng.hash_stats()
        assert hs["live"] == a["tot_walkers"], (it, hs, a["tot_walkers"])
        assert 0 <= hs["tombstones"] and hs["live"] + hs["tombstones"] <= hs["entries"], (it, hs)
    hs = eng.hash_stats()
    print(f"hash after 40 iterations: {hs}", flush=True)
    assert hs["rebuilds"] >= 1, hs
    assert assert_same_walkers(eng, orc) > 20


def test_hash_full_is_an_error():
    """More live determinants than 80% of the hash entries: NGPU_ERR_HASH_FULL."""
    from neci_stable_amd._lib import NgpuError
    spec = SPECS["hubbard10"](init_walkers=4000, n_hash_buckets=8)
    eng = engine(spec)
    rng = np.random.default_rng(3)
    dets = np.unique(random_dets(spec, 60, rng), axis=0)
    with pytest.raises(NgpuError) as e:
        eng.upload(eng.rows(dets, np.ones((len(dets), 1))))
    assert e.value.code == 6


@pytest.mark.parametrize("name,nrep,fp32,switch", [("hubbard10", 1, False, False), ("hubbard10", 2, False, False),
                                                    ("mol8o6e_unif_fast", 1, False, False),
                                                    ("mol8o6e_unif_fast", 1, True, False),
                                                    ("mol8o6e_unif_fast", 1, True, True),
                                                    ("mol33o4e_w2", 1, False, False),
                                                    ("mol33o4e_w2", 1, False, True),
                                                    ("mol8o6e_uniform", 1, False, False)])
def test_semistochastic_iterations_match_oracle(name, nrep, fp32, switch):
    """Iterations with a semistochastic core space (deterministic projection of
    the core amplitudes, core-to-core spawns cancelled): identical walker sets
    and signs after every iteration.  The core projection sums in the same
    fixed order on both sides, so amplitudes agree bit for bit; statistics that
    sum real amplitudes in a different order agree to 1e-10.  switch: the core
    is first set in the other precision, then set again with the same
    determinants (only the stored values change); a different core is refused."""
    tau = 0.01 if name.startswith("hub") else 0.02
    spec = SPECS[name](init_walkers=4000, n_replicas=nrep, seed=31)
    eng = engine(spec)
    orc = oracle(spec, tau)
    n0 = 30.0 if name.startswith("hub") else 400.0
    eng.start_single_part(n0)
    orc.add_walkers(orc.ref[None, :], np.full((1, nrep), n0), [(1 << nrep) - 1])
    rng = np.random.default_rng(8)
    if name.startswith("hub"):
        core = hopping_neighbourhood(spec, orc.ref, 2)
    else:
        core = np.unique(np.concatenate([orc.ref[None, :], random_dets(spec, 150, rng)]), axis=0)
    if switch:
        from neci_stable_amd._lib import NgpuError
        eng.set_core(core, fp32=not fp32)
        with pytest.raises(NgpuError) as e:
            eng.set_core(core[1:], fp32=fp32)
        assert e.value.code == 7
    nnz = eng.set_core(core, fp32=fp32)
    assert nnz == orc.set_core(core, fp32=fp32) and nnz > 2 * len(core)
    shift = np.linspace(0.2, 0.4, nrep)
    for it in range(1, 31):
        a = eng.iterate(it, tau, shift)
        b = orc.iterate(it, shift)
        assert b["rc"] == 0
        for f in EXACT_FIELDS + ["enum_cyc"]:
            np.testing.assert_allclose(a[f][:nrep], b[f][:nrep], rtol=1e-10, atol=1e-9, err_msg=f)
        for f in ["n_valid_excits", "n_invalid_excits", "tot_walkers", "n_attempts", "n_spawned",
                  "n_unique", "n_new"]:
            assert a[f] == b[f], (it, f, a[f], b[f])
    n = assert_same_walkers(eng, orc)
    assert n > len(core)
