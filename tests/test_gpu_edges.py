"""Edge cases of the spawned list and of the walker list, engine against the
oracle on identical inputs (bit for bit, as the other parity tests):
* explicit spawned lists (ngpu_annihilate) that are empty, hold one spawn
  onto a resident determinant or onto a new one, or cancel exactly (every
  determinant receives +v and -v, in shuffled order);
* a population driven to extinction by a shift far below every diagonal
  element (death factor tau*(H_ii - S) between 0.8 and 1.0, below the
  abort limit of 2), after which the engine iterates over an empty walker
  list;
* an engine that starts from an empty upload.
"""
import numpy as np
import pytest

from common import internal_to_neci_flags, sort_walkers
from test_gpu_parity import SPECS, assert_same_walkers, assert_stats, engine, oracle, random_dets

pytestmark = pytest.mark.gpu


def _resident(spec, eng, orc, rng, n=200):
    dets = np.unique(random_dets(spec, n, rng), axis=0)
    sg = rng.integers(-6, 7, (len(dets), 1)).astype(float)
    sg[sg == 0] = 1.0
    fl = (np.abs(sg[:, 0]) > 3).astype(np.uint32)
    eng.upload(eng.rows(dets, sg, internal_to_neci_flags(fl, 1)))
    orc.add_walkers(dets, sg, fl)
    return dets


def _not_resident(spec, dets, rng, n):
    cand = np.unique(random_dets(spec, 4 * n, rng), axis=0)
    res = set(map(tuple, dets.tolist()))
    keep = [c for c in cand.tolist() if tuple(c) not in res]
    return np.array(keep[:n], dtype=np.int64).reshape(-1, spec.W)


@pytest.mark.parametrize("name", ["hubbard10", "mol33o4e_w2"])
@pytest.mark.parametrize("case", ["empty", "one_resident", "one_new", "cancelling"])
def test_annihilate_edge_lists(name, case):
    spec = SPECS[name](init_walkers=3000)
    eng = engine(spec)
    orc = oracle(spec, 0.01)
    rng = np.random.default_rng(31)
    dets = _resident(spec, eng, orc, rng)
    if case == "empty":
        sdet = np.zeros((0, spec.W), np.int64)
        sval = np.zeros(0)
        sinfo = np.zeros(0, np.uint32)
    elif case == "one_resident":
        sdet, sval, sinfo = dets[:1], np.array([-2.0]), np.zeros(1, np.uint32)
    elif case == "one_new":  # from an initiator parent (info bit 8), so it is kept
        sdet, sval, sinfo = _not_resident(spec, dets, rng, 1), np.array([3.0]), np.array([1 << 8], np.uint32)
    else:
        pool = np.concatenate([dets[:40], _not_resident(spec, dets, rng, 40)])
        v = rng.integers(1, 4, len(pool)).astype(float)
        sdet = np.concatenate([pool, pool])
        sval = np.concatenate([v, -v])
        perm = rng.permutation(len(sdet))
        sdet, sval = sdet[perm], sval[perm]
        sinfo = (rng.random(len(sdet)) < 0.5).astype(np.uint32) << 8
    a, md, ms, mf = eng.annihilate(7, sdet, sval, sinfo)
    b, md2, ms2, mf2 = orc.annihilate(7, sdet, sval, sinfo)
    assert b["rc"] == 0
    for x, y in zip(sort_walkers(md, ms, mf), sort_walkers(md2, ms2, mf2)):
        np.testing.assert_array_equal(x, y)
    assert_stats(a, b, 1)
    n = assert_same_walkers(eng, orc)
    if case in ("empty", "cancelling"):  # nothing is born and the resident signs are untouched
        assert a["n_new"] == 0 and n == len(dets)
    if case == "one_new":
        assert a["n_new"] == 1 and n == len(dets) + 1
    eng.close()


def test_population_dies_out_then_iterates_empty():
    tau = 0.01
    spec = SPECS["hubbard10"](init_walkers=4000, seed=9)
    eng = engine(spec)
    orc = oracle(spec, tau)
    eng.start_single_part(30.0)
    orc.add_walkers(orc.ref[None, :], np.full((1, 1), 30.0), [1])
    for it in range(1, 4):  # grow a little first
        assert_stats(eng.iterate(it, tau, 0.3), orc.iterate(it, 0.3), 1)
    # H_ii - H_ref lies in [0, 20] (U = 4, up to 5 doubly occupied sites, none
    # in the Neel reference), so tau*(H_ii - H_ref + 80) is in [0.8, 1.0]
    empty_at = None
    for it in range(4, 80):
        a = eng.iterate(it, tau, -80.0)
        b = orc.iterate(it, -80.0)
        assert b["rc"] == 0
        assert_stats(a, b, 1)
        if a["tot_walkers"] == 0:
            assert a["tot_parts"][0] == 0.0
            if empty_at:  # the walkers alive at the start of the first empty iteration still spawned
                assert a["n_attempts"] == 0 and a["n_spawned"] == 0
            empty_at = empty_at or it
        if empty_at and it >= empty_at + 5:
            break
    assert empty_at is not None, "population did not die out"
    assert assert_same_walkers(eng, orc) == 0
    # and it can be refilled
    eng.start_single_part(10.0)
    orc.add_walkers(orc.ref[None, :], np.full((1, 1), 10.0), [1])
    a = eng.iterate(it + 1, tau, 0.3)
    assert_stats(a, orc.iterate(it + 1, 0.3), 1)
    eng.close()


def test_iterate_from_empty_upload():
    tau = 0.01
    spec = SPECS["mol33o4e_w2"](init_walkers=3000)
    eng = engine(spec)
    orc = oracle(spec, tau)
    eng.upload(eng.rows(np.zeros((0, spec.W), np.int64), np.zeros((0, 1))))
    for it in range(1, 4):
        a = eng.iterate(it, tau, 0.0)
        b = orc.iterate(it, 0.0)
        assert b["rc"] == 0
        assert_stats(a, b, 1)
        assert a["tot_walkers"] == 0 and a["n_attempts"] == 0
    eng.close()
