"""Capacity and stability limits, forced from the host so that every kernel
still exits normally, engine against the oracle (which stops with the same
status as the reference's stop_all in each case):
* death factor tau*(H_ii - S) above 2 (attempt_die_normal,
  fcimc_pointed_fns.F90:649-663): NGPU_ERR_DEATH_PROB;
* a spawned list longer than MaxSpawned: NGPU_ERR_SPAWN_FULL;
* more determinants than MaxWalkers, from an upload and from growth:
  NGPU_ERR_WALKERS_FULL, at the same iteration as the oracle, with identical
  statistics before it.
"""
import numpy as np
import pytest

from test_gpu_parity import SPECS, assert_stats, engine, oracle

pytestmark = pytest.mark.gpu
TAU = 0.01


def _start(spec, n0):
    eng = engine(spec)
    orc = oracle(spec, TAU)
    eng.start_single_part(n0)
    orc.add_walkers(orc.ref[None, :], np.full((1, 1), n0), [1])
    return eng, orc


def _fails_like_oracle(eng, orc, shift, code, max_iter=10):
    from neci_stable_amd._lib import NgpuError
    for it in range(1, max_iter + 1):
        b = orc.iterate(it, shift)
        if b["rc"]:
            assert b["rc"] == code
            with pytest.raises(NgpuError) as e:
                eng.iterate(it, TAU, shift)
            assert e.value.code == code
            return it
        assert_stats(eng.iterate(it, TAU, shift), b, 1)
    raise AssertionError("the oracle never reached the limit")


def test_death_probability_above_two():
    # H_ii - H_ref >= 0 for the Hubbard chain, so tau*(H_ii + 300) >= 3
    eng, orc = _start(SPECS["hubbard10"](init_walkers=4000, seed=9), 30.0)
    assert _fails_like_oracle(eng, orc, -300.0, 3) == 1
    eng.close()


def test_spawn_buffer_full():
    eng, orc = _start(SPECS["hubbard10"](init_walkers=4000, seed=9, max_spawned=64), 3000.0)
    assert _fails_like_oracle(eng, orc, 0.0, 1) == 1
    eng.close()


def test_walker_list_full_from_growth():
    eng, orc = _start(SPECS["hubbard10"](init_walkers=4000, seed=9, max_walkers=40), 3000.0)
    it = _fails_like_oracle(eng, orc, 0.0, 2)
    print(f"walker list full at iteration {it}", flush=True)
    eng.close()


def test_walker_list_full_from_upload():
    from neci_stable_amd._lib import NgpuError
    from test_gpu_parity import random_dets
    spec = SPECS["hubbard10"](init_walkers=4000, max_walkers=40)
    eng = engine(spec)
    dets = np.unique(random_dets(spec, 80, np.random.default_rng(5)), axis=0)
    assert len(dets) > 40
    with pytest.raises(NgpuError) as e:
        eng.upload(eng.rows(dets, np.ones((len(dets), 1))))
    assert e.value.code == 2
    eng.close()


def test_async_iterations_halt_at_the_failing_iteration():
    """Iterations queued without waiting (as the host drivers run them): the
    status arrives at the next read of the cycle statistics, the engine
    reports the oracle's failing iteration (ngpu_get_error), and the walker
    list is the one the failing iteration left, however many iterations were
    queued after it."""
    from neci_stable_amd._lib import NgpuError
    spec = SPECS["hubbard10"](init_walkers=4000, seed=9, max_walkers=40)
    eng, orc = _start(spec, 3000.0)
    ref = engine(spec)
    ref.start_single_part(3000.0)
    it_fail = None
    for it in range(1, 11):
        if orc.iterate(it, 0.0)["rc"]:
            it_fail = it
            break
    assert it_fail is not None
    for it in range(1, it_fail + 1):  # the same engine state, run to the failure with waits
        if it < it_fail:
            ref.iterate(it, TAU, 0.0)
        else:
            with pytest.raises(NgpuError):
                ref.iterate(it, TAU, 0.0)
    assert eng.error() == (0, 0)
    for it in range(1, it_fail + 6):
        eng.iterate(it, TAU, 0.0, wait=False)
    with pytest.raises(NgpuError) as e:
        eng.cycle_stats()
    assert e.value.code == 2
    assert eng.error() == (2, it_fail)
    with pytest.raises(NgpuError):  # still halted in the next cycle
        eng.iterate(it_fail + 6, TAU, 0.0)
    a = eng.download()
    b = ref.download()
    from common import sort_walkers
    for x, y in zip(sort_walkers(*a[:3]), sort_walkers(*b[:3])):
        np.testing.assert_array_equal(x, y)
    eng.close()
    ref.close()

