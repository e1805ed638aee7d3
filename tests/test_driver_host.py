"""Host driver accounting (CPU only): FciMCPar over a scripted backend with
the engine's cycle API.

* A run that stops inside an update cycle keeps that cycle's sums; the next
  run finishes the cycle and the shift update averages over all StepsSft
  iterations (calculate_new_shift_wrapper, fcimc_iter_utilities.F90:929-933).
* The running totals count every iteration exactly once, so per-step figures
  over a window of iterations (the bench's timed steps after its warm-up) are
  exact whatever the window's alignment to update cycles: with 5 warm-up and
  20 timed steps, attempts per step equal the walkers per step.
* The batched driver reads every driver at the end of each of its update
  cycles, also when the drivers' cycles are not aligned.
"""
import numpy as np
import pytest

from neci_stable_amd.fcimc import FciMCPar, merge_cycle_sums


class ScriptedBackend:
    """Iteration it has TotParts = 1000 + it walkers, as many attempts, and
    max_cyc_spawn = it; cycle_stats() returns the engine's cycle sums."""

    S = 1
    hii = 0.0

    def __init__(self):
        self.sum = None
        self.n = 0
        self.last = None
        self.iters = []

    def iterate(self, it, tau, shift, wait=False):
        self.iters.append(it)
        w = 1000.0 + it
        st = dict(tot_parts=np.array([w]), nborn=np.array([1.0]), ndied=np.array([0.0]),
                  nannihil=np.array([0.0]), naborted=np.array([0.0]), enum_cyc=np.array([-w]),
                  hf_cyc=np.array([w]), inst_no_at_hf=np.array([float(it)]), max_cyc_spawn=float(it),
                  bloom_size=np.array([0.0, 0.0, float(it)]), tot_walkers=it, n_attempts=int(w),
                  n_spawned=it, n_unique=1, n_new=2, err=0)
        self.last = st
        self.sum = st if self.sum is None else merge_cycle_sums(self.sum, st)
        self.n += 1

    def cycle_stats(self):
        out = (self.sum, self.last, self.n)
        self.sum, self.n = None, 0
        return out


def test_merge_rules():
    a = dict(tot_parts=np.array([1.0]), max_cyc_spawn=3.0, bloom_size=np.array([1.0, 5.0, 0.0]),
             tot_walkers=7, inst_no_at_hf=np.array([2.0]), n_attempts=4, err=1)
    b = dict(tot_parts=np.array([2.0]), max_cyc_spawn=2.0, bloom_size=np.array([3.0, 1.0, 0.0]),
             tot_walkers=9, inst_no_at_hf=np.array([5.0]), n_attempts=6, err=4)
    m = merge_cycle_sums(a, b)
    assert m["tot_parts"][0] == 3.0 and m["n_attempts"] == 10
    assert m["max_cyc_spawn"] == 3.0 and list(m["bloom_size"]) == [3.0, 5.0, 0.0]
    assert m["tot_walkers"] == 9 and m["inst_no_at_hf"][0] == 5.0 and m["err"] == 5


@pytest.mark.parametrize("warmup,steps", [(5, 20), (3, 17), (10, 10)])
def test_partial_cycles_and_totals(warmup, steps):
    b = ScriptedBackend()
    seen = []
    d = FciMCPar(b, tau=0.01, target_walkers=10.0, steps_shift=10, shift_damp=0.1,
                 log=lambda row: seen.append(row))
    d.run(warmup)
    t0 = dict(d.totals)
    d.run(steps)
    n = warmup + steps
    assert b.iters == list(range(1, n + 1))
    # every update cycle covers all ten of its iterations
    assert [r["iter"] for r in seen] == list(range(10, n + 1, 10))
    for r in seen:
        its = np.arange(r["iter"] - 9, r["iter"] + 1)
        assert r["attempts"] == int(np.sum(1000 + its))
        assert r["spawned"] == int(np.sum(its))
    # the timed window's per-step counts are exact
    dt = {k: d.totals[k] - t0[k] for k in d.totals}
    its = np.arange(warmup + 1, n + 1)
    assert dt["iters"] == steps
    assert dt["attempts"] == int(np.sum(1000 + its))
    assert dt["walker_iters"] == float(np.sum(1000 + its))
    assert dt["attempts"] / steps == pytest.approx(dt["walker_iters"] / steps)
    assert d.walker_iters == float(np.sum(1000 + np.arange(1, n + 1)))


class ScriptedBatch:
    """ngpu_batch_iterate over scripted backends."""

    def __init__(self, backends):
        self.backends = backends

    def iterate(self, iters, taus, shifts):
        for b, it, t, sh in zip(self.backends, iters, taus, shifts):
            b.iterate(it, t, sh)


def test_batched_driver_with_unaligned_cycles():
    from neci_stable_amd.fcimc import BatchedFciMCPar
    backends = [ScriptedBackend() for _ in range(3)]
    seen = [[] for _ in backends]
    drivers = [FciMCPar(b, tau=0.01, target_walkers=10.0, steps_shift=10, shift_damp=0.1,
                        log=lambda row, k=k: seen[k].append(row["iter"])) for k, b in enumerate(backends)]
    drivers[0].run(13)  # the first driver three iterations into a cycle
    drivers[1].run(5)   # the second five
    runner = BatchedFciMCPar(ScriptedBatch(backends), drivers)
    runner.run(37)
    assert [d.iter for d in drivers] == [50, 42, 37]
    assert seen[0] == [10, 20, 30, 40, 50]
    assert seen[1] == [10, 20, 30, 40]
    assert seen[2] == [10, 20, 30]
    for k, d in enumerate(drivers):
        assert d.totals["iters"] == d.iter
