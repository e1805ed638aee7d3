"""Full-size molecular runs, checked through properties that hold at any size.

* 26 electrons in 40 spatial orbitals (80 spin orbitals, two-word
  determinants; synthetic Hamiltonian rank 100, seed 11), PCHB doubles,
  5e7 walkers, spawn buffer 1.5x the walkers, 2^27 hash buckets, initiator 3,
  a 20 000-determinant semistochastic core with FP32-stored Hamiltonian.
* 14 electrons in 30 spatial orbitals (seed 7, rank 60), 5e6 walkers, a
  50 000-determinant core of the most populated determinants: projected
  energies with the core Hamiltonian in FP64 and stored in FP32 agree within
  their blocking error bars.

After every phase: no engine error, the determinants downloaded are unique
and carry nel electrons, TotParts equals the sum of |sign| over them (exactly
for integer walkers; to 1e-12 relative once core amplitudes, which are real,
are summed), and every core determinant is resident and flagged.
"""
import time

import numpy as np
import pytest

from neci_stable_amd import Spec, synthetic_molecule
from neci_stable_amd.fcimc import FciMCPar, reblock

pytestmark = pytest.mark.gpu

DETERM_BIT = 41  # flag bit of a core determinant, run 1, in the engine's default layout (bit_rep_data.F90:102-118)


def as_records(d):
    """(n, W) int64 determinant words as one structured value per row, which
    numpy sorts and searches lexicographically."""
    d = np.ascontiguousarray(d, dtype=np.int64)
    return d.view([(f"w{w}", "<i8") for w in range(d.shape[1])]).ravel()


def check_state(eng, spec, last, core=None):
    d, s, f, _, _ = eng.download()
    rec = np.sort(as_records(d))
    assert not np.any(rec[1:] == rec[:-1]), "a determinant is stored twice"
    assert np.all(np.bitwise_count(d.view(np.uint64)).sum(axis=1) == spec.nel)
    tot = np.abs(s).sum(axis=0)
    if core is None:  # integer walkers: the sum is exact in any order
        np.testing.assert_array_equal(tot, last["tot_parts"][:spec.S])
    else:  # core amplitudes are real: summation order shows in the last bits
        np.testing.assert_allclose(tot, last["tot_parts"][:spec.S], rtol=1e-12)
    if core is not None:
        order = np.argsort(as_records(d), kind="stable")
        srt = as_records(d)[order]
        c = as_records(core)
        at = np.clip(np.searchsorted(srt, c), 0, len(srt) - 1)
        assert np.all(srt[at] == c), "core determinant missing from the walker list"
        assert np.all((f[order[at]] >> DETERM_BIT) & 1)
    return d, s


def grow(eng, tau, target, n0, shift_damp=0.5, max_iter=4000, t_max=240.0):
    drv = FciMCPar(eng, tau=tau, target_walkers=target, steps_shift=10, shift_damp=shift_damp)
    eng.start_single_part(n0)
    t0 = time.time()
    while drv.single_part.any():
        drv.run(10)
        assert drv.iter < max_iter and time.time() - t0 < t_max, (drv.iter, drv.rows[-1]["tot_parts"])
        if drv.iter % 200 == 0:
            print(f"  iteration {drv.iter}: {drv.rows[-1]['tot_parts']} walkers", flush=True)
    print(f"grown to {drv.rows[-1]['tot_parts']} walkers on {drv.rows[-1]['tot_walkers']} dets "
          f"in {drv.iter} iterations, {time.time() - t0:.1f} s", flush=True)
    return drv


def check_top_weights(core, d, s, n):
    """The device's choice holds the n largest weights sum_r |c_r| (equal
    weights at the boundary may be other determinants than a storage-order
    choice would take)."""
    w = np.abs(s).sum(axis=1)
    dt = np.dtype([(f"w{i}", np.int64) for i in range(d.shape[1])])
    dv = np.ascontiguousarray(d).view(dt).ravel()
    cv = np.ascontiguousarray(core).view(dt).ravel()
    o = np.argsort(dv)
    pos = np.searchsorted(dv[o], cv)
    assert len(core) == n and np.all(pos < len(dv)) and np.array_equal(dv[o][pos], cv)
    chosen = np.sort(w[o][pos])
    top = np.sort(np.partition(w, len(w) - n)[len(w) - n:])
    np.testing.assert_array_equal(chosen, top)


def most_populated(d, s, n):
    o = np.argsort(-np.abs(s).sum(axis=1), kind="stable")[:n]
    return np.ascontiguousarray(d[np.sort(o)])


def test_two_word_26e40o_5e7_walkers_fp32_core():
    from neci_stable_amd.engine import Engine
    mol = synthetic_molecule(40, 26, 100, 11)
    spec = Spec(mol, init_walkers=5e7, seed=11, particles=0, psingles=0.1, memory_fac_spawn=1.5,
                memory_fac_part=3.0, n_hash_buckets=2 ** 27)
    assert spec.W == 2
    eng = Engine(spec)
    drv = grow(eng, 0.0002, 5e7, 1e5)
    d, s = check_state(eng, spec, drv.last)
    t0 = time.time()
    core, nnz = eng.select_core(20000, fp32=True)  # chosen and built on the device
    print(f"core: {len(core)} dets chosen and built in {time.time() - t0:.2f} s, {nnz} stored elements (FP32)",
          flush=True)
    check_top_weights(core, d, s, 20000)
    eng.set_spawn_precision(True)  # FP32-stored spawned amplitudes, FP64 sums
    t0 = time.time()
    w0 = drv.walker_iters
    drv.run(20)
    dt = time.time() - t0
    print(f"26e/40o, 5e7 walkers, 20 iterations: {dt / 20 * 1e3:.2f} ms/iteration, "
          f"{(drv.walker_iters - w0) / dt:.3g} walker-iterations/s", flush=True)
    assert drv.last["tot_parts"][0] > 4e7
    check_state(eng, spec, drv.last, core)
    eng.close()


def test_14e30o_core50k_fp32_matches_fp64():
    from neci_stable_amd.engine import Engine
    mol = synthetic_molecule(30, 14, 60, 7)
    spec = Spec(mol, init_walkers=5e6, seed=7, particles=0, psingles=0.1, memory_fac_spawn=6.0,
                memory_fac_part=8.0, memory_fac_hash=4.0)
    tau = 0.001
    eng = Engine(spec)
    drv = grow(eng, tau, 5e6, 1e3)
    drv.run(200)
    d, s = check_state(eng, spec, drv.last)
    f = eng.download()[2]
    core, _ = eng.select_core(50000)  # chosen on the device
    check_top_weights(core, d, s, 50000)
    start_rows = eng.rows(d, s, f)
    shift0 = drv.shift.copy()
    it0 = drv.iter
    eng.close()
    res = {}
    finals = {}
    for fp32 in (False, True):
        e = Engine(spec)
        e.upload(start_rows)
        nnz = e.set_core(core, fp32=fp32)
        run = FciMCPar(e, tau=tau, target_walkers=5e6, steps_shift=10, shift_damp=0.5, diag_shift=0.0)
        run.shift[:] = shift0
        run.single_part[:] = False
        run.iter = it0
        t0 = time.time()
        run.run(1000)
        dt = time.time() - t0
        pe = np.array([r["proj_e"][0] for r in run.rows[20:]])  # correlation energy per cycle
        mean, err = reblock(pe)
        res[fp32] = (mean, err)
        print(f"core FP{'32' if fp32 else '64'}: nnz {nnz}, E_proj {mean:.10f} +- {err:.6f}, "
              f"{dt:.1f} s for 1000 iterations", flush=True)
        dd, ss = check_state(e, spec, run.last, core)
        finals[fp32] = (dd, ss)
        e.close()
    # the FP32-stored Hamiltonian changes the core amplitudes only at FP32 rounding level
    (d64, s64), (d32, s32) = finals[False], finals[True]
    if d64.shape == d32.shape and np.array_equal(d64, d32):
        rel = np.max(np.abs(s64 - s32)) / np.max(np.abs(s64))
        print(f"same walker set after 1000 iterations; max |c64 - c32| / max |c64| = {rel:.3e}", flush=True)
        assert rel < 1e-5
    (m64, e64), (m32, e32) = res[False], res[True]
    assert abs(m32 - m64) < 3.0 * np.hypot(e32, e64) + 1e-9, res

