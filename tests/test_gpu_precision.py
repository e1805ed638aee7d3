"""FP32 storage of the spawned amplitudes and of the core Hamiltonian against
the all-FP64 engine, on a real-coefficient run (ALLREALCOEFF, where spawned
amplitudes are not integers and FP32 storage rounds them).

Synthetic molecular 10e/20o (seed 7, rank 60), PCHB doubles with heat-bath
singles, 1e5 walkers held by TARGET-SHIFTDAMP, a 2000-determinant
semistochastic core.  From one grown walker list:
* one iteration each way from the same start: the same determinants, and
  amplitudes that differ at FP32 rounding level (and do differ);
* 2000 iterations each way: shift energies within 2 sigma (blocking
  analysis) of each other.
"""
import numpy as np
import pytest

from neci_stable_amd import Spec, synthetic_molecule
from neci_stable_amd.fcimc import FciMCPar, reblock

pytestmark = pytest.mark.gpu


def _spec():
    mol = synthetic_molecule(20, 10, 60, 7)
    return Spec(mol, init_walkers=1e5, seed=5, particles=0, psingles=0.1, all_real_coeff=True,
                memory_fac_spawn=12.0, memory_fac_part=12.0, memory_fac_hash=6.0)


def _sorted(eng):
    d, s, _, _, _ = eng.download()
    o = np.lexsort(d.T[::-1])
    return d[o], s[o]


def test_fp32_spawns_and_core_against_fp64():
    from neci_stable_amd.engine import Engine
    tau = 0.005
    spec = _spec()
    eng = Engine(spec)
    eng.start_single_part(100.0)
    drv = FciMCPar(eng, tau=tau, target_walkers=1e5, steps_shift=10, shift_damp=0.5, target_shiftdamp=True)
    settled = 0  # grown until five consecutive update cycles average within 10 % of the target
    while settled < 5:
        drv.run(10)
        assert drv.iter < 3000
        tot = float(drv.rows[-1]["tot_parts"][0])
        settled = settled + 1 if abs(tot / 1e5 - 1.0) < 0.1 else 0
        if drv.iter % 100 == 0:
            print(f"  iteration {drv.iter}: {tot:.1f} walkers", flush=True)
    drv.run(100)
    d, s, f, _, _ = eng.download()
    start = eng.rows(d, s, f)
    core = np.ascontiguousarray(d[np.sort(np.argsort(-np.abs(s[:, 0]), kind="stable")[:2000])])
    shift0, it0 = drv.shift.copy(), drv.iter
    eng.close()
    print(f"grown: {len(d)} determinants, {np.abs(s).sum():.1f} walkers at iteration {it0}", flush=True)

    def fresh(fp32):
        e = Engine(_spec())
        e.upload(start)
        e.set_core(core, fp32=fp32)
        e.set_spawn_precision(fp32)
        r = FciMCPar(e, tau=tau, target_walkers=1e5, steps_shift=10, shift_damp=0.5, diag_shift=0.0,
                     target_shiftdamp=True)
        r.shift[:] = shift0
        r.single_part[:] = False
        r.iter = it0
        return e, r

    # one iteration from the same start
    one = {}
    for fp32 in (False, True):
        e, r = fresh(fp32)
        r.run(1)
        one[fp32] = _sorted(e)
        e.close()
    (d64, s64), (d32, s32) = one[False], one[True]
    np.testing.assert_array_equal(d64, d32)
    diff = np.abs(s32 - s64)
    assert np.count_nonzero(diff) > 100
    rel = float(np.max(diff / (np.abs(s64) + 1.0)))
    print(f"one iteration: {np.count_nonzero(diff)} of {len(s64)} amplitudes differ, "
          f"max |c32 - c64| / (|c64| + 1) = {rel:.2e}", flush=True)
    assert rel < 1e-5

    # energies over 2000 iterations: the shift (blocked over update cycles)
    # and the projected energy as the ratio of the accumulated ENumCyc and
    # HFCyc sums
    res = {}
    for fp32 in (False, True):
        e, r = fresh(fp32)
        r.run(2000)
        sh = np.array([row["shift"][0] for row in r.rows[20:]])
        res[fp32] = reblock(sh)
        pe = float(r.sum_enum[0] / r.sum_noathf[0])
        print(f"{'FP32' if fp32 else 'FP64'} spawns and core: shift {res[fp32][0]:.10f} "
              f"+- {res[fp32][1]:.6f}, E_proj (ratio of sums) {pe:.10f}", flush=True)
        e.close()
    (m64, e64), (m32, e32) = res[False], res[True]
    assert abs(m32 - m64) <= 2.0 * np.hypot(e32, e64) + 1e-9, res
