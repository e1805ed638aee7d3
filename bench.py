#!/usr/bin/env python3
"""FCIQMC walker-iterations/s on one MI355X.

Workload (config 1 of the measurement plan): synthetic C1 molecular
Hamiltonian, 10 electrons in 20 spatial orbitals (seed 7, Cholesky rank 60),
initiator FCIQMC (threshold 3), PCHB doubles (UNIF-UNIF particles, FAST-FAST
spin-orbital-resolved holes) + on-the-fly heat-bath singles with pSingles 0.1,
1e6 target walkers, tau = 0.005, SHIFTDAMP 0.5, MEMORYFACHASH 2, FP64 signs.

A step is one FCIQMC iteration (spawn, merge, death, annihilation, hash
statistics) over the resident walker list; every 10 steps the host runs the
shift update (NECI's update cycle).  The walker population is grown to the
target untimed, then W warm-up steps, then K timed steps.
value = sum over timed steps of TotParts / wall time of those steps.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "walker-iterations/s (spawn+annihilate+death) on 1 GPU; batched vs plain FP64"
STAGES = ["prepass", "scan", "spawn", "sort", "merge", "annihilate", "insert", "stats"]
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(stage, c, W=1, S=1, nel=10, nsp=20, psingles=0.1, nbasis=40):
    """Minimum compulsory bytes of one launch of a stage for per-iteration counts c.

    N_d occupied dets, N_a attempts, N_s spawns, N_u unique spawned dets,
    N_new inserted dets.  Records: key 8W, sign 8S, flags 4."""
    K = 8 * W
    Nd, Na, Ns, Nu, Nn = c["Nd"], c["Na"], c["Ns"], c["Nu"], c["Nn"]
    if stage == "prepass":   # read key, signs, flags, diagH, offdiagH; write signs, flags, counts, pinfo
        return Nd * (K + 8 * S + 4 + 16) + Nd * (8 * S + 4 + 4 * S + S)
    if stage == "scan":
        return Nd * S * (4 + 8)
    if stage == "spawn":
        nvirt = nsp - nel // 2
        single = nvirt * (8 + 2 * nel * 8) + (8 + 2 * nel * 8)   # heat-bath list + H_ij gathers
        double = 16 + 8 + 8 + 16                                 # AB alias entry, prob, tgtOrbs, 2 UMAT gathers
        per_attempt = K + 8 + 1 + psingles * single + (1 - psingles) * double
        return Na * per_attempt + Ns * (K + 8 + 4)
    if stage == "sort":      # one read+write pass of (key, index)
        return 2 * Ns * (K + 4)
    if stage == "merge":
        return Ns * (K + 4 + 8 + 4) + Nu * (K + 8 * S + 4)
    if stage == "annihilate":
        return Nu * (K + 8 * S + 4) + Nu * 64 + Nu * 8 * S + Nn * (K + 8 * S + 4)
    if stage == "insert":
        return Nn * (K + 8 * S + 4) + Nn * (K + 8 * S + 4 + 24) + Nn * 64
    if stage == "stats":
        return Nd * (K + 8 * S + 4)
    return 0


def build_workload(args):
    from neci_stable_amd import Spec, synthetic_molecule
    mol = synthetic_molecule(20, 10, 60, 7)
    spec = Spec(mol, init_walkers=args.walkers, seed=1234, trunc_initiator=True, init_walk_no=3.0,
                particles=0, psingles=0.1, pparallel=0.5, memory_fac_spawn=8.0,
                memory_fac_part=10.0, memory_fac_hash=2.0)
    return spec


def cpu_baseline(spec, eng, drv, seconds):
    """The oracle (single-threaded CPU restatement) on the walker set the GPU
    reached, for as many iterations as fit in the time budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from oracle_ctypes import Oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from common import neci_to_internal_flags
    dets, signs, flags, _, _ = eng.download()
    orc = Oracle(spec, drv.tau)
    t_load = time.perf_counter()
    orc.add_walkers_bulk(dets, signs, neci_to_internal_flags(flags, spec.S))
    t_load = time.perf_counter() - t_load
    shift = drv.shift.copy()
    it0 = drv.iter + 1
    walker_iters = 0.0
    n = 0
    t0 = time.perf_counter()
    while True:
        st = orc.iterate(it0 + n, shift)
        if st["rc"]:
            raise RuntimeError(f"oracle iteration failed: {st['rc']}")
        walker_iters += float(np.sum(st["tot_parts"]))
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return dict(value=walker_iters / dt, unit="walker-iterations/s", cores=1, kind="port",
                sample=f"{n} iterations of the same workload from the GPU's walker set "
                       f"({len(dets)} determinants, {walker_iters / n:.3g} walkers/iteration), "
                       f"single-threaded oracle, {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--walkers", type=float, default=1e6)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-growth-iters", type=int, default=6000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one GPU per process (one rank's walkers); "
                                   "multi-GPU runs are not part of this build", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    from neci_stable_amd.engine import Engine
    from neci_stable_amd.fcimc import FciMCPar

    spec = build_workload(args)
    tau = 0.005
    eng = Engine(spec)
    eng.start_single_part(100.0)
    drv = FciMCPar(eng, tau=tau, target_walkers=args.walkers, steps_shift=10, shift_damp=0.5)
    tg = time.perf_counter()
    while drv.single_part.any() and drv.iter < args.max_growth_iters:
        drv.run(10)
    t_growth = time.perf_counter() - tg
    print(f"# growth: {drv.iter} iterations, {t_growth:.2f} s, TotParts {drv.rows[-1]['tot_parts']}",
          file=sys.stderr, flush=True)
    drv.run(args.warmup)
    print(f"# warmup done: TotParts {drv.rows[-1]['tot_parts']}, dets {drv.rows[-1]['tot_walkers']}",
          file=sys.stderr, flush=True)
    nrow0 = len(drv.rows)
    eng.set_profiling(True)
    w0 = drv.walker_iters
    t0 = time.perf_counter()
    drv.run(args.steps)
    dt = time.perf_counter() - t0
    walker_iters = drv.walker_iters - w0
    ms, nprof = eng.stage_times()
    eng.set_profiling(False)
    rows = drv.rows[nrow0:]
    K = args.steps
    counts = dict(Nd=float(np.mean([r["tot_walkers"] for r in rows])),
                  Na=float(sum(r["attempts"] for r in rows)) / K,
                  Ns=float(sum(r["spawned"] for r in rows)) / K,
                  Nu=float(sum(r["unique"] for r in rows)) / K,
                  Nn=float(sum(r["new"] for r in rows)) / K)
    per_stage = {STAGES[i]: ms[i] / max(nprof, 1) for i in range(len(STAGES))}
    dom = max(per_stage, key=per_stage.get)
    b = algorithmic_bytes(dom, counts, psingles=spec.psingles)
    achieved = b / (per_stage[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom}_config1.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    line = {
        "metric": METRIC, "value": walker_iters / dt, "unit": "walker-iterations/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": dt * 1e3 / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "synthetic molecular 10e/20o (seed 7, rank 60), initiator 3, "
                               "PCHB unif-unif:fast-fast + heat-bath singles (psingles 0.1), tau 0.005, plain FP64",
                   "target_walkers": args.walkers,
                   "mean_tot_parts": walker_iters / K, "mean_determinants": counts["Nd"],
                   "attempts_per_step": counts["Na"], "spawns_per_step": counts["Ns"],
                   "growth_iterations": drv.iter - K - args.warmup},
        "stage_ms_per_step": per_stage,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": b},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(spec, eng, drv, args.cpu_seconds)
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
