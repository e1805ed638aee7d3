#!/usr/bin/env python3
"""FCIQMC walker-iterations/s on one MI355X.

Workload: synthetic C1 molecular
Hamiltonian, 10 electrons in 20 spatial orbitals (seed 7, Cholesky rank 60),
initiator FCIQMC (threshold 3), PCHB doubles (UNIF-UNIF particles, FAST-FAST
spin-orbital-resolved holes) + on-the-fly heat-bath singles with pSingles 0.1,
1e6 target walkers, tau = 0.005, SHIFTDAMP 0.5, MEMORYFACHASH 2, FP64 signs.

A step is one FCIQMC iteration (spawn, merge, death, annihilation, hash
statistics) over the resident walker list; every 10 steps the host runs the
shift update (NECI's update cycle, with TARGET-SHIFTDAMP so that the
population settles at TOTALWALKERS).  The walker population is grown to the
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
STAGES = ["prepass", "scan", "spawn", "spawn_singles", "sort", "merge", "annihilate", "insert", "stats"]
KERNELS = {"prepass": "k_prepass", "spawn": "k_spawn", "spawn_singles": "k_spawn_singles",
           "annihilate": "k_merge_annihilate", "insert": "k_insert", "stats": "k_stats"}
SINGLE_KERNEL_STAGES = set(KERNELS)
HBM_PEAK_GBS = 8000.0


def sort_key(nbasis, nel):
    """(bytes per key, 8-bit passes through global memory) of the spawn sort:
    for the 32-bit colex-rank key (C(nbasis, nel) fits 32 bits and its
    binomial table the engine's LDS copy) the top-byte pass and the segment
    sort's read and write, counted as two passes; else the determinant words,
    one LSD pass per 8 bits."""
    import math
    if nel is not None and math.comb(nbasis, nel) <= 2 ** 32 and nbasis * (nel + 1) <= 512:
        return 4, 2
    W = (nbasis + 63) // 64
    return 8 * W, ((nbasis + 7) // 8 if W == 1 else (nbasis - 64 + 7) // 8 + 8)


def algorithmic_bytes(stage, c, W=1, S=1, nbasis=40, psingles=0.1, nel=None):
    """Compulsory HBM bytes of one iteration's launch(es) of a stage: every input
    read once and every output written once, for the per-iteration counts c
    (Nd occupied dets, Na attempts, Ns spawns, Nu unique spawned dets, Nn new
    dets).  Records: key 8W, signs 8S, flags 4.  The integral and heat-bath
    tables (under 1 MB here) stay in L2 and count once, i.e. zero; the PCHB
    alias rows (21 MB at 40 spin orbitals) do not fit one XCD's 4 MB L2, so
    every double attempt counts its alias entry as the minimum an alias
    draw reads: bias 8 + alias 4 + probability 8 bytes (T_gen of a PCHB
    double).  Queued single attempts and their share of the spawns are their
    expected values psingles*Na and psingles*Ns."""
    K, Sg = 8 * W, 8 * S
    Nd, Na, Ns, Nu, Nn = c["Nd"], c["Na"], c["Ns"], c["Nu"], c["Nn"]
    Nq, Nsq = psingles * Na, psingles * Ns
    if stage == "prepass":     # read key, signs, flags, diagH, offdiagH; write signs, flags, counts, pinfo
        return Nd * (K + Sg + 4 + 16) + Nd * (Sg + 4 + 4 * S + S)
    if stage == "scan":        # read counts, write offsets
        return Nd * S * (4 + 8)
    if stage == "spawn":       # parents once (key, pinfo, offset); one alias entry per double attempt;
        #                        queue (16 B) and double spawns written
        return Nd * (K + S + 8 * S) + (Na - Nq) * 20 + Nq * 16 + (Ns - Nsq) * (K + 12)
    if stage == "spawn_singles":  # queue entry, parent key/pinfo/offset, single spawns written
        return Nq * (16 + K + 1 + 8) + Nsq * (K + 12)
    if stage == "sort":        # per pass: histogram read of keys, scatter read+write of (key, index)
        kb, passes = sort_key(nbasis, nel)
        return passes * Ns * (kb + 2 * (kb + 4))
    if stage == "merge":       # fused into the annihilate stage
        return 0.0
    if stage == "annihilate":  # merge and annihilation in one pass: sorted keys + index, gathered value + info
        #                        per spawn; per unique det one hash bucket (64W B) and its state row
        #                        (S signs + flag word) read and written; new list written
        return Ns * (K + 4 + 8 + 4) + Nu * 64 * W + Nu * 2 * (Sg + 8) + Nn * (K + Sg + 4)
    if stage == "insert":      # new list, free-queue entry, bucket line, walker row (key, signs, flags, 3x8)
        return Nn * (K + Sg + 4) + Nn * 8 + Nn * 64 * W + Nn * (K + Sg + 4 + 24)
    if stage == "stats":       # flags, signs, key per slot
        return Nd * (4 + Sg + K)
    return 0.0


def iteration_bytes(c, W=1, S=1, nbasis=40, psingles=0.1, nel=None):
    """Compulsory HBM bytes of one whole iteration: the stages' figures of
    algorithmic_bytes summed, except that the sort counts one read and one
    write of (key, index) per spawn, the algorithmic minimum, not one per
    radix pass."""
    tot = sum(algorithmic_bytes(st, c, W, S, nbasis, psingles, nel) for st in STAGES if st != "sort")
    return tot + 2 * c["Ns"] * (sort_key(nbasis, nel)[0] + 4)


def _kernel_short(name):
    """'void ngpu::k_spawn<1, 1>(...)' -> 'k_spawn<1, 1>'."""
    n = name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").strip()
    return n.split("::")[-1]


def summarize_trace(path, last, per_iter_kernel="k_end_iter"):
    """Steady-state table from a rocprofv3 kernel-trace CSV: per kernel, the
    dispatches of the last `last` iterations (an iteration ends with one
    dispatch of per_iter_kernel), average and per-iteration duration."""
    import csv
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    ends = [i for i, r in enumerate(rows) if per_iter_kernel in r["Kernel_Name"]]
    if len(ends) <= last:
        raise SystemExit(f"only {len(ends)} iterations in {path}")
    window = rows[ends[-last - 1] + 1: ends[-1] + 1]
    agg = {}
    for r in window:
        k = _kernel_short(r["Kernel_Name"])
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        a = agg.setdefault(k, [0, 0.0])
        a[0] += 1
        a[1] += d
    total = sum(v[1] for v in agg.values()) / last
    print(f"| kernel | dispatches / iteration | average (us) | per iteration (us) | share |")
    print("|---|---|---|---|---|")
    for k, (n, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{k}` | {n / last:g} | {t / n:.1f} | {t / last:.1f} | {100 * t / last / total:.1f} % |")
    print(f"| all kernels | | | {total:.1f} | 100 % |")


def summarize_pmc(path, last_dispatches):
    """Average counter value per dispatch of each kernel over its last
    `last_dispatches` dispatches, from a rocprofv3 counter-collection CSV."""
    import csv
    rows = list(csv.DictReader(open(path)))
    per = {}
    for r in rows:
        key = (_kernel_short(r["Kernel_Name"]), r["Counter_Name"])
        per.setdefault(key, []).append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    print("| kernel | counter | dispatches | average per dispatch |")
    print("|---|---|---|---|")
    for (k, c), v in sorted(per.items(), key=lambda kv: -sum(x[1] for x in kv[1][-last_dispatches:])):
        v.sort()
        tail = [x[1] for x in v[-last_dispatches:]]
        print(f"| `{k}` | {c} | {len(tail)} | {sum(tail) / len(tail):.1f} |")


def _pmc_means(path, last_dispatches):
    import csv
    per = {}
    for r in csv.DictReader(open(path)):
        per.setdefault(_kernel_short(r["Kernel_Name"]), []).append((int(r["Dispatch_Id"]),
                                                                     float(r["Counter_Value"])))
    out = {}
    for k, v in per.items():
        v.sort()
        tail = [x[1] for x in v[-last_dispatches:]]
        out[k] = sum(tail) / len(tail)
    return out


def traffic_table(fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, copy_bytes, last_dispatches):
    """HBM bytes per launch from FETCH_SIZE and WRITE_SIZE (KiB per dispatch),
    each scaled by what the counter reads for a device copy of copy_bytes
    (read copy_bytes, write copy_bytes)."""
    cf = _pmc_means(cal_fetch_csv, 5)
    cw = _pmc_means(cal_write_csv, 5)
    kc = max(cf, key=cf.get)  # the copy kernel dominates its own run
    sf = copy_bytes / (cf[kc] * 1024.0)
    sw = copy_bytes / (cw[kc] * 1024.0)
    f = _pmc_means(fetch_csv, last_dispatches)
    w = _pmc_means(write_csv, last_dispatches)
    table = {"fetch_scale": sf, "write_scale": sw, "copy_bytes": copy_bytes, "kernels": {}}
    for k in sorted(set(f) & set(w)):
        name = k.split("<")[0]
        table["kernels"][name] = {"instance": k, "fetch_kib": f[k], "write_kib": w[k],
                                  "hbm_bytes_per_launch": sf * f[k] * 1024.0 + sw * w[k] * 1024.0}
    return table


def copy_calibration(nbytes, reps):
    """A plain device copy of known size, for scaling FETCH_SIZE/WRITE_SIZE."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def build_workload(args):
    from neci_stable_amd import Spec, synthetic_molecule
    mol = synthetic_molecule(20, 10, 60, 7)
    spec = Spec(mol, init_walkers=args.walkers, seed=1234, trunc_initiator=True, init_walk_no=3.0,
                particles=0, psingles=0.1, pparallel=0.5, memory_fac_spawn=8.0,
                memory_fac_part=10.0, memory_fac_hash=args.memory_fac_hash)
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


def batched_workload(n_mol_replicas=16, n_hub=16, seed=2024, mol_walkers=5e6, ncore=50000):
    """32 independent problems of uneven size: the 14e/30o molecular
    Hamiltonian (seed 7, rank 60) with n_mol_replicas replicas of distinct
    random streams and a core of the ncore most populated determinants, and
    n_hub half-filled Hubbard chains of 12-18 sites, U/t uniform in [2, 8],
    walker targets log-uniform in [1e4, 2e6]."""
    import numpy as np
    from neci_stable_amd import HubbardChain, Spec, synthetic_molecule
    rng = np.random.default_rng(seed)
    hubs = []
    for i in range(n_hub):
        L = int(rng.integers(12, 19))
        u = float(rng.uniform(2.0, 8.0))
        target = float(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
        hubs.append(dict(L=L, u=u, target=target, seed=seed + 1 + i))
    mol = synthetic_molecule(30, 14, 60, 7)
    hub_specs = [(Spec(HubbardChain(h["L"], h["L"], h["u"], bhub=-1.0), init_walkers=h["target"], seed=h["seed"],
                       memory_fac_part=4.0, memory_fac_spawn=4.0, memory_fac_hash=2.0), 0.01, h) for h in hubs]
    mol_kw = dict(particles=0, psingles=0.1, memory_fac_spawn=4.0, memory_fac_part=3.0, memory_fac_hash=2.0)
    return mol, mol_kw, hub_specs, mol_walkers, ncore


def run_batched(args):
    """The same engines iterated two ways, from the same grown populations:
    plain (each problem its own engine, iterated alone with one launch per
    kernel, one problem after the other) and batched (all engines as the
    branches of one captured graph per iteration).  The molecular replicas
    are engines of one replica each with distinct seeds."""
    import numpy as np
    from neci_stable_amd import Spec
    from neci_stable_amd.engine import Batch, Engine
    from neci_stable_amd.fcimc import BatchedFciMCPar, FciMCPar

    mol, mol_kw, hub_specs, mol_walkers, ncore = batched_workload(args.mol_replicas, args.hub_systems,
                                                                  mol_walkers=args.mol_walkers)
    tau_mol = 0.001

    def grown(spec, tau, target, n0):
        e = Engine(spec)
        e.start_single_part(n0)
        d = FciMCPar(e, tau=tau, target_walkers=target, steps_shift=10, shift_damp=0.5)
        while d.single_part.any() and d.iter < args.max_growth_iters:
            d.run(10)
        d.run(args.warmup)
        return e, d

    def with_core(e, d):
        dets, sg, _, _, _ = e.download()
        o = np.argsort(-np.abs(sg).sum(axis=1), kind="stable")[:ncore]
        core = np.ascontiguousarray(dets[np.sort(o)])
        e.set_core(core, fp32=False)
        d.run(10)
        return core

    def timed(fn, drivers, k):
        w0 = sum(d.walker_iters for d in drivers)
        t0 = time.perf_counter()
        fn(k)
        dt = time.perf_counter() - t0
        return (sum(d.walker_iters for d in drivers) - w0) / dt, dt

    K = args.steps
    engines = []
    cores = []
    for r in range(args.mol_replicas):
        spec = Spec(mol, init_walkers=mol_walkers, seed=7000 + r, **mol_kw)
        e, d = grown(spec, tau_mol, mol_walkers, 1e3)
        cores.append(with_core(e, d))
        engines.append((e, d))
    for spec, tau, h in hub_specs:
        engines.append(grown(spec, tau, h["target"], 30.0))
    print(f"# grown: {len(engines)} engines", file=sys.stderr, flush=True)
    drivers = [d for _, d in engines]
    # plain FP64: FP64 core Hamiltonian and spawned amplitudes
    p_rate, p_dt = timed(lambda k: [d.run(k) for d in drivers], drivers, K)
    # batched FP32/FP64: core Hamiltonian and spawned amplitudes stored in FP32,
    # sums, energies and the shift in FP64
    for i, (e, _) in enumerate(engines):
        if i < len(cores):
            e.set_core(cores[i], fp32=True)
        e.set_spawn_precision(True)
    batch = Batch([e for e, _ in engines])
    runner = BatchedFciMCPar(batch, drivers)
    runner.run(10)  # capture and first replays
    b_rate, b_dt = timed(runner.run, drivers, K)
    batch.close()
    walkers = [float(np.sum(d.last["tot_parts"])) for d in drivers]
    for e, _ in engines:
        e.close()
    line = {"metric": "walker-iterations/s on 1 GPU, batched vs plain FP64", "unit": "walker-iterations/s",
            "n_gpus": 1, "steps": K, "batched": b_rate, "plain": p_rate, "ratio": b_rate / p_rate,
            "ms_per_step": {"batched": b_dt * 1e3 / K, "plain": p_dt * 1e3 / K},
            "dtype": {"plain": "f64 walkers, core Hamiltonian and spawned amplitudes",
                      "batched": "f64 walkers and sums; f32-stored core Hamiltonian and spawned amplitudes"},
            "data": "synthetic",
            "config": {"workload": f"{args.mol_replicas} replicas of molecular 14e/30o (seed 7, rank 60, "
                                   f"{ncore}-det core, one engine each) + {len(hub_specs)} Hubbard chains "
                                   f"of 12-18 sites",
                       "hubbard": [dict(sites=h["L"], u=round(h["u"], 3), target=round(h["target"]))
                                   for _, _, h in hub_specs],
                       "walkers_per_problem": walkers}}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=50, help="steps of the all-stage timing pass")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--walkers", type=float, default=1e6)
    ap.add_argument("--memory-fac-hash", type=float, default=0.7,
                    help="MEMORYFACHASH: walker hash buckets per TOTALWALKERS (the reference's default 0.7)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-growth-iters", type=int, default=6000)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a device copy of this many bytes (PMC traffic scale)")
    ap.add_argument("--summarize-trace", default=None, help="rocprofv3 kernel-trace CSV to summarise")
    ap.add_argument("--summarize-pmc", default=None, help="rocprofv3 counter-collection CSV to summarise")
    ap.add_argument("--last", type=int, default=60, help="iterations (or dispatches) summarised")
    ap.add_argument("--batched", action="store_true",
                    help="batched (one graph per iteration) vs plain form on the 32-problem workload")
    ap.add_argument("--mol-replicas", type=int, default=16)
    ap.add_argument("--hub-systems", type=int, default=16)
    ap.add_argument("--mol-walkers", type=float, default=5e6)
    ap.add_argument("--traffic", nargs=4, default=None,
                    metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="HBM bytes per launch from counter runs, scaled by a copy calibration")
    args = ap.parse_args()
    if args.copy_calibration and not args.traffic:
        copy_calibration(args.copy_calibration, 5)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.last)
        return
    if args.traffic:
        print(json.dumps(traffic_table(*args.traffic, copy_bytes=args.copy_calibration or 268435456,
                                       last_dispatches=args.last), indent=1))
        return
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc, args.last)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one GPU per process (one rank's walkers); "
                                   "multi-GPU runs are not part of this build", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.batched:
        run_batched(args)
        return
    import numpy as np
    from neci_stable_amd.engine import Engine
    from neci_stable_amd.fcimc import FciMCPar

    spec = build_workload(args)
    tau = 0.005
    eng = Engine(spec)
    eng.start_single_part(100.0)
    # TARGET-SHIFTDAMP holds the population at TOTALWALKERS; the run is grown
    # until five consecutive update cycles average within 10 % of it
    drv = FciMCPar(eng, tau=tau, target_walkers=args.walkers, steps_shift=10, shift_damp=0.5,
                   target_shiftdamp=True)
    tg = time.perf_counter()
    settled = 0
    while settled < 5 and drv.iter < args.max_growth_iters:
        drv.run(10)
        av = float(np.sum(drv.rows[-1]["tot_parts"]))
        settled = settled + 1 if abs(av / args.walkers - 1.0) < 0.1 else 0
    t_growth = time.perf_counter() - tg
    growth_iters = drv.iter
    print(f"# growth: {drv.iter} iterations, {t_growth:.2f} s, TotParts {drv.rows[-1]['tot_parts']}",
          file=sys.stderr, flush=True)
    drv.run(args.warmup)
    print(f"# warmup done: TotParts {drv.rows[-1]['tot_parts']}, dets {drv.rows[-1]['tot_walkers']}",
          file=sys.stderr, flush=True)
    tuned = None
    if not args.no_autotune:
        from neci_stable_amd.autotune import autotune
        tuned = autotune(drv)
        print(f"# autotuned launch shape: {tuned['shape']}", file=sys.stderr, flush=True)
    # every stage timed over a short pass first (the breakdown, and the choice
    # of the roofline kernel); the timed region then carries events on that
    # kernel's launches only, since events on every launch slow the iteration
    eng.set_profiling(True)
    drv.run(args.profile_steps)
    ms_all, nprof_all = eng.stage_times()
    eng.set_profiling(False)
    per_stage = {STAGES[i]: ms_all[i] / max(nprof_all, 1) for i in range(len(STAGES))}
    dom = max((k for k in per_stage if k in SINGLE_KERNEL_STAGES), key=per_stage.get)
    nrow0 = len(drv.rows)
    eng.set_profiling(1 << STAGES.index(dom))
    w0 = drv.walker_iters
    t0 = time.perf_counter()
    drv.run(args.steps)
    dt = time.perf_counter() - t0
    walker_iters = drv.walker_iters - w0
    ms, nprof = eng.stage_times()
    eng.set_profiling(False)
    dom_ms = ms[STAGES.index(dom)] / max(nprof, 1)
    rows = drv.rows[nrow0:]
    K = args.steps
    counts = dict(Nd=float(np.mean([r["tot_walkers"] for r in rows])),
                  Na=float(sum(r["attempts"] for r in rows)) / K,
                  Ns=float(sum(r["spawned"] for r in rows)) / K,
                  Nu=float(sum(r["unique"] for r in rows)) / K,
                  Nn=float(sum(r["new"] for r in rows)) / K)
    # the roofline is that of one kernel: the slowest stage made of a single
    # kernel, its launches timed in the timed region itself
    b = algorithmic_bytes(dom, counts, W=eng.W, S=eng.S, nbasis=spec.nbasis, psingles=spec.psingles, nel=spec.nel)
    achieved = b / (dom_ms * 1e-3) / 1e9
    it_bytes = iteration_bytes(counts, W=eng.W, S=eng.S, nbasis=spec.nbasis, psingles=spec.psingles, nel=spec.nel)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_mol10e20o.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["kernels"].get(KERNELS[dom], {}).get("hbm_bytes_per_launch")
    line = {
        "metric": METRIC, "value": walker_iters / dt, "unit": "walker-iterations/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": dt * 1e3 / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "synthetic molecular 10e/20o (seed 7, rank 60), initiator 3, "
                               "PCHB unif-unif:fast-fast + heat-bath singles (psingles 0.1), tau 0.005, "
                               "target-shiftdamp 0.5, plain FP64",
                   "target_walkers": args.walkers,
                   "mean_tot_parts": walker_iters / K, "mean_determinants": counts["Nd"],
                   "attempts_per_step": counts["Na"], "spawns_per_step": counts["Ns"],
                   "unique_spawned_per_step": counts["Nu"], "new_dets_per_step": counts["Nn"],
                   "growth_iterations": growth_iters,
                   "launch_shape": eng.launch_shape(),
                   "autotuned": tuned is not None},
        "stage_ms_per_step": per_stage,
        "stage_ms_note": f"every stage timed over {args.profile_steps} steps before the timed region",
        "roofline": {"bound": "hbm", "kernel": KERNELS[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": b, "launch_ms": dom_ms},
        # the whole iteration against the same peak: compulsory bytes of every
        # stage per iteration over the measured time per iteration
        "iteration_roofline": {"bound": "hbm", "achieved": it_bytes / (dt / K) / 1e9, "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": it_bytes / (dt / K) / 1e9 / HBM_PEAK_GBS,
                               "algorithmic_bytes_per_iteration": it_bytes},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(spec, eng, drv, args.cpu_seconds)
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
