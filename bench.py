#!/usr/bin/env python3
"""FCIQMC walker-iterations/s on one MI355X, batched against plain FP64.

Default workload (the metric's configuration): 32 independent problems of
uneven size -- 16 replicas (distinct seeds) of the synthetic C1 molecular
Hamiltonian with 14 electrons in 30 spatial orbitals (seed 7, Cholesky rank
60; initiator threshold 3, PCHB doubles with UNIF-UNIF particles and
FAST-FAST holes, heat-bath singles with pSingles 0.1, tau 0.001 (0.005 is
past this Hamiltonian's stable time step), 5e6 walkers each, a
semistochastic core of the 50 000 most populated determinants) and
16 half-filled Hubbard chains of 12-18 sites (U/t in [2, 8], walker targets
log-uniform in [1e4, 2e6], tau 0.01).

A step is one FCIQMC iteration of every problem (spawn, merge, death,
annihilation, hash statistics; the core projection for the molecular ones).
Every 10 steps the host runs each problem's update cycle (TARGET-SHIFTDAMP
for the molecular replicas).  Populations are set up untimed, then W warm-up
steps, then K timed steps, once in the plain FP64 form and once in the
batched FP32/FP64 form.  value = the batched form's sum over timed steps and
problems of TotParts / wall time of those steps.

--workload mol10e20o runs one 10e/20o molecular engine (1e6 walkers) in plain
FP64 instead.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "walker-iterations/s (spawn+annihilate+death) on 1 GPU; batched vs plain FP64"
STAGES = ["prepass", "scan", "spawn", "spawn_singles", "sort", "merge", "annihilate", "insert", "stats", "core"]
KERNELS = {"prepass": "k_prepass", "spawn": "k_spawn", "spawn_singles": "k_spawn_singles",
           "annihilate": "k_merge_annihilate", "insert": "k_insert", "stats": "k_stats", "core": "k_core_spmv"}
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


def algorithmic_bytes(stage, c, W=1, S=1, nbasis=40, psingles=0.1, nel=None, val_bytes=8):
    """Compulsory HBM bytes of one iteration's launch(es) of a stage: every input
    read once and every output written once, for the per-iteration counts c
    (Nd occupied dets, Na attempts, Ns spawns, Nu unique spawned dets, Nn new
    dets; Nc stored core Hamiltonian elements and Ncore core dets, when the
    run has a core space).  Records: key 8W, signs 8S, flags 4; a spawned
    amplitude and a core Hamiltonian element take val_bytes (8 in FP64, 4
    when stored in FP32).  A hash probe reads the matching entry's key and
    slot index, K + 8 bytes, not the whole bucket line.  The integral and heat-bath
    tables (under 1 MB here) stay in L2 and count once, i.e. zero; the PCHB
    alias rows (21 MB at 40 spin orbitals) do not fit one XCD's 4 MB L2, so
    every double attempt counts its alias entry as the minimum an alias
    draw reads: bias 8 + alias 4 + probability 8 bytes (T_gen of a PCHB
    double).  Queued single attempts and their share of the spawns are their
    expected values psingles*Na and psingles*Ns."""
    K, Sg, V = 8 * W, 8 * S, val_bytes
    Nd, Na, Ns, Nu, Nn = c["Nd"], c["Na"], c["Ns"], c["Nu"], c["Nn"]
    Nq, Nsq = psingles * Na, psingles * Ns
    if stage == "prepass":     # read key, signs, flags, diagH, offdiagH; write signs, flags, counts, pinfo
        return Nd * (K + Sg + 4 + 16) + Nd * (Sg + 4 + 4 * S + S)
    if stage == "scan":        # read counts, write offsets
        return Nd * S * (4 + 8)
    if stage == "spawn":       # parents once (key, pinfo, offset); one alias entry per double attempt;
        #                        queue (16 B) and double spawns written
        return Nd * (K + S + 8 * S) + (Na - Nq) * 20 + Nq * 16 + (Ns - Nsq) * (K + V + 4)
    if stage == "spawn_singles":  # queue entry, parent key/pinfo/offset, single spawns written
        return Nq * (16 + K + 1 + 8) + Nsq * (K + V + 4)
    if stage == "sort":        # per pass: histogram read of keys, scatter read+write of (key, index)
        kb, passes = sort_key(nbasis, nel)
        return passes * Ns * (kb + 2 * (kb + 4))
    if stage == "merge":       # fused into the annihilate stage
        return 0.0
    if stage == "annihilate":  # merge and annihilation in one pass: sorted keys + index, gathered value + info
        #                        per spawn; per unique det one hash probe (K + 8) and its state row
        #                        (S signs + flag word) read and written; new list written
        return Ns * (K + 4 + V + 4) + Nu * (K + 8) + Nu * 2 * (Sg + 8) + Nn * (K + Sg + 4)
    if stage == "insert":      # new list, free-queue entry, hash entry, walker row (key, signs, flags, 3x8)
        return Nn * (K + Sg + 4) + Nn * 8 + Nn * (K + 8) + Nn * (K + Sg + 4 + 24)
    if stage == "stats":       # flags, signs, key per slot
        return Nd * (4 + Sg + K)
    if stage == "core":        # core H.v: each stored element (value + int32 column) once, the core
        #                        amplitudes read and the projected ones written
        return c.get("Nc", 0.0) * (V + 4) + 2 * Sg * c.get("Ncore", 0.0)
    return 0.0


def iteration_bytes(c, W=1, S=1, nbasis=40, psingles=0.1, nel=None, val_bytes=8):
    """Compulsory HBM bytes of one whole iteration: the stages' figures of
    algorithmic_bytes summed, except that the sort counts one read and one
    write of (key, index) per spawn, the algorithmic minimum, not one per
    radix pass."""
    tot = sum(algorithmic_bytes(st, c, W, S, nbasis, psingles, nel, val_bytes) for st in STAGES if st != "sort")
    return tot + 2 * c["Ns"] * (sort_key(nbasis, nel)[0] + 4)


def _kernel_short(name):
    """'void ngpu::k_spawn<1, 1>(...)' -> 'k_spawn<1, 1>'."""
    n = name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").strip()
    return n.split("::")[-1]


def summarize_trace(path, last, per_iter_kernel="k_end_iter", ends_per_iter=1):
    """Steady-state table from a rocprofv3 kernel-trace CSV: per kernel, the
    dispatches of the last `last` iterations (an iteration ends with
    ends_per_iter dispatches of per_iter_kernel: one for a single engine, one
    per launch group for a batch), average and per-iteration duration."""
    import csv
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    ends = [i for i, r in enumerate(rows) if _kernel_short(r["Kernel_Name"]) == per_iter_kernel]
    ends = ends[ends_per_iter - 1::ends_per_iter] if len(ends) % ends_per_iter == 0 else \
        ends[len(ends) % ends_per_iter + ends_per_iter - 1::ends_per_iter]
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
        row = {"instance": k, "fetch_kib": f[k], "write_kib": w[k],
               "hbm_bytes_per_launch": sf * f[k] * 1024.0 + sw * w[k] * 1024.0}
        # instances of one kernel (other generators, word counts): the busiest
        if name not in table["kernels"] or row["hbm_bytes_per_launch"] > table["kernels"][name]["hbm_bytes_per_launch"]:
            table["kernels"][name] = row
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


def cpu_threads():
    """Host threads of the CPU baseline: every core the process may use, at
    most 16 (a one-GPU box's share of its host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, n))


def cpu_baseline(spec, tau, dets, signs, flags, shift, it0, seconds, core=None, threads=None):
    """The oracle (the CPU restatement of PerformFCIMCycPar) on a walker set
    the GPU reached, run as P hash-sharded states on P threads (one per host
    core, as the reference runs one MPI rank per core), for as many
    iterations as fit in the time budget.  Returns the cpu_baseline object."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from oracle_ctypes import ShardedOracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from common import neci_to_internal_flags
    P = threads or cpu_threads()
    orc = ShardedOracle(spec, tau, P)
    t_set = time.perf_counter()
    orc.add_walkers_bulk(dets, signs, neci_to_internal_flags(flags, spec.S))
    if core is not None:
        orc.set_core(core)
    t_set = time.perf_counter() - t_set
    shift = np.array(shift, dtype=np.float64).copy()
    walker_iters = 0.0
    n = 0
    t0 = time.perf_counter()
    while True:
        st = orc.iterate(it0 + n, shift)
        if st["rc"]:
            raise RuntimeError(f"oracle iteration failed: {st['rc']}")
        walker_iters += float(np.sum(st["tot_parts"][:spec.S]))
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    orc.close()
    return dict(value=walker_iters / dt, unit="walker-iterations/s", cores=P, kind="port",
                sample=f"{n} iterations of the oracle, hash-sharded over {P} threads, from the GPU's walker set "
                       f"({len(dets)} determinants, {walker_iters / n:.3g} walkers per iteration"
                       + (f", {len(core)}-determinant core" if core is not None else "")
                       + f"), {dt:.1f} s (set-up {t_set:.1f} s not counted)")


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
    mol_kw = dict(trunc_initiator=True, init_walk_no=3.0, particles=0, psingles=0.1, pparallel=0.5,
                  memory_fac_spawn=4.0, memory_fac_part=3.0, memory_fac_hash=2.0)
    return mol, mol_kw, hub_specs, mol_walkers, ncore


# tau of the 14e/30o molecular problems.  The 10e/20o workload's 0.005 is past
# this Hamiltonian's stable time step: from a population settled at tau 0.001
# (1e5 walkers, its shift held), one iteration grows the population by 0.1 %
# at tau 0.001, 3.6 % at 0.002, 13 % at 0.003, 45 % at 0.004 and 80 % at
# 0.005 (the oracle; the engine at 4.6e6 walkers: x1.15, x1.43, x1.68, x1.84,
# x1.93 in the first five iterations at 0.005, then the spawn list overflows).
TAU_MOL = 0.001


def run_batched(args):
    """Config 3: 32 problems iterated two ways from the same populations.
    Plain FP64: each problem its own engine, iterated alone with one launch
    per kernel, one problem after the other, core Hamiltonian and spawned
    amplitudes in FP64.  Batched FP32/FP64: all engines as the branches of
    one captured graph per iteration, core Hamiltonian and spawned
    amplitudes stored in FP32, every sum, the energies and the shift in
    FP64.  The molecular replicas start from one population grown to the
    target (a common starting point, as replicas restarted from one
    POPSFILE) and then follow their own random streams (distinct seeds)."""
    import numpy as np
    from neci_stable_amd import Spec
    from neci_stable_amd.engine import Batch, Engine
    from neci_stable_amd.fcimc import BatchedFciMCPar, FciMCPar

    mol, mol_kw, hub_specs, mol_walkers, ncore = batched_workload(args.mol_replicas, args.hub_systems,
                                                                  mol_walkers=args.mol_walkers, ncore=args.ncore)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    t_setup = time.perf_counter()
    # one molecular population grown to the target under TARGET-SHIFTDAMP
    g_spec = Spec(mol, init_walkers=mol_walkers, seed=6999, **mol_kw)
    g = Engine(g_spec)
    g.start_single_part(1000.0)
    gd = FciMCPar(g, tau=TAU_MOL, target_walkers=mol_walkers, steps_shift=10, shift_damp=0.5,
                  target_shiftdamp=True)
    settled = 0
    while settled < 3 and gd.iter < args.max_growth_iters:
        gd.run(10)
        av = float(np.sum(gd.rows[-1]["tot_parts"]))
        settled = settled + 1 if abs(av / mol_walkers - 1.0) < 0.1 else 0
    dets, sg, fl, _, _ = g.download()
    core = None
    if ncore:
        o = np.argsort(-np.abs(sg).sum(axis=1), kind="stable")[:ncore]
        core = np.ascontiguousarray(dets[np.sort(o)])
    log(f"# molecular population: {gd.iter} iterations, {np.abs(sg).sum():.4g} walkers on {len(dets)} "
        f"determinants, shift {gd.shift[0]:.4f}")
    rows = g.rows(dets, sg, fl)
    g.close()
    engines = []
    for r in range(args.mol_replicas):
        e = Engine(Spec(mol, init_walkers=mol_walkers, seed=7000 + r, **mol_kw))
        e.upload(rows)
        if core is not None:
            e.set_core(core, fp32=False)
        d = FciMCPar(e, tau=TAU_MOL, target_walkers=mol_walkers, steps_shift=10, shift_damp=0.5,
                     target_shiftdamp=True, diag_shift=float(gd.shift[0]))
        d.iter = gd.iter
        d.old_av_walkers = gd.old_av_walkers
        engines.append((e, d))
    del rows
    for spec, tau, h in hub_specs:
        e = Engine(spec)
        e.start_single_part(30.0)
        d = FciMCPar(e, tau=tau, target_walkers=h["target"], steps_shift=10, shift_damp=0.5)
        while d.single_part.any() and d.iter < args.max_growth_iters:
            d.run(10)
        engines.append((e, d))
    drivers = [d for _, d in engines]
    for d in drivers:
        d.run(args.warmup)
    mol_engines = [e for e, _ in engines[:args.mol_replicas]]
    tuned = None
    if not args.no_autotune and mol_engines:
        # the plain form's launch shape: tuned on one molecular replica and
        # given to all of them (the same system at the same size)
        from neci_stable_amd.autotune import autotune
        tp_ = autotune(drivers[0], steps=10, warm=10, repeats=1)  # whole update cycles: the drivers stay aligned
        for e in mol_engines[1:]:
            e.set_launch_shape(**tp_["shape"])
        tuned = {"plain": tp_["shape"]}
        log(f"# autotuned launch shape, plain molecular engines: {tp_['shape']}")
    t_setup = time.perf_counter() - t_setup
    log(f"# {len(engines)} engines set up in {t_setup:.1f} s")

    def totals():
        return sum(d.totals["walker_iters"] for d in drivers)

    K = args.steps
    # plain FP64: each engine alone, one after the other
    per_plain = []
    w0 = totals()
    tp = 0.0
    for e, d in engines:
        a = d.totals["walker_iters"]
        t0 = time.perf_counter()
        d.run(K)
        dt = time.perf_counter() - t0
        tp += dt
        per_plain.append((dt * 1e3 / K, (d.totals["walker_iters"] - a) / K))
    p_rate = (totals() - w0) / tp
    log(f"# plain FP64: {tp * 1e3 / K:.2f} ms per iteration of all problems")
    # batched FP32/FP64
    nnz0 = 0
    for i, (e, _) in enumerate(engines):
        if i < args.mol_replicas and core is not None:
            nnz = e.set_core(core, fp32=True)
            nnz0 = nnz0 or nnz
        e.set_spawn_precision(True)
    batch = Batch([e for e, _ in engines])
    runner = BatchedFciMCPar(batch, drivers)
    runner.run(max(args.warmup, 10))  # capture and first replays
    if tuned is not None:
        # the batched form's launch shape: the molecular group's engines tuned
        # together on the batch's own iterations (regrouped and recaptured
        # per candidate)
        from neci_stable_amd.autotune import tune_engines
        tb_ = tune_engines(mol_engines, runner.run, steps=10, warm=3, repeats=1)
        tuned["batched"] = tb_["shape"]
        log(f"# autotuned launch shape, batched molecular group: {tb_['shape']}")
        runner.run(max(args.warmup, 10))
    gids, ngroups = batch.groups()
    gm = int(gids[0])                  # the molecular replicas' launch group
    mol = [i for i in range(len(engines)) if gids[i] == gm]
    # one captured graph per iteration (the batch's default form)
    w0 = totals()
    t0 = time.perf_counter()
    runner.run(K)
    tg = time.perf_counter() - t0
    g_rate = (totals() - w0) / tg
    log(f"# batched FP32/FP64, graph replay: {tg * 1e3 / K:.2f} ms per iteration of all problems, "
        f"{ngroups} launch groups")
    # the roofline kernel: every stage of the molecular group timed over a
    # short pass of direct launches, then the timed region carries events on
    # the slowest single-kernel stage's launches only
    batch.set_graph(False)
    batch.set_profiling(True)
    runner.run(10)
    ms_all, nprof_all = batch.group_stage_times(gm)
    per_stage = {STAGES[i]: ms_all[i] / max(nprof_all, 1) for i in range(len(STAGES))}
    dom = max((k for k in per_stage if k in SINGLE_KERNEL_STAGES), key=per_stage.get)
    batch.set_profiling(1 << STAGES.index(dom))
    c0 = [dict(drivers[i].totals) for i in mol]
    w0 = totals()
    t0 = time.perf_counter()
    runner.run(K)
    tb = time.perf_counter() - t0
    b_rate = (totals() - w0) / tb
    ms, nprof = batch.group_stage_times(gm)
    dom_ms = ms[STAGES.index(dom)] / max(nprof, 1)
    batch.set_profiling(False)
    log(f"# batched FP32/FP64, direct launches with events on {KERNELS[dom]}: {tb * 1e3 / K:.2f} ms per iteration")
    # algorithmic bytes of one launch of the dominant kernel: the sum over
    # the molecular group's engines of each engine's per-iteration figure
    b = 0.0
    e0 = engines[0][0]
    for j, i in enumerate(mol):
        cnt = window_counts(c0[j], drivers[i].totals)
        cnt["Nc"], cnt["Ncore"] = float(nnz0), float(len(core) if core is not None else 0)
        b += algorithmic_bytes(dom, cnt, W=e0.W, S=e0.S, nbasis=e0.nbasis, psingles=e0.psingles, nel=e0.nel,
                               val_bytes=4)
    achieved = b / (dom_ms * 1e-3) / 1e9
    # HBM bytes per launch of that kernel's group instance from the FETCH_SIZE /
    # WRITE_SIZE counter runs of this workload (profiles/), when collected
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_batched_14e30o_hubbard.json")
    if os.path.exists(tpath) and args.mol_replicas == 16 and args.hub_systems == 16:
        with open(tpath) as f:
            traffic = json.load(f)["kernels"].get("km" + KERNELS[dom][1:], {}).get("hbm_bytes_per_launch")
    d0 = engines[0][1]
    batch.close()
    walkers = [float(np.sum(d.last["tot_parts"])) for d in drivers]
    cpu = None
    if not args.no_cpu:
        dets0, sg0, fl0, _, _ = e0.download()
        import dataclasses
        cpu = cpu_baseline(dataclasses.replace(e0.spec, spawn_fp32=False), TAU_MOL, dets0, sg0, fl0, d0.shift,
                           d0.iter + 1, args.cpu_seconds, core=core)
    for e, _ in engines:
        e.close()
    line = {
        "metric": METRIC, "value": b_rate, "unit": "walker-iterations/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": tb * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "precision": "walkers, sums, energies and shift in FP64; core Hamiltonian and spawned amplitudes "
                     "stored in FP32 (the batched form)",
        "data": "synthetic",
        "config": {"workload": f"batched: {args.mol_replicas} replicas of molecular 14e/30o (seed 7, rank 60, "
                               f"initiator 3, PCHB unif-unif + heat-bath singles, psingles 0.1, tau {TAU_MOL}, "
                               f"{ncore}-determinant core, distinct seeds) + {len(hub_specs)} Hubbard chains of "
                               f"12-18 sites (U/t in [2, 8], targets 1e4-2e6, tau 0.01)",
                   "mol_target_walkers": mol_walkers, "walkers_per_problem": walkers,
                   "total_walkers": float(sum(walkers)),
                   "autotuned_launch_shape": tuned},
        "plain": {"value": p_rate, "ms_per_step": tp * 1e3 / K,
                  "form": "each problem its own engine, iterated alone with one launch per kernel, one problem "
                          "after the other, FP64 core Hamiltonian and spawned amplitudes",
                  "ms_per_problem": [round(x[0], 4) for x in per_plain]},
        "batched_over_plain": b_rate / p_rate,
        "graph_replay": {"value": g_rate, "ms_per_step": tg * 1e3 / K,
                         "form": "the same batched chain replayed as one captured graph per iteration, no events"},
        "launch_groups": int(ngroups),
        "roofline": {"bound": "hbm", "kernel": KERNELS[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": b, "launch_ms": dom_ms,
                     "where": f"one launch covers the {len(mol)} molecular replicas; its launches timed with "
                              "events over the timed region"},
        "stage_ms_per_step": per_stage,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def window_counts(t0, t1):
    """Per-iteration counts over the iterations between two snapshots of a
    driver's running totals (every iteration counted, whatever the update
    cycle boundaries)."""
    n = t1["iters"] - t0["iters"]
    if n <= 0:
        raise ValueError("no iterations between the snapshots")
    return dict(Nd=(t1["det_iters"] - t0["det_iters"]) / n, Na=(t1["attempts"] - t0["attempts"]) / n,
                Ns=(t1["spawned"] - t0["spawned"]) / n, Nu=(t1["unique"] - t0["unique"]) / n,
                Nn=(t1["new"] - t0["new"]) / n, walkers=(t1["walker_iters"] - t0["walker_iters"]) / n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=["batched", "mol10e20o"], default="batched",
                    help="batched: the 32-problem batched-vs-plain workload (the metric's configuration); "
                         "mol10e20o: one 10e/20o molecular engine, plain FP64")
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
    ap.add_argument("--per-iter-kernel", default="k_end_iter", help="kernel that ends an iteration (trace summary)")
    ap.add_argument("--ends-per-iter", type=int, default=1,
                    help="dispatches of --per-iter-kernel per iteration (launch groups of a batch)")
    ap.add_argument("--mol-replicas", type=int, default=16)
    ap.add_argument("--ncore", type=int, default=50000, help="core determinants of the molecular problems")
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
        summarize_trace(args.summarize_trace, args.last, args.per_iter_kernel, args.ends_per_iter)
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
    if args.workload == "batched":
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
    tot0 = dict(drv.totals)
    eng.set_profiling(1 << STAGES.index(dom))
    w0 = drv.walker_iters
    t0 = time.perf_counter()
    drv.run(args.steps)
    dt = time.perf_counter() - t0
    walker_iters = drv.walker_iters - w0
    ms, nprof = eng.stage_times()
    eng.set_profiling(False)
    dom_ms = ms[STAGES.index(dom)] / max(nprof, 1)
    K = args.steps
    counts = window_counts(tot0, drv.totals)  # every timed iteration, whatever the cycle boundaries
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
        dets, sg, fl, _, _ = eng.download()
        line["cpu_baseline"] = cpu_baseline(spec, tau, dets, sg, fl, drv.shift, drv.iter + 1, args.cpu_seconds)
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
