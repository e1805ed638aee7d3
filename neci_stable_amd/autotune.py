"""Launch-shape autotuner.

The engine's kernels split their work over a number of blocks fixed at launch
and the radix sort ranks tiles of a fixed number of elements per thread
(ngpu_set_launch_shape); the heat-bath singles run 4, 8 or 16 lanes per
attempt.  The best values depend on the walker count and the system, so they
are measured on the running calculation: one knob at a time (radix tile, the
chunked kernels' blocks, the grid-stride kernels' blocks, the lanes per
single attempt, then the merge-annihilation kernel's blocks), each candidate
timed over a few iterations of the driver (or of a batch, for the engines of
one launch group), the fastest kept.  The walker dynamics do not depend on
the shape, so tuning while the calculation runs changes nothing but its
speed.
"""
import time


def _time_iterations(run, steps):
    t0 = time.perf_counter()
    run(steps)
    return (time.perf_counter() - t0) / steps * 1e3


def tune_engines(engines, run, steps=20, warm=5, repeats=2, log=None):
    """Tune the launch shape shared by engines (Engine objects of one shape,
    e.g. the replicas of one system), timing run(n), which iterates them n
    times (a driver's run, or a batch's: the engines of a batch are regrouped
    and their graph recaptured when their shape changes).  One knob at a
    time, the fastest candidate kept.  Returns {"shape": {...}, "trials": [...]}."""
    base = engines[0].launch_shape()
    cu = base["n_cu"]
    knobs = [
        ("radix_items", [2, 4, 8]),
        ("chunk_blocks", [2 * cu, 4 * cu, 8 * cu]),
        ("grid_blocks", [3 * cu, 4 * cu, 8 * cu]),
        ("single_lanes", [4, 8, 16]),
        ("merge_blocks", [2 * cu, 3 * cu, 4 * cu]),
    ]
    best = {k: base[k] for k in ("chunk_blocks", "grid_blocks", "radix_items", "single_lanes", "merge_blocks")}
    trials = []
    for name, values in knobs:
        scores = {}
        for v in values:
            shape = dict(best)
            shape[name] = v
            for e in engines:
                e.set_launch_shape(**shape)
            ms = []
            for _ in range(repeats):
                run(warm)
                ms.append(_time_iterations(run, steps))
            scores[v] = min(ms)
            trials.append(dict(shape, ms_per_iteration=scores[v]))
            if log:
                log(f"autotune {shape}: {scores[v]:.4f} ms/iteration")
        best[name] = min(scores, key=scores.get)
    for e in engines:
        e.set_launch_shape(**best)
    return {"shape": best, "trials": trials}


def autotune(driver, steps=20, warm=5, repeats=2, log=None):
    """Tune the launch shape of driver's engine (driver: FciMCPar over an
    Engine).  Runs (warm + steps) * repeats iterations per candidate, about
    25 candidates' worth in all.  Returns {"shape": {...}, "trials": [...]}."""
    return tune_engines([driver.b], driver.run, steps, warm, repeats, log)
