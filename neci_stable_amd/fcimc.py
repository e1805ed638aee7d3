"""Host driver: FciMCPar's main loop and update cycle around the GPU iteration.

Per iteration the engine runs PerformFCIMCycPar on the device; every
StepsSft iterations the host performs the update cycle of
calculate_new_shift_wrapper / update_shift (fcimc_iter_utilities.F90:883-1260)
in FP64 from the cycle-accumulated device statistics:

  growth   = (sum_cycle TotParts / StepsSft) / previous cycle average  (AVGROWTHRATE default)
  DiagSft -= ln(growth) * SftDamp / (tau * StepsSft)                      (:1157)
  E_proj   = ENumCyc / HFCyc (+ Hii for the total energy)                 (:1205-1215)

During the single-particle phase the shift is held until TotParts exceeds
TOTALWALKERS, then jumps to the cycle projected energy (tJumpShift, :1063-1083).
With TARGET-SHIFTDAMP (tTargetShiftdamp, Calc.F90:1532-1550) there is no
single-particle phase and the update gains a term that pulls the population
to TOTALWALKERS (:1147-1154):

  DiagSft -= (ln(growth) * SftDamp + ln(TotParts / TOTALWALKERS) * SftDamp2) / (tau * StepsSft)

with SftDamp2 = SftDamp^2 / 4 unless given.
"""
import math
import time

import numpy as np

# statistics that the cycle keeps as a maximum, or as the last iteration's value
_CYCLE_MAX = ("max_cyc_spawn", "bloom_size")
_CYCLE_LAST = ("tot_walkers", "inst_no_at_hf", "rc")


def merge_cycle_sums(a, b):
    """Cycle sums of two consecutive runs of iterations (a, then b) as one:
    sums add, maxima take the larger, the last iteration's values come from b
    (the accumulation rules of the engine's cycle block)."""
    out = {}
    for k, v in b.items():
        if k not in a:
            out[k] = v
        elif k in _CYCLE_MAX:
            out[k] = np.maximum(a[k], v) if isinstance(v, np.ndarray) else max(a[k], v)
        elif k in _CYCLE_LAST:
            out[k] = v
        elif k == "err":
            out[k] = int(a[k]) | int(v)
        else:
            out[k] = a[k] + v
    return out


class OracleCycleAdapter:
    """Gives a synchronous backend (iterate(it, shift) -> stats) the cycle API
    of the engine: asynchronous iterate + cycle_stats() sums."""

    def __init__(self, backend, nrep):
        self.b = backend
        self.S = nrep
        self.hii = backend.hii
        self._reset()

    def _reset(self):
        self.sum = None
        self.n = 0
        self.last = None

    def iterate(self, it, tau, shift, wait=False):
        st = self.b.iterate(it, shift)
        if st.get("rc", 0) or st.get("err", 0):
            raise RuntimeError(f"iteration {it} failed with status {st.get('rc')}/{st.get('err')}")
        self.last = st
        self.n += 1
        if self.sum is None:
            self.sum = {k: (np.array(v, dtype=float).copy() if isinstance(v, np.ndarray) else v)
                        for k, v in st.items()}
        else:
            self.sum = merge_cycle_sums(self.sum, st)
        return st if wait else None

    def cycle_stats(self):
        out = (self.sum, self.last, self.n)
        self._reset()
        return out


class FciMCPar:
    """Runs iterations of a backend with the engine's cycle API: iterate(it,
    tau, shift, wait=False) and cycle_stats() -> (sums, last, n_iters)."""

    def __init__(self, backend, *, tau, target_walkers, steps_shift=10, shift_damp=0.1,
                 diag_shift=0.0, jump_shift=True, hii=None, log=None, target_shiftdamp=False,
                 shift_damp2=None):
        self.b = backend
        self.tau = float(tau)
        self.target = float(target_walkers)
        self.steps_shift = int(steps_shift)
        self.shift_damp = float(shift_damp)
        self.S = backend.S
        self.shift = np.full(self.S, float(diag_shift))
        self.target_shiftdamp = bool(target_shiftdamp)
        if self.target_shiftdamp:
            self.shift_damp2 = self.shift_damp ** 2 / 4.0 if shift_damp2 is None else float(shift_damp2)
            if not (0.0 < self.shift_damp < 1.0 and 0.0 < self.shift_damp2 < self.shift_damp):
                raise ValueError("shift damping factors must lie in (0, 1) with SftDamp2 < SftDamp")
        self.single_part = np.full(self.S, not self.target_shiftdamp)
        self.vary_shift_iter = np.zeros(self.S, dtype=np.int64)
        self.jump_shift = jump_shift
        self.hii = backend.hii if hii is None else hii
        self.old_av_walkers = None
        self.iter = 0
        self.rows = []        # FCIMCStats-like rows per update cycle
        self.shift_hist = []  # (iter, shift, proj_e) once every replica varies its shift
        self.sum_enum = np.zeros(self.S)
        self.sum_noathf = np.zeros(self.S)
        self.walker_iters = 0.0
        # running totals over every iteration read so far (per-step figures
        # of any window of iterations are differences of these)
        self.totals = dict(iters=0, walker_iters=0.0, attempts=0, spawned=0, unique=0, new=0, det_iters=0.0)
        self._pending = None  # sums of the current update cycle's iterations read so far
        self.last = None
        self.log = log

    def _update_shift(self, sm, last, n):
        av = sm["tot_parts"] / float(n)
        growth = None
        if self.old_av_walkers is not None and np.all(self.old_av_walkers != 0):
            growth = av / self.old_av_walkers
        hf = sm["hf_cyc"]
        en = sm["enum_cyc"]
        with np.errstate(invalid="ignore", divide="ignore"):
            proje = np.where(np.abs(hf) > 1e-13, en / np.where(hf == 0, 1, hf), np.nan)
        defer = np.zeros(self.S, dtype=bool)
        tot_parts = last["tot_parts"]
        for r in range(self.S):
            if self.single_part[r] and tot_parts[r] > self.target:
                self.single_part[r] = False
                self.vary_shift_iter[r] = self.iter
                if self.jump_shift and np.isfinite(proje[r]):
                    self.shift[r] = proje[r]
                    defer[r] = True
            if not self.single_part[r] and not defer[r] and growth is not None and growth[r] > 0:
                d = math.log(growth[r]) * self.shift_damp
                if self.target_shiftdamp and av[r] > 0:
                    d += math.log(av[r] / self.target) * self.shift_damp2
                self.shift[r] -= d / (self.tau * self.steps_shift)
        self.old_av_walkers = av
        if not np.all(self.single_part):
            self.sum_enum += en
            self.sum_noathf += hf
        row = dict(iter=self.iter, shift=self.shift.copy(), tot_parts=tot_parts.copy(),
                   tot_walkers=last["tot_walkers"], proj_e=proje, hf=hf / n, enum=en / n,
                   nborn=sm["nborn"], ndied=sm["ndied"], nannihil=sm["nannihil"],
                   naborted=sm["naborted"], attempts=sm["n_attempts"], spawned=sm["n_spawned"],
                   unique=sm["n_unique"], new=sm["n_new"],
                   single_part=self.single_part.copy())
        self.rows.append(row)
        if not np.any(self.single_part):
            self.shift_hist.append((self.iter, self.shift.copy(), proje.copy()))
        if self.log:
            self.log(row)

    def _chunk_done(self, sm, last, n):
        """Statistics of the iterations run since the last read.  The update
        cycle (calculate_new_shift_wrapper) averages over all StepsSft
        iterations of the cycle (fcimc_iter_utilities.F90:929-933), so a
        chunk that ends inside a cycle is kept and merged with the rest of
        the cycle; the running totals count every iteration as it is read."""
        self.walker_iters += float(np.sum(sm["tot_parts"]))
        t = self.totals
        t["iters"] += n
        t["walker_iters"] += float(np.sum(sm["tot_parts"]))
        t["attempts"] += int(sm["n_attempts"])
        t["spawned"] += int(sm["n_spawned"])
        t["unique"] += int(sm["n_unique"])
        t["new"] += int(sm["n_new"])
        t["det_iters"] += float(last["tot_walkers"]) * n
        self.last = last
        if self._pending is not None:
            sm, n = merge_cycle_sums(self._pending[0], sm), self._pending[1] + n
            self._pending = None
        if self.iter % self.steps_shift == 0:
            self._update_shift(sm, last, n)
        else:
            self._pending = (sm, n)

    def run(self, n_iter):
        """n_iter iterations, update cycle every StepsSft; returns wall seconds."""
        t0 = time.perf_counter()
        done = 0
        while done < n_iter:
            k = min(self.steps_shift - self.iter % self.steps_shift, n_iter - done)
            for _ in range(k):
                self.iter += 1
                self.b.iterate(self.iter, self.tau, self.shift, wait=False)
            done += k
            self._chunk_done(*self.b.cycle_stats())
        return time.perf_counter() - t0

    def projected_energy(self):
        """ProjectionE = SumENum / SumNoatHF (+ Hii) over the varying-shift phase."""
        with np.errstate(invalid="ignore", divide="ignore"):
            return self.sum_enum / self.sum_noathf + self.hii


class BatchedFciMCPar:
    """FciMCPar for several engines at once: one ngpu_batch_iterate per
    iteration runs every engine's iteration; each driver keeps its own shift
    and update cycle (all drivers share StepsSft; their cycles need not be
    aligned)."""

    def __init__(self, batch, drivers):
        self.batch = batch
        self.drivers = list(drivers)
        steps = {d.steps_shift for d in self.drivers}
        if len(steps) != 1:
            raise ValueError("batched drivers must share StepsSft")
        self.steps_shift = steps.pop()

    def run(self, n_iter):
        t0 = time.perf_counter()
        done = 0
        while done < n_iter:
            # no driver passes the end of its update cycle without a read
            # (drivers may be at different points of their cycles)
            k = min(min(self.steps_shift - d.iter % self.steps_shift for d in self.drivers), n_iter - done)
            for _ in range(k):
                for d in self.drivers:
                    d.iter += 1
                self.batch.iterate([d.iter for d in self.drivers], [d.tau for d in self.drivers],
                                   [d.shift for d in self.drivers])
            done += k
            for d in self.drivers:
                d._chunk_done(*d.b.cycle_stats())
        return time.perf_counter() - t0

    @property
    def walker_iters(self):
        return sum(d.walker_iters for d in self.drivers)


def reblock(x):
    """Flyvbjerg-Petersen reblocking (the method of ErrorAnalysis.F90:374-1145):
    returns (mean, error) taking the plateau of the blocked standard error."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    if n < 4:
        return float(np.mean(x)), float("nan")
    mean = float(np.mean(x))
    errs = []
    y = x.copy()
    while y.size >= 4:
        m = y.size
        err = np.sqrt(np.var(y, ddof=1) / m)
        errs.append((m, err))
        if m % 2:
            y = y[:-1]
        y = 0.5 * (y[0::2] + y[1::2])
    # plateau: first blocking level where the error stops growing beyond its own uncertainty
    best = errs[-1][1]
    for m, err in errs:
        e_err = err / np.sqrt(2.0 * (m - 1))
        nxt = [e for mm, e in errs if mm < m]
        if all(e <= err + 2 * e_err for e in nxt):
            best = err
            break
    return mean, float(best)


def _plateau(errs):
    """Index of the first blocking level whose error no later level exceeds
    by more than twice its own uncertainty (the plateau of Flyvbjerg-Petersen
    blocking); the last level when none qualifies."""
    for i, (m, err) in enumerate(errs):
        e_err = err / np.sqrt(2.0 * (m - 1))
        if all(e <= err + 2 * e_err for _, e in errs[i + 1:]):
            return i
    return len(errs) - 1


def _blocks(x):
    """Blocking levels of x: each level halves the previous one by averaging
    neighbours (reblock_data, ErrorAnalysis.F90:944-984, block length 2)."""
    out = []
    y = np.asarray(x, dtype=np.float64)
    while y.size >= 4:
        out.append(y)
        if y.size % 2:
            y = y[:-1]
        y = 0.5 * (y[0::2] + y[1::2])
    return out


def proje_error(num, den):
    """Projected energy and its error from the per-cycle numerator and
    denominator series (error_analysis, ErrorAnalysis.F90:250-300): both series
    are reblocked to the longer of their two plateau block lengths, the
    means and errors taken there (analyze_data, :912-942), and the ratio's
    error combines the relative errors with the covariance correction:
    E = <N>/<D>, err = |E| sqrt((sD/<D>)^2 + (sN/<N>)^2 - 2 cov(N, D)/(n <D> <N>))."""
    bn, bd = _blocks(num), _blocks(den)
    if not bn:
        return float(np.mean(num) / np.mean(den)), float("nan")

    def errs(levels):
        return [(y.size, float(np.sqrt(max(np.mean(y * y) - np.mean(y) ** 2, 0.0) / (y.size - 1))))
                for y in levels]
    lvl = max(_plateau(errs(bn)), _plateau(errs(bd)))
    yn, yd = bn[lvl], bd[lvl]
    n = yn.size
    mn, md = float(np.mean(yn)), float(np.mean(yd))
    sn = float(np.sqrt(max(np.mean(yn * yn) - mn * mn, 0.0) / (n - 1)))
    sd = float(np.sqrt(max(np.mean(yd * yd) - md * md, 0.0) / (n - 1)))
    cov = float(np.sum((yn - mn) * (yd - md)) / (n - 1))
    e = mn / md
    var = (sd / md) ** 2 + (sn / mn) ** 2 - 2.0 * cov / (n * md * mn)
    return e, abs(e) * float(np.sqrt(max(var, 0.0)))

