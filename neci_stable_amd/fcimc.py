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
"""
import math
import time

import numpy as np

SUM_FIELDS = ["nborn", "ndied", "nannihil", "naborted", "nremoved", "enum_cyc", "hf_cyc",
              "noat_doubs", "spawn_from_sing", "no_init_dets", "no_init_walk", "no_noninit_dets",
              "no_noninit_walk"]


class CycleAccumulator:
    def __init__(self, nrep):
        self.nrep = nrep
        self.reset()

    def reset(self):
        self.n = 0
        self.sums = {f: np.zeros(self.nrep) for f in SUM_FIELDS}
        self.sum_tot_parts = np.zeros(self.nrep)
        self.attempts = 0
        self.last = None

    def add(self, st):
        self.n += 1
        for f in SUM_FIELDS:
            self.sums[f] += st[f]
        self.sum_tot_parts += st["tot_parts"]
        self.attempts += st["n_attempts"]
        self.last = st


class FciMCPar:
    """Runs NMCyc iterations of a backend exposing iterate(it, tau, shift) and
    cycle statistics.  backend: neci_stable_amd.Engine (GPU) or any object with
    the same iterate/stats API."""

    def __init__(self, backend, *, tau, target_walkers, steps_shift=10, shift_damp=0.1,
                 diag_shift=0.0, jump_shift=True, hii=None, log=None):
        self.b = backend
        self.tau = float(tau)
        self.target = float(target_walkers)
        self.steps_shift = int(steps_shift)
        self.shift_damp = float(shift_damp)
        self.S = backend.S
        self.shift = np.full(self.S, float(diag_shift))
        self.single_part = np.ones(self.S, dtype=bool)
        self.vary_shift_iter = np.zeros(self.S, dtype=np.int64)
        self.jump_shift = jump_shift
        self.hii = backend.hii if hii is None else hii
        self.old_av_walkers = None
        self.acc = CycleAccumulator(self.S)
        self.iter = 0
        self.rows = []       # FCIMCStats-like rows per update cycle
        self.shift_hist = []  # (iter, shift, proj_e) for blocking analysis
        self.sum_enum = np.zeros(self.S)
        self.sum_noathf = np.zeros(self.S)
        self.walker_iters = 0.0
        self.log = log

    def _update_shift(self):
        a = self.acc
        av = a.sum_tot_parts / float(a.n)
        growth = None
        if self.old_av_walkers is not None and np.all(self.old_av_walkers != 0):
            growth = av / self.old_av_walkers
        hf = a.sums["hf_cyc"]
        en = a.sums["enum_cyc"]
        proje = np.where(np.abs(hf) > 1e-13, en / np.where(hf == 0, 1, hf), np.nan)
        defer = np.zeros(self.S, dtype=bool)
        tot_parts = a.last["tot_parts"]
        for r in range(self.S):
            if self.single_part[r]:
                if tot_parts[r] > self.target:
                    self.single_part[r] = False
                    self.vary_shift_iter[r] = self.iter
                    if self.jump_shift and np.isfinite(proje[r]):
                        self.shift[r] = proje[r]
                        defer[r] = True
            if not self.single_part[r] and not defer[r] and growth is not None:
                if growth[r] > 0:
                    self.shift[r] -= math.log(growth[r]) * self.shift_damp / (self.tau * self.steps_shift)
        self.old_av_walkers = av
        if not np.all(self.single_part):
            self.sum_enum += en
            self.sum_noathf += hf
        row = dict(iter=self.iter, shift=self.shift.copy(), tot_parts=tot_parts.copy(),
                   tot_walkers=a.last["tot_walkers"], proj_e=proje, hf=hf / a.n,
                   nborn=a.sums["nborn"], ndied=a.sums["ndied"], nannihil=a.sums["nannihil"],
                   naborted=a.sums["naborted"], single_part=self.single_part.copy())
        self.rows.append(row)
        if not np.any(self.single_part):
            self.shift_hist.append((self.iter, self.shift.copy(), proje.copy()))
        if self.log:
            self.log(row)
        a.reset()

    def run(self, n_iter):
        t0 = time.perf_counter()
        for _ in range(n_iter):
            self.iter += 1
            st = self.b.iterate(self.iter, self.tau, self.shift)
            self.acc.add(st)
            self.walker_iters += float(np.sum(st["tot_parts"]))
            if self.iter % self.steps_shift == 0:
                self._update_shift()
        return time.perf_counter() - t0

    def projected_energy(self):
        """ProjectionE = SumENum / SumNoatHF (+ Hii) over the varying-shift phase."""
        with np.errstate(invalid="ignore", divide="ignore"):
            return self.sum_enum / self.sum_noathf + self.hii


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
