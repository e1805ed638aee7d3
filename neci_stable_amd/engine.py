"""Python view of one GPU engine handle (a thin layer over the C ABI)."""
import ctypes

import numpy as np

from . import _lib
from ._lib import IterInT, IterOutT, SysT, check, ptr
from .config import Spec
from .systems import encode_det


class Engine:
    """One rank's walkers resident on the GPU, set up from a Spec (or from a
    system plus Spec keyword arguments: tau is per iteration, the rest mirror
    TOTALWALKERS, MEMORYFAC*, TRUNCINITIATOR/ADDTOINITIATOR, ALLREALCOEFF,
    REALSPAWNCUTOFF, SETOCCUPIEDTHRESH, SEED and the PCHB options)."""

    def __init__(self, spec_or_system, **kw):
        spec = spec_or_system if isinstance(spec_or_system, Spec) else Spec(spec_or_system, **kw)
        self.spec = spec
        self.lib = _lib.load_library()
        self.system = spec.system
        self.nel, self.nbasis, self.W, self.S = spec.nel, spec.nbasis, spec.W, spec.S
        self.seed = int(spec.seed)
        self.max_walkers, self.max_spawned, self.n_hash_buckets = (
            int(spec.max_walkers), int(spec.max_spawned), int(spec.n_hash_buckets))
        self.model = spec.model
        self.params = spec.params
        self.psingles, self.particles, self.pparallel = spec.psingles, spec.particles, spec.pparallel
        sysv = SysT(nel=self.nel, nbasis=self.nbasis, n_replicas=self.S, model=self.model,
                    t_trunc_initiator=int(spec.trunc_initiator), t_all_real_coeff=int(spec.all_real_coeff),
                    t_real_spawn_cutoff=int(spec.t_real_spawn_cutoff), flag_layout=spec.flag_layout,
                    initiator_walk_no=spec.init_walk_no, real_spawn_cutoff=spec.real_spawn_cutoff,
                    occupied_thresh=spec.occupied_thresh, av_mc_excits=spec.av_mc_excits,
                    max_walkers=self.max_walkers, max_spawned=self.max_spawned,
                    n_hash_buckets=self.n_hash_buckets, seed=self.seed & 0xFFFFFFFF)
        h = ctypes.c_void_p()
        check(self.lib.ngpu_create(ctypes.byref(h), ctypes.byref(sysv)), "ngpu_create")
        self.h = h
        system = spec.system
        if self.model == _lib.MODEL_HUBBARD:
            nbr = np.ascontiguousarray(system.nbr, dtype=np.int32)
            check(self.lib.ngpu_set_lattice(h, ptr(nbr), system.nsites, nbr.shape[1], system.bhub,
                                            system.uhub), "ngpu_set_lattice")
        else:
            tm = np.ascontiguousarray(system.tmat_colmajor())
            um = np.ascontiguousarray(system.umat, dtype=np.float64)
            check(self.lib.ngpu_set_integrals(h, ptr(tm), ptr(um), um.size, system.ecore),
                  "ngpu_set_integrals")
            if spec.excitgen == "uniform":
                check(self.lib.ngpu_set_excitgen_uniform(h, spec.psingles), "ngpu_set_excitgen_uniform")
            else:
                check(self.lib.ngpu_set_excitgen_pchb(h, spec.singles, spec.particles, spec.holes, spec.psingles,
                                                      spec.pparallel), "ngpu_set_excitgen_pchb")
        if spec.spawn_fp32:
            check(self.lib.ngpu_set_spawn_precision(h, 1), "ngpu_set_spawn_precision")
        self.random_hash = spec.random_hash
        check(self.lib.ngpu_set_hash(h, ptr(self.random_hash)), "ngpu_set_hash")
        self.ref_orbs = spec.ref_orbs
        self.ref = encode_det(self.ref_orbs, self.nbasis)
        # Hii = H of the reference (the reference must be set once to evaluate it)
        check(self.lib.ngpu_set_reference(h, ptr(self.ref), 0.0), "ngpu_set_reference")
        self.hii = float(self.diag(self.ref[None, :])[0][0])
        check(self.lib.ngpu_set_reference(h, ptr(self.ref), self.hii), "ngpu_set_reference")
        self._out = IterOutT()

    # -- lifetime
    def close(self):
        if getattr(self, "batch", None) is not None:
            self.batch.close()
        if getattr(self, "h", None):
            self.lib.ngpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- walkers
    def rows(self, dets, signs, flags=None):
        """CurrentDets rows from (n, W) det words and (n, S) signs."""
        dets = np.asarray(dets, dtype=np.int64).reshape(-1, self.W)
        signs = np.asarray(signs, dtype=np.float64).reshape(-1, self.S)
        n = dets.shape[0]
        rows = np.zeros((n, self.W + self.S + 1), dtype=np.int64)
        rows[:, :self.W] = dets
        rows[:, self.W:self.W + self.S] = signs.view(np.int64)
        if flags is not None:
            rows[:, -1] = np.asarray(flags, dtype=np.int64)
        return rows

    def upload(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        check(self.lib.ngpu_upload_walkers(self.h, ptr(rows), rows.shape[0]), "ngpu_upload_walkers")

    def start_single_part(self, n_initial=1.0, initiator_flag=True):
        """Initial walkers on the reference (TStartSinglePart / InitialPart)."""
        sg = np.full((1, self.S), float(n_initial))
        fl = 0
        if initiator_flag:
            for r in range(self.S):
                fl |= 1 << [13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 57, 58][r]
        self.upload(self.rows(self.ref[None, :], sg, [fl]))

    def download(self, cap=None):
        n = ctypes.c_int64()
        if cap is None:
            check(self.lib.ngpu_download_walkers(self.h, None, None, None, 0, ctypes.byref(n)),
                  "ngpu_download_walkers")
            cap = n.value
        rows = np.zeros((max(cap, 1), self.W + self.S + 1), dtype=np.int64)
        dh = np.zeros(max(cap, 1))
        oh = np.zeros(max(cap, 1))
        check(self.lib.ngpu_download_walkers(self.h, ptr(rows), ptr(dh), ptr(oh), cap, ctypes.byref(n)),
              "ngpu_download_walkers")
        m = n.value
        dets = rows[:m, :self.W].copy()
        signs = rows[:m, self.W:self.W + self.S].copy().view(np.float64)
        flags = rows[:m, -1].copy()
        return dets, signs, flags, dh[:m].copy(), oh[:m].copy()

    def download_dev(self, rows_ptr, diag_ptr, offdiag_ptr, cap, n_ptr):
        """ngpu_download_walkers_dev: rows compacted into device memory at the
        given addresses (int or None), asynchronous on the engine's stream."""
        check(self.lib.ngpu_download_walkers_dev(self.h, rows_ptr, diag_ptr, offdiag_ptr, int(cap), n_ptr),
              "ngpu_download_walkers_dev")

    def sync(self):
        """Wait for the engine's queued work (ngpu_get_error waits on its stream)."""
        st = ctypes.c_int32()
        check(self.lib.ngpu_get_error(self.h, ctypes.byref(st), None), "ngpu_get_error")
        return st.value

    # -- iteration
    def iterate(self, it, tau, shift, wait=True):
        inp = IterInT()
        inp.tau = float(tau)
        sh = np.broadcast_to(np.asarray(shift, dtype=np.float64), (self.S,))
        for r in range(self.S):
            inp.diag_sft[r] = float(sh[r])
        inp.iter = int(it) & 0xFFFFFFFF
        if wait:
            check(self.lib.ngpu_iterate(self.h, ctypes.byref(inp), ctypes.byref(self._out)), "ngpu_iterate")
            return self._out.to_dict(self.S)
        check(self.lib.ngpu_iterate(self.h, ctypes.byref(inp), None), "ngpu_iterate")
        return None

    def stats(self):
        check(self.lib.ngpu_get_stats(self.h, ctypes.byref(self._out)), "ngpu_get_stats")
        return self._out.to_dict(self.S)

    def cycle_stats(self):
        """(sums over the cycle, last iteration, number of iterations); starts a new cycle."""
        sm = IterOutT()
        n = ctypes.c_int32()
        check(self.lib.ngpu_get_cycle_stats(self.h, ctypes.byref(sm), ctypes.byref(self._out), ctypes.byref(n)),
              "ngpu_get_cycle_stats")
        return sm.to_dict(self.S), self._out.to_dict(self.S), n.value

    def set_core(self, dets, fp32=False):
        """Semistochastic core space (ngpu_set_core); returns the number of stored H_ij."""
        dets = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.W)
        nnz = ctypes.c_int64()
        check(self.lib.ngpu_set_core(self.h, ptr(dets), dets.shape[0], int(bool(fp32)), ctypes.byref(nnz)),
              "ngpu_set_core")
        self.ncore = dets.shape[0]
        return nnz.value

    def set_core_build(self, method):
        """How set_core builds the core Hamiltonian: CORE_BUILD_AUTO (the
        default), CORE_BUILD_JOIN (connections by subset join) or
        CORE_BUILD_SCAN (every pair tested)."""
        check(self.lib.ngpu_set_core_build(self.h, int(method)), "ngpu_set_core_build")

    def core_hamil(self):
        """The stored core Hamiltonian as CSR: (rowptr, cols, vals)."""
        nnz = ctypes.c_int64()
        check(self.lib.ngpu_get_core_hamil(self.h, None, None, None, 0, ctypes.byref(nnz)), "ngpu_get_core_hamil")
        rowptr = np.zeros(self.ncore + 1, dtype=np.int64)
        cols = np.zeros(max(nnz.value, 1), dtype=np.int32)
        vals = np.zeros(max(nnz.value, 1), dtype=np.float64)
        check(self.lib.ngpu_get_core_hamil(self.h, ptr(rowptr), ptr(cols), ptr(vals), nnz.value, ctypes.byref(nnz)),
              "ngpu_get_core_hamil")
        return rowptr, cols[:nnz.value].copy(), vals[:nnz.value].copy()

    def most_populated(self, target):
        """The target most populated determinants (ngpu_most_populated), weight
        descending, equal weights in determinant order."""
        out = np.zeros((max(int(target), 1), self.W), dtype=np.int64)
        n = ctypes.c_int64()
        check(self.lib.ngpu_most_populated(self.h, int(target), ptr(out), ctypes.byref(n)), "ngpu_most_populated")
        return out[:n.value].copy()

    def select_core(self, target, fp32=False):
        """Core space of the target most populated determinants, chosen on the
        device (ngpu_select_core); returns (core determinants, stored H_ij)."""
        out = np.zeros((max(int(target), 1), self.W), dtype=np.int64)
        n = ctypes.c_int64()
        nnz = ctypes.c_int64()
        check(self.lib.ngpu_select_core(self.h, int(target), int(bool(fp32)), ptr(out), ctypes.byref(n),
                                        ctypes.byref(nnz)), "ngpu_select_core")
        self.ncore = n.value
        return out[:n.value].copy(), nnz.value

    def error(self):
        """Sticky error of the iteration chain: (status, iteration that failed),
        (0, 0) while the engine runs (ngpu_get_error)."""
        st = ctypes.c_int32()
        it = ctypes.c_uint32()
        check(self.lib.ngpu_get_error(self.h, ctypes.byref(st), ctypes.byref(it)), "ngpu_get_error")
        return st.value, it.value

    def hash_stats(self):
        """Walker-hash diagnostics: dict(live, tombstones, entries, rebuilds)."""
        v = [ctypes.c_int64() for _ in range(4)]
        check(self.lib.ngpu_get_hash_stats(self.h, *[ctypes.byref(x) for x in v]), "ngpu_get_hash_stats")
        return dict(live=v[0].value, tombstones=v[1].value, entries=v[2].value, rebuilds=v[3].value)

    def set_spawn_precision(self, fp32):
        """Spawned amplitudes stored in FP32 (True) or FP64 (ngpu_set_spawn_precision)."""
        check(self.lib.ngpu_set_spawn_precision(self.h, int(bool(fp32))), "ngpu_set_spawn_precision")
        self.spec.spawn_fp32 = bool(fp32)

    def set_checks(self, mask):
        """Invariant checks after every iteration (ngpu_set_checks): CHECK_TABLE,
        CHECK_PGEN, CHECK_SELFTEST bits; 0 switches them off."""
        check(self.lib.ngpu_set_checks(self.h, int(mask)), "ngpu_set_checks")

    def check_counts(self):
        """Cumulative counts: iterations, bad electron counts, hash misses,
        hole mismatches, live-count mismatches, pgen tests, pgen mismatches."""
        c = np.zeros(7, dtype=np.int64)
        check(self.lib.ngpu_get_check_counts(self.h, ptr(c), 7), "ngpu_get_check_counts")
        return dict(zip(("iterations", "bad_nel", "hash_misses", "hole_mismatches", "live_mismatches",
                         "pgen_tests", "pgen_mismatches"), (int(x) for x in c)))

    def set_launch_shape(self, chunk_blocks=0, grid_blocks=0, radix_items=0, single_lanes=0, merge_blocks=0):
        check(self.lib.ngpu_set_launch_shape(self.h, int(chunk_blocks), int(grid_blocks), int(radix_items),
                                             int(single_lanes), int(merge_blocks)), "ngpu_set_launch_shape")

    def launch_shape(self):
        v = [ctypes.c_int32() for _ in range(6)]
        check(self.lib.ngpu_get_launch_shape(self.h, *[ctypes.byref(x) for x in v]), "ngpu_get_launch_shape")
        return dict(chunk_blocks=v[0].value, grid_blocks=v[1].value, radix_items=v[2].value,
                    single_lanes=v[3].value, merge_blocks=v[4].value, n_cu=v[5].value)

    def set_graph(self, on=True):
        """Replay one captured graph per iteration (ngpu_set_graph)."""
        check(self.lib.ngpu_set_graph(self.h, int(on)), "ngpu_set_graph")

    N_STAGES = 10
    ALL_STAGES = (1 << N_STAGES) - 1

    def set_profiling(self, on=True):
        """Stage timing: True for every stage, False for none, or a bit mask of stages."""
        mask = self.ALL_STAGES if on is True else (0 if on is False else int(on))
        check(self.lib.ngpu_set_profiling(self.h, mask), "ngpu_set_profiling")

    def stage_times(self):
        ms = np.zeros(self.N_STAGES)
        n = ctypes.c_int64()
        check(self.lib.ngpu_get_stage_times(self.h, ptr(ms), self.N_STAGES, ctypes.byref(n)),
              "ngpu_get_stage_times")
        return ms, n.value

    # -- batched plugin-surface entry points
    def find_walker_hash(self, dets):
        dets = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.W)
        out = np.zeros(dets.shape[0], dtype=np.int32)
        check(self.lib.ngpu_find_walker_hash_batch(self.h, ptr(dets), dets.shape[0], ptr(out)),
              "ngpu_find_walker_hash_batch")
        return out

    def helement(self, di, dj):
        di = np.ascontiguousarray(di, dtype=np.int64).reshape(-1, self.W)
        dj = np.ascontiguousarray(dj, dtype=np.int64).reshape(-1, self.W)
        out = np.zeros(di.shape[0])
        check(self.lib.ngpu_get_spawn_helement_batch(self.h, ptr(di), ptr(dj), di.shape[0], ptr(out)),
              "ngpu_get_spawn_helement_batch")
        return out

    def diag(self, dets):
        dets = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.W)
        d = np.zeros(dets.shape[0])
        o = np.zeros(dets.shape[0])
        check(self.lib.ngpu_diag_helement_batch(self.h, ptr(dets), dets.shape[0], ptr(d), ptr(o)),
              "ngpu_diag_helement_batch")
        return d, o

    def generate_excitations(self, dets, draws):
        dets = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.W)
        draws = np.ascontiguousarray(draws, dtype=np.float64).reshape(-1, 6)
        n = dets.shape[0]
        dj = np.zeros((n, self.W), dtype=np.int64)
        ic = np.zeros(n, dtype=np.int32)
        ex = np.zeros((n, 4), dtype=np.int32)
        par = np.zeros(n, dtype=np.int32)
        pg = np.zeros(n)
        check(self.lib.ngpu_generate_excitation_batch(self.h, ptr(dets), ptr(draws), n, ptr(dj), ptr(ic),
                                                      ptr(ex), ptr(par), ptr(pg)),
              "ngpu_generate_excitation_batch")
        return dj, ic, ex, par, pg

    def calc_pgen(self, dets, ic, ex):
        dets = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.W)
        ic = np.ascontiguousarray(ic, dtype=np.int32)
        ex = np.ascontiguousarray(ex, dtype=np.int32).reshape(-1, 4)
        out = np.zeros(dets.shape[0])
        check(self.lib.ngpu_calc_pgen_batch(self.h, ptr(dets), ptr(ic), ptr(ex), dets.shape[0], ptr(out)),
              "ngpu_calc_pgen_batch")
        return out

    def attempt_die(self, kii, sign, u, tau, shift):
        kii, sign, u = (np.ascontiguousarray(x, dtype=np.float64) for x in (kii, sign, u))
        out = np.zeros(kii.size)
        check(self.lib.ngpu_attempt_die_batch(self.h, ptr(kii), ptr(sign), ptr(u), kii.size, tau, shift,
                                              ptr(out)), "ngpu_attempt_die_batch")
        return out

    def attempt_create(self, hel, pgen, psign, u, tau):
        hel, pgen, psign, u = (np.ascontiguousarray(x, dtype=np.float64) for x in (hel, pgen, psign, u))
        out = np.zeros(hel.size)
        check(self.lib.ngpu_attempt_create_batch(self.h, ptr(hel), ptr(pgen), ptr(psign), ptr(u), hel.size,
                                                 tau, ptr(out)), "ngpu_attempt_create_batch")
        return out

    def annihilate(self, it, sdet, sval, sinfo):
        sdet = np.ascontiguousarray(sdet, dtype=np.int64).reshape(-1, self.W)
        sval = np.ascontiguousarray(sval, dtype=np.float64)
        sinfo = np.ascontiguousarray(sinfo, dtype=np.uint32)
        n = sdet.shape[0]
        md = np.zeros((max(n, 1), self.W), dtype=np.int64)
        ms = np.zeros((max(n, 1), self.S))
        mf = np.zeros(max(n, 1), dtype=np.uint32)
        nm = ctypes.c_int64()
        out = IterOutT()
        check(self.lib.ngpu_annihilate(self.h, it, ptr(sdet), ptr(sval), ptr(sinfo), n, ctypes.byref(out),
                                       ptr(md), ptr(ms), ptr(mf), ctypes.byref(nm)), "ngpu_annihilate")
        k = nm.value
        return out.to_dict(self.S), md[:k], ms[:k], mf[:k]


class Batch:
    """Several engines iterated as one captured graph per iteration
    (ngpu_batch_create / ngpu_batch_iterate).  Engines keep their own
    statistics and walkers; each one's results are those of running it alone."""

    def __init__(self, engines):
        self.engines = list(engines)
        self.lib = _lib.load_library()
        hs = (ctypes.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        b = ctypes.c_void_p()
        check(self.lib.ngpu_batch_create(ctypes.byref(b), hs, len(self.engines)), "ngpu_batch_create")
        self.b = b
        self._ins = (IterInT * len(self.engines))()
        for e in self.engines:
            e.batch = self

    def iterate(self, its, taus, shifts):
        """One iteration of every engine: its[i], taus[i], shifts[i] (scalar or per replica)."""
        for i, e in enumerate(self.engines):
            inp = self._ins[i]
            inp.tau = float(taus[i])
            sh = np.broadcast_to(np.asarray(shifts[i], dtype=np.float64), (e.S,))
            for r in range(e.S):
                inp.diag_sft[r] = float(sh[r])
            inp.iter = int(its[i]) & 0xFFFFFFFF
        check(self.lib.ngpu_batch_iterate(self.b, self._ins), "ngpu_batch_iterate")

    def groups(self):
        """Launch group of each engine (ngpu_batch_get_groups): engines of one
        group share every launch of the iteration chain."""
        g = np.zeros(len(self.engines), dtype=np.int32)
        n = ctypes.c_int32()
        check(self.lib.ngpu_batch_get_groups(self.b, ptr(g), ctypes.byref(n)), "ngpu_batch_get_groups")
        return g, n.value

    def group_flags(self, group):
        """ngpu_batch_get_group_flags: GROUP_CORE | GROUP_SHARED_CORE bits."""
        f = ctypes.c_int32()
        check(self.lib.ngpu_batch_get_group_flags(self.b, int(group), ctypes.byref(f)), "ngpu_batch_get_group_flags")
        return f.value

    def set_graph(self, on=True):
        """Replay one captured graph per iteration (True) or launch directly."""
        check(self.lib.ngpu_batch_set_graph(self.b, int(bool(on))), "ngpu_batch_set_graph")

    def set_profiling(self, on=True):
        """Stage timing of the batch's launches (while launching directly):
        True for every stage, False for none, or a bit mask of stages."""
        mask = Engine.ALL_STAGES if on is True else (0 if on is False else int(on))
        check(self.lib.ngpu_batch_set_profiling(self.b, mask), "ngpu_batch_set_profiling")

    def group_stage_times(self, group):
        """(ms per stage summed over the timed iterations, iterations) of one
        launch group, held by the group's first engine."""
        g, _ = self.groups()
        first = int(np.flatnonzero(g == group)[0])
        return self.engines[first].stage_times()

    def close(self):
        if getattr(self, "b", None):
            self.lib.ngpu_batch_destroy(self.b)
            self.b = None
            for e in self.engines:
                e.batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
