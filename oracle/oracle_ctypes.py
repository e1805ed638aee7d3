"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MAXR = 16


class OSys(ctypes.Structure):
    _fields_ = [
        ("nel", ctypes.c_int32), ("nbasis", ctypes.c_int32), ("nrep", ctypes.c_int32),
        ("model", ctypes.c_int32), ("nsites", ctypes.c_int32), ("maxnbr", ctypes.c_int32),
        ("nbr", ctypes.c_void_p), ("bhub", ctypes.c_double), ("uhub", ctypes.c_double),
        ("tmat", ctypes.c_void_p), ("umat", ctypes.c_void_p), ("umat_len", ctypes.c_int64),
        ("ecore", ctypes.c_double), ("tau", ctypes.c_double), ("avmc", ctypes.c_double),
        ("init_walk_no", ctypes.c_double), ("real_spawn_cutoff", ctypes.c_double),
        ("occ_thresh", ctypes.c_double), ("t_trunc_init", ctypes.c_int32),
        ("t_real_coeff", ctypes.c_int32), ("t_real_spawn_cutoff", ctypes.c_int32),
        ("spawn_fp32", ctypes.c_int32), ("hii", ctypes.c_double), ("ref", ctypes.c_uint64 * 2),
        ("singles", ctypes.c_int32), ("particles", ctypes.c_int32), ("psingles", ctypes.c_double),
        ("pparallel", ctypes.c_double), ("pchb_ready", ctypes.c_int32), ("uniform", ctypes.c_int32),
        ("random_hash2", ctypes.c_void_p), ("nbuckets", ctypes.c_int64), ("seed", ctypes.c_uint32),
        ("pad2", ctypes.c_uint32), ("max_walkers", ctypes.c_int64), ("max_spawned", ctypes.c_int64),
        ("holes", ctypes.c_int32), ("pad3", ctypes.c_int32),
    ]


_R = ctypes.c_double * MAXR
_FIELDS = ["nborn", "ndied", "nannihil", "naborted", "nremoved", "tot_parts", "norm_sq", "enum_cyc",
           "hf_cyc", "noat_doubs", "spawn_from_sing", "inst_no_at_hf", "no_init_dets", "no_init_walk",
           "no_noninit_dets", "no_noninit_walk"]


class OStats(ctypes.Structure):
    _fields_ = [(f, _R) for f in _FIELDS] + [
        ("n_valid_excits", ctypes.c_double), ("n_invalid_excits", ctypes.c_double),
        ("max_cyc_spawn", ctypes.c_double), ("bloom_count", ctypes.c_double * 3),
        ("bloom_size", ctypes.c_double * 3), ("tot_walkers", ctypes.c_int64),
        ("n_attempts", ctypes.c_int64), ("n_spawned", ctypes.c_int64), ("n_unique", ctypes.c_int64),
        ("n_new", ctypes.c_int64), ("err", ctypes.c_int32), ("pad", ctypes.c_int32),
    ]

    def to_dict(self, nrep):
        d = {f: np.array(getattr(self, f)[:nrep]) for f in _FIELDS}
        for f in ["n_valid_excits", "n_invalid_excits", "max_cyc_spawn", "tot_walkers", "n_attempts",
                  "n_spawned", "n_unique", "n_new", "err"]:
            d[f] = getattr(self, f)
        d["bloom_count"] = np.array(self.bloom_count[:])
        d["bloom_size"] = np.array(self.bloom_size[:])
        return d


_vp = ctypes.c_void_p
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle library missing: run make -C oracle")
        L = ctypes.CDLL(LIB)
        sig = {
            "o_philox4x32_10": (None, [_vp, _vp, _vp]),
            "o_find_walker_hash": (ctypes.c_int32, [_vp, ctypes.c_int32, _vp, ctypes.c_int64]),
            "o_umat_ind": (ctypes.c_uint64, [ctypes.c_int32] * 4),
            "o_decode": (ctypes.c_int32, [_vp, ctypes.c_int32, _vp]),
            "o_make_double": (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, ctypes.c_int32, _vp, _vp]),
            "o_make_single": (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp]),
            "o_get_helement": (ctypes.c_double, [_vp, _vp, _vp, ctypes.c_int32]),
            "o_diag_helement": (ctypes.c_double, [_vp, _vp]),
            "o_offdiag_helement": (ctypes.c_double, [_vp, _vp]),
            "o_build_pchb": (ctypes.c_int32, [_vp]),
            "o_free_pchb": (None, []),
            "o_alias_build": (None, [_vp, ctypes.c_int32, _vp, _vp, _vp]),
            "o_get_pchb_row": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp]),
            "o_gen_excit": (ctypes.c_int32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
            "o_calc_pgen": (ctypes.c_double, [_vp, _vp, ctypes.c_int32, _vp]),
            "o_hb_weight": (ctypes.c_double, [_vp, _vp, ctypes.c_int32, ctypes.c_int32]),
            "o_create": (_vp, [_vp]),
            "o_destroy": (None, [_vp]),
            "o_add_walker": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint32]),
            "o_get_walkers": (ctypes.c_int64, [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int64]),
            "o_add_walkers": (ctypes.c_int32, [_vp, _vp, _vp, _vp, ctypes.c_int64]),
            "o_iterate": (ctypes.c_int32, [_vp, ctypes.c_uint32, _vp, _vp]),
            "o_annihilate_list": (ctypes.c_int32, [_vp, ctypes.c_uint32, _vp, _vp, _vp, ctypes.c_int64, _vp,
                                                   _vp, _vp, _vp, _vp]),
            "o_set_core": (ctypes.c_int32, [_vp, _vp, ctypes.c_int64, ctypes.c_int32, _vp]),
            "o_core_project": (None, [_vp, _vp, _vp, ctypes.c_double, _vp]),
            "o_core_csr": (ctypes.c_int64, [_vp, _vp, _vp, _vp]),
            "o_most_populated": (ctypes.c_int64, [_vp, ctypes.c_int64, _vp]),
            "o_gen_excit_ctx": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32,
                                                 _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
            "o_mt_create": (_vp, [_vp, ctypes.c_int32]),
            "o_mt_destroy": (None, [_vp]),
            "o_mt_add_walkers": (ctypes.c_int32, [_vp, _vp, _vp, _vp, ctypes.c_int64]),
            "o_mt_get_walkers": (ctypes.c_int64, [_vp, _vp, _vp, _vp, ctypes.c_int64]),
            "o_mt_set_core": (ctypes.c_int32, [_vp, _vp, ctypes.c_int64, ctypes.c_int32, _vp]),
            "o_mt_iterate": (ctypes.c_int32, [_vp, ctypes.c_uint32, _vp, _vp]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """CPU restatement of PerformFCIMCycPar over the same system as an Engine.

    Built from an engine's host-side parameters so the two see identical
    inputs (integrals, lattice, reference, RandomHash2, pSingles, seed).
    """

    def __init__(self, spec, tau, max_walkers=None, max_spawned=None):
        """spec: a neci_stable_amd.config.Spec or an Engine (uses engine.spec)."""
        spec = getattr(spec, "spec", spec)
        L = lib()
        self.L = L
        self.W = spec.W
        self.S = spec.S
        s = spec.system
        self.keep = []
        o = OSys()
        o.nel, o.nbasis, o.nrep = spec.nel, spec.nbasis, spec.S
        o.model = spec.model
        if spec.model == 0:
            nbr = np.ascontiguousarray(s.nbr, dtype=np.int32)
            self.keep.append(nbr)
            o.nsites, o.maxnbr, o.nbr, o.bhub, o.uhub = s.nsites, nbr.shape[1], nbr.ctypes.data, s.bhub, s.uhub
        else:
            tm = np.ascontiguousarray(s.tmat_colmajor())
            um = np.ascontiguousarray(s.umat, dtype=np.float64)
            self.keep += [tm, um]
            o.tmat, o.umat, o.umat_len, o.ecore = tm.ctypes.data, um.ctypes.data, um.size, s.ecore
            o.singles, o.particles = 1, spec.particles
            o.holes = int(getattr(spec, "holes", 0))
            o.psingles, o.pparallel = spec.psingles, spec.pparallel
            o.uniform = int(getattr(spec, "excitgen", "pchb") == "uniform")
        o.tau = float(tau)
        o.avmc = spec.av_mc_excits
        o.init_walk_no = spec.init_walk_no
        o.real_spawn_cutoff = spec.real_spawn_cutoff
        o.occ_thresh = spec.occupied_thresh
        o.t_trunc_init = int(spec.trunc_initiator)
        o.t_real_coeff = int(spec.all_real_coeff)
        o.t_real_spawn_cutoff = int(spec.t_real_spawn_cutoff)
        o.spawn_fp32 = int(getattr(spec, "spawn_fp32", False))
        W = spec.W
        ref = np.zeros(W, dtype=np.uint64)
        for orb in spec.ref_orbs:
            ref[(orb - 1) >> 6] |= np.uint64(1) << np.uint64((orb - 1) & 63)
        self.ref = ref.view(np.int64)
        o.ref[0] = int(ref[0])
        o.ref[1] = int(ref[1]) if W > 1 else 0
        rh = np.ascontiguousarray(spec.random_hash, dtype=np.int32)
        self.keep.append(rh)
        o.random_hash2 = rh.ctypes.data
        o.nbuckets = int(spec.n_hash_buckets)
        o.seed = int(spec.seed) & 0xFFFFFFFF
        o.max_walkers = int(max_walkers or spec.max_walkers)
        o.max_spawned = int(max_spawned or spec.max_spawned)
        self.sys = o
        if spec.model == 1 and not o.uniform:
            L.o_build_pchb(ctypes.byref(o))
        o.hii = 0.0
        self.hii = self.diag(spec.ref_orbs)
        o.hii = self.hii
        self.st = L.o_create(ctypes.byref(o))
        self.out = OStats()

    def close(self):
        if self.st:
            self.L.o_destroy(self.st)
            self.st = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_walkers(self, dets, signs, flags):
        dets = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.W).view(np.uint64)
        signs = np.ascontiguousarray(signs, dtype=np.float64).reshape(-1, self.S)
        for i in range(dets.shape[0]):
            d = np.ascontiguousarray(dets[i])
            sg = np.ascontiguousarray(signs[i])
            rc = self.L.o_add_walker(self.st, p(d), p(sg), int(flags[i]))
            if rc:
                raise RuntimeError(f"o_add_walker {rc}")

    def add_walkers_bulk(self, dets, signs, flags):
        dets = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.W).view(np.uint64)
        signs = np.ascontiguousarray(signs, dtype=np.float64).reshape(-1, self.S)
        flags = np.ascontiguousarray(flags, dtype=np.uint32)
        rc = self.L.o_add_walkers(self.st, p(dets), p(signs), p(flags), dets.shape[0])
        if rc:
            raise RuntimeError(f"o_add_walkers {rc}")

    def walkers(self, cap=10_000_000):
        n = self.L.o_get_walkers(self.st, None, None, None, None, None, 0)
        d = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        s = np.zeros((max(n, 1), self.S))
        f = np.zeros(max(n, 1), dtype=np.uint32)
        dh = np.zeros(max(n, 1))
        oh = np.zeros(max(n, 1))
        self.L.o_get_walkers(self.st, p(d), p(s), p(f), p(dh), p(oh), n)
        return d[:n].view(np.int64), s[:n], f[:n], dh[:n], oh[:n]

    def iterate(self, it, shift):
        sh = np.zeros(MAXR)
        sh[:self.S] = np.broadcast_to(np.asarray(shift, dtype=np.float64), (self.S,))
        rc = self.L.o_iterate(self.st, int(it) & 0xFFFFFFFF, p(sh), ctypes.byref(self.out))
        d = self.out.to_dict(self.S)
        d["rc"] = rc
        return d

    def annihilate(self, it, sdet, sval, sinfo):
        sdet = np.ascontiguousarray(sdet, dtype=np.int64).reshape(-1, self.W).view(np.uint64)
        sval = np.ascontiguousarray(sval, dtype=np.float64)
        sinfo = np.ascontiguousarray(sinfo, dtype=np.uint32)
        n = sdet.shape[0]
        md = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        ms = np.zeros((max(n, 1), self.S))
        mf = np.zeros(max(n, 1), dtype=np.uint32)
        nm = ctypes.c_int64()
        rc = self.L.o_annihilate_list(self.st, it, p(sdet), p(sval), p(sinfo), n, ctypes.byref(self.out),
                                      p(md), p(ms), p(mf), ctypes.byref(nm))
        d = self.out.to_dict(self.S)
        d["rc"] = rc
        k = nm.value
        return d, md[:k].view(np.int64), ms[:k], mf[:k]

    def set_core(self, dets, fp32=False):
        """Semistochastic core space; returns the number of stored H_ij."""
        dets = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.W).view(np.uint64)
        self.ncore = dets.shape[0]
        nnz = ctypes.c_int64()
        rc = self.L.o_set_core(self.st, p(dets), self.ncore, int(bool(fp32)), ctypes.byref(nnz))
        if rc:
            raise RuntimeError(f"o_set_core {rc}")
        return nnz.value

    def most_populated(self, target):
        """generate_space_most_populated: the target most populated
        determinants (weight descending, ties by determinant)."""
        out = np.zeros((max(int(target), 1), self.W), dtype=np.uint64)
        n = self.L.o_most_populated(self.st, int(target), p(out))
        return out[:n].view(np.int64)

    def core_csr(self):
        nnz = self.L.o_core_csr(self.st, None, None, None)
        rp = np.zeros(self.ncore + 1, dtype=np.int64)
        cols = np.zeros(max(nnz, 1), dtype=np.int32)
        vals = np.zeros(max(nnz, 1))
        self.L.o_core_csr(self.st, p(rp), p(cols), p(vals))
        return rp, cols[:nnz], vals[:nnz]

    def core_project(self, x, shift, tau):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.ncore, self.S)
        sh = np.ascontiguousarray(np.broadcast_to(np.asarray(shift, dtype=np.float64), (self.S,)))
        y = np.zeros_like(x)
        self.L.o_core_project(self.st, p(x), p(sh), float(tau), p(y))
        return y

    def gen_excit(self, det, draws):
        det = np.ascontiguousarray(det, dtype=np.int64).view(np.uint64)
        draws = np.ascontiguousarray(draws, dtype=np.float64)
        nJ = np.zeros(64, dtype=np.int32)
        dj = np.zeros(self.W, dtype=np.uint64)
        ic = ctypes.c_int32()
        ex = np.zeros(4, dtype=np.int32)
        tp = ctypes.c_int32()
        pg = ctypes.c_double()
        ok = self.L.o_gen_excit(ctypes.byref(self.sys), p(det), p(draws), p(nJ), p(dj), ctypes.byref(ic),
                                p(ex), ctypes.byref(tp), ctypes.byref(pg))
        if not ok:
            return None
        return dj.view(np.int64), ic.value, ex.copy(), tp.value, pg.value

    def gen_excit_ctx(self, det, draws, it, rep, k):
        """gen_excit with the draw context of spawn attempt k (rejection-loop draws); adds fail."""
        det = np.ascontiguousarray(det, dtype=np.int64).view(np.uint64)
        draws = np.ascontiguousarray(draws, dtype=np.float64)
        nJ = np.zeros(64, dtype=np.int32)
        dj = np.zeros(self.W, dtype=np.uint64)
        ic = ctypes.c_int32()
        ex = np.zeros(4, dtype=np.int32)
        tp = ctypes.c_int32()
        pg = ctypes.c_double()
        fail = ctypes.c_int32()
        ok = self.L.o_gen_excit_ctx(ctypes.byref(self.sys), p(det), p(draws), int(it) & 0xFFFFFFFF, int(rep),
                                    int(k) & 0xFFFFFFFF, p(nJ), p(dj), ctypes.byref(ic), p(ex), ctypes.byref(tp),
                                    ctypes.byref(pg), ctypes.byref(fail))
        if not ok:
            return None, fail.value
        return (dj.view(np.int64), ic.value, ex.copy(), tp.value, pg.value), fail.value

    def hb_weight(self, det, src, orb):
        """The heat-bath single weight H(src -> orb) relative to the reference
        (signed, no parity): the value a heat-bath single spawns with."""
        det = np.ascontiguousarray(det, dtype=np.int64).view(np.uint64)
        return self.L.o_hb_weight(ctypes.byref(self.sys), p(det), int(src), int(orb))

    def calc_pgen(self, det, ic, ex):
        det = np.ascontiguousarray(det, dtype=np.int64).view(np.uint64)
        ex = np.ascontiguousarray(ex, dtype=np.int32)
        return self.L.o_calc_pgen(ctypes.byref(self.sys), p(det), int(ic), p(ex))

    def helement(self, orbs_i, orbs_j):
        a = np.ascontiguousarray(orbs_i, dtype=np.int32)
        b = np.ascontiguousarray(orbs_j, dtype=np.int32)
        return self.L.o_get_helement(ctypes.byref(self.sys), p(a), p(b), len(a))

    def diag(self, orbs):
        a = np.ascontiguousarray(orbs, dtype=np.int32)
        return self.L.o_diag_helement(ctypes.byref(self.sys), p(a))

    def offdiag(self, det):
        d = np.ascontiguousarray(det, dtype=np.int64).view(np.uint64)
        return self.L.o_offdiag_helement(ctypes.byref(self.sys), p(d))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().o_philox4x32_10(c, k, o)
    return list(o)


def find_walker_hash(orbs, rh, L):
    a = np.ascontiguousarray(orbs, dtype=np.int32)
    r = np.ascontiguousarray(rh, dtype=np.int32)
    return lib().o_find_walker_hash(p(a), len(a), p(r), int(L))


def alias_build(w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    n = w.size
    b = np.zeros(n)
    a = np.zeros(n, dtype=np.int32)
    pr = np.zeros(n)
    lib().o_alias_build(p(w), n, p(b), p(a), p(pr))
    return b, a, pr


def make_double(nI, e1, e2, t1, t2):
    a = np.ascontiguousarray(nI, dtype=np.int32)
    nJ = np.zeros(len(a), dtype=np.int32)
    ex = np.zeros(4, dtype=np.int32)
    par = lib().o_make_double(p(a), len(a), e1, e2, t1, t2, p(nJ), p(ex))
    return nJ, ex, par


def make_single(nI, elec, tgt):
    a = np.ascontiguousarray(nI, dtype=np.int32)
    nJ = np.zeros(len(a), dtype=np.int32)
    ex = np.zeros(2, dtype=np.int32)
    par = lib().o_make_single(p(a), len(a), elec, tgt, p(nJ), p(ex))
    return nJ, ex, par


class ShardedOracle:
    """The oracle's iteration over P hash-sharded states on P threads (one
    per host core, as the reference runs one MPI rank per core).  Same
    system set-up as Oracle; walkers are routed to their owner shard."""

    def __init__(self, spec, tau, nshards, max_walkers=None, max_spawned=None):
        base = Oracle(spec, tau, max_walkers, max_spawned)  # system tables and PCHB build
        self.base = base
        self.L = base.L
        self.W, self.S, self.ref, self.hii = base.W, base.S, base.ref, base.hii
        self.P = int(nshards)
        self.m = self.L.o_mt_create(ctypes.byref(base.sys), self.P)
        if not self.m:
            raise RuntimeError("o_mt_create failed")
        self.out = OStats()
        self.ncore = 0

    def close(self):
        if getattr(self, "m", None):
            self.L.o_mt_destroy(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_walkers_bulk(self, dets, signs, flags):
        dets = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.W).view(np.uint64)
        signs = np.ascontiguousarray(signs, dtype=np.float64).reshape(-1, self.S)
        flags = np.ascontiguousarray(flags, dtype=np.uint32)
        rc = self.L.o_mt_add_walkers(self.m, p(dets), p(signs), p(flags), dets.shape[0])
        if rc:
            raise RuntimeError(f"o_mt_add_walkers {rc}")

    def set_core(self, dets, fp32=False):
        dets = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.W).view(np.uint64)
        nnz = ctypes.c_int64()
        rc = self.L.o_mt_set_core(self.m, p(dets), dets.shape[0], int(bool(fp32)), ctypes.byref(nnz))
        if rc:
            raise RuntimeError(f"o_mt_set_core {rc}")
        self.ncore = dets.shape[0]
        return nnz.value

    def walkers(self, cap=20_000_000):
        n = self.L.o_mt_get_walkers(self.m, None, None, None, 0)
        d = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        s = np.zeros((max(n, 1), self.S))
        f = np.zeros(max(n, 1), dtype=np.uint32)
        n = self.L.o_mt_get_walkers(self.m, p(d), p(s), p(f), n)
        return d[:n].view(np.int64), s[:n], f[:n]

    def iterate(self, it, shift):
        sh = np.ascontiguousarray(np.broadcast_to(np.asarray(shift, dtype=np.float64), (self.S,)))
        rc = self.L.o_mt_iterate(self.m, int(it) & 0xFFFFFFFF, p(sh), ctypes.byref(self.out))
        d = self.out.to_dict(self.S)
        d["rc"] = rc
        return d

