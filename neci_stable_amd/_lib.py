"""ctypes binding of the C ABI in include/neci_gpu.h.

The engine library (neci_stable_amd/libneci_gpu.so, gfx950 code objects) is
the only compute path: if it is missing or cannot be loaded, every entry point
raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

MAX_REPLICAS = 16
LIB_NAME = "libneci_gpu.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

MODEL_HUBBARD = 0
MODEL_MOLECULAR = 1
SINGLES_HEATBATH = 1
PART_UNIF_UNIF = 0
PART_UNIF_FAST = 1
PART_UNIF_FULL = 2
PART_FULL_FULL = 3
HOLES_FAST_FAST = 0
HOLES_FULL_FULL = 1
CHECK_TABLE = 1
CHECK_PGEN = 2
CHECK_SELFTEST = 4
GROUP_CORE = 1
GROUP_SHARED_CORE = 2
CORE_BUILD_AUTO = 0
CORE_BUILD_JOIN = 1
CORE_BUILD_SCAN = 2
FLAGS_MNECI = 0
FLAGS_NECI = 1

STATUS = {
    0: "NGPU_OK",
    1: "NGPU_ERR_SPAWN_FULL (increase MEMORYFACSPAWN)",
    2: "NGPU_ERR_WALKERS_FULL (increase MEMORYFACPART)",
    3: "NGPU_ERR_DEATH_PROB (death probability > 2: reduce the time step)",
    4: "NGPU_ERR_ARG",
    5: "NGPU_ERR_HIP",
    6: "NGPU_ERR_HASH_FULL (increase MEMORYFACHASH)",
    7: "NGPU_ERR_STATE",
    8: "NGPU_ERR_EXCITGEN (uniform generator found no allowed orbital)",
    9: "NGPU_ERR_CHECK (an invariant check failed)",
}


class NgpuError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: {STATUS.get(code, code)}")


class SysT(ctypes.Structure):
    _fields_ = [
        ("nel", ctypes.c_int32), ("nbasis", ctypes.c_int32), ("n_replicas", ctypes.c_int32),
        ("model", ctypes.c_int32), ("t_trunc_initiator", ctypes.c_int32),
        ("t_all_real_coeff", ctypes.c_int32), ("t_real_spawn_cutoff", ctypes.c_int32),
        ("flag_layout", ctypes.c_int32), ("initiator_walk_no", ctypes.c_double),
        ("real_spawn_cutoff", ctypes.c_double), ("occupied_thresh", ctypes.c_double),
        ("av_mc_excits", ctypes.c_double), ("max_walkers", ctypes.c_int64),
        ("max_spawned", ctypes.c_int64), ("n_hash_buckets", ctypes.c_int64),
        ("seed", ctypes.c_uint32), ("pad", ctypes.c_int32),
    ]


class IterInT(ctypes.Structure):
    _fields_ = [("tau", ctypes.c_double), ("diag_sft", ctypes.c_double * MAX_REPLICAS),
                ("iter", ctypes.c_uint32), ("pad", ctypes.c_int32)]


_R = ctypes.c_double * MAX_REPLICAS
REPLICA_FIELDS = ["nborn", "ndied", "nannihil", "naborted", "nremoved", "tot_parts", "norm_sq",
                  "enum_cyc", "hf_cyc", "noat_doubs", "spawn_from_sing", "inst_no_at_hf",
                  "no_init_dets", "no_init_walk", "no_noninit_dets", "no_noninit_walk"]


class IterOutT(ctypes.Structure):
    _fields_ = [(f, _R) for f in REPLICA_FIELDS] + [
        ("n_valid_excits", ctypes.c_double), ("n_invalid_excits", ctypes.c_double),
        ("max_cyc_spawn", ctypes.c_double), ("bloom_count", ctypes.c_double * 3),
        ("bloom_size", ctypes.c_double * 3), ("tot_walkers", ctypes.c_int64),
        ("n_attempts", ctypes.c_int64), ("n_spawned", ctypes.c_int64), ("n_unique", ctypes.c_int64),
        ("n_new", ctypes.c_int64), ("err", ctypes.c_int32), ("pad", ctypes.c_int32),
    ]

    def to_dict(self, nrep):
        d = {f: np.array(getattr(self, f)[:nrep]) for f in REPLICA_FIELDS}
        for f in ["n_valid_excits", "n_invalid_excits", "max_cyc_spawn", "tot_walkers",
                  "n_attempts", "n_spawned", "n_unique", "n_new", "err"]:
            d[f] = getattr(self, f)
        d["bloom_count"] = np.array(self.bloom_count[:])
        d["bloom_size"] = np.array(self.bloom_size[:])
        return d


P = ctypes.POINTER
_vp = ctypes.c_void_p

# (name, restype, argtypes)
_SIGS = [
    ("ngpu_create", ctypes.c_int, [P(_vp), P(SysT)]),
    ("ngpu_destroy", ctypes.c_int, [_vp]),
    ("ngpu_fcidump_header", ctypes.c_int, [ctypes.c_char_p, P(ctypes.c_int32), P(ctypes.c_int32),
                                           P(ctypes.c_int32), P(ctypes.c_int32), _vp, ctypes.c_int32]),
    ("ngpu_read_fcidump", ctypes.c_int, [ctypes.c_char_p, ctypes.c_double, _vp, ctypes.c_int64, _vp,
                                         ctypes.c_int64, P(ctypes.c_double), P(ctypes.c_int64)]),
    ("ngpu_set_integrals", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_double]),
    ("ngpu_set_lattice", ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                        ctypes.c_double]),
    ("ngpu_set_excitgen_uniform", ctypes.c_int, [_vp, ctypes.c_double]),
    ("ngpu_set_spawn_precision", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("ngpu_set_excitgen", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                         ctypes.c_double]),
    ("ngpu_set_excitgen_pchb", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_double, ctypes.c_double]),
    ("ngpu_set_hash", ctypes.c_int, [_vp, _vp]),
    ("ngpu_set_reference", ctypes.c_int, [_vp, _vp, ctypes.c_double]),
    ("ngpu_set_core", ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int32, P(ctypes.c_int64)]),
    ("ngpu_set_core_build", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("ngpu_get_core_hamil", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int64, P(ctypes.c_int64)]),
    ("ngpu_most_populated", ctypes.c_int, [_vp, ctypes.c_int64, _vp, P(ctypes.c_int64)]),
    ("ngpu_select_core", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, _vp, P(ctypes.c_int64),
                                        P(ctypes.c_int64)]),
    ("ngpu_upload_walkers", ctypes.c_int, [_vp, _vp, ctypes.c_int64]),
    ("ngpu_download_walkers", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int64, P(ctypes.c_int64)]),
    ("ngpu_download_walkers_dev", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int64, _vp]),
    ("ngpu_iterate", ctypes.c_int, [_vp, P(IterInT), P(IterOutT)]),
    ("ngpu_get_stats", ctypes.c_int, [_vp, P(IterOutT)]),
    ("ngpu_get_cycle_stats", ctypes.c_int, [_vp, P(IterOutT), P(IterOutT), P(ctypes.c_int32)]),
    ("ngpu_get_error", ctypes.c_int, [_vp, P(ctypes.c_int32), P(ctypes.c_uint32)]),
    ("ngpu_get_hash_stats", ctypes.c_int, [_vp, P(ctypes.c_int64), P(ctypes.c_int64), P(ctypes.c_int64),
                                           P(ctypes.c_int64)]),
    ("ngpu_set_checks", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("ngpu_get_check_counts", ctypes.c_int, [_vp, _vp, ctypes.c_int32]),
    ("ngpu_set_profiling", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("ngpu_set_graph", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("ngpu_set_launch_shape", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32]),
    ("ngpu_get_launch_shape", ctypes.c_int, [_vp, P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int32),
                                             P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int32)]),
    ("ngpu_batch_create", ctypes.c_int, [P(_vp), _vp, ctypes.c_int32]),
    ("ngpu_batch_iterate", ctypes.c_int, [_vp, _vp]),
    ("ngpu_batch_destroy", ctypes.c_int, [_vp]),
    ("ngpu_batch_get_groups", ctypes.c_int, [_vp, _vp, P(ctypes.c_int32)]),
    ("ngpu_batch_get_group_flags", ctypes.c_int, [_vp, ctypes.c_int32, P(ctypes.c_int32)]),
    ("ngpu_batch_set_graph", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("ngpu_batch_set_profiling", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("ngpu_get_stage_times", ctypes.c_int, [_vp, _vp, ctypes.c_int32, P(ctypes.c_int64)]),
    ("ngpu_find_walker_hash_batch", ctypes.c_int, [_vp, _vp, ctypes.c_int64, _vp]),
    ("ngpu_get_spawn_helement_batch", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, _vp]),
    ("ngpu_diag_helement_batch", ctypes.c_int, [_vp, _vp, ctypes.c_int64, _vp, _vp]),
    ("ngpu_generate_excitation_batch", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, _vp, _vp, _vp,
                                                      _vp, _vp]),
    ("ngpu_calc_pgen_batch", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int64, _vp]),
    ("ngpu_encode_child", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp, _vp]),
    ("ngpu_attempt_die_batch", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_double,
                                              ctypes.c_double, _vp]),
    ("ngpu_attempt_create_batch", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int64,
                                                 ctypes.c_double, _vp]),
    ("ngpu_annihilate", ctypes.c_int, [_vp, ctypes.c_uint32, _vp, _vp, _vp, ctypes.c_int64,
                                       P(IterOutT), _vp, _vp, _vp, P(ctypes.c_int64)]),
]
EXPORTED = [s[0] for s in _SIGS]

_lib = None


def load_library(path=None):
    """Load the HIP engine; raises if it is absent (no fallback path exists)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP engine library {p} not found: run __graft_entry__.build() first")
    lib = ctypes.CDLL(p)
    for name, res, args in _SIGS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code, where):
    if code != 0:
        raise NgpuError(code, where)


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
