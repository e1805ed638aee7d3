"""Shared builders for the parity tests (systems, walker-set conversion)."""
import numpy as np

from neci_stable_amd import HubbardChain, Spec, synthetic_molecule

MNECI_INIT = [13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 57, 58]
MNECI_DETERM = [41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 61, 62]

_MOL_CACHE = {}


def molecule(nspat, nel, rank=8, seed=7):
    key = (nspat, nel, rank, seed)
    if key not in _MOL_CACHE:
        _MOL_CACHE[key] = synthetic_molecule(nspat, nel, rank, seed)
    return _MOL_CACHE[key]


def hubbard_spec(nsites=10, nel=10, u=4.0, **kw):
    kw.setdefault("init_walkers", 2000)
    return Spec(HubbardChain(nsites, nel, u, bhub=-1.0), **kw)


def mol_spec(nspat=8, nel=6, rank=8, mseed=7, **kw):
    kw.setdefault("init_walkers", 2000)
    return Spec(molecule(nspat, nel, rank, mseed), **kw)


def internal_to_neci_flags(flags, S):
    out = np.zeros(len(flags), dtype=np.int64)
    for i, f in enumerate(flags):
        v = 0
        for r in range(S):
            if (int(f) >> r) & 1:
                v |= 1 << MNECI_INIT[r]
            if int(f) & (1 << 16):
                v |= 1 << MNECI_DETERM[r]
        out[i] = v
    return out


def neci_to_internal_flags(flags, S):
    out = np.zeros(len(flags), dtype=np.uint32)
    for i, f in enumerate(flags):
        v = 0
        for r in range(S):
            if (int(f) >> MNECI_INIT[r]) & 1:
                v |= 1 << r
            if (int(f) >> MNECI_DETERM[r]) & 1:
                v |= 1 << 16
        out[i] = v
    return out


def sort_walkers(dets, *arrs):
    dets = np.asarray(dets).view(np.int64)
    keys = [dets[:, w] for w in range(dets.shape[1] - 1, -1, -1)]
    o = np.lexsort(keys)
    return (dets[o],) + tuple(np.asarray(a)[o] for a in arrs)


class DeviceBuffer:
    """Device memory from the HIP runtime the engine library itself links
    (/opt/rocm/lib/libamdhip64.so.7), for tests of the entry points that take
    device pointers.  (torch carries its own HIP runtime; two runtimes in one
    process do not share the device, so tests avoid torch here.)"""

    _hip = None

    @classmethod
    def hip(cls):
        import ctypes
        if cls._hip is None:
            h = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes, fill_byte=None):
        import ctypes
        self.nbytes = max(int(nbytes), 8)
        self.ptr = ctypes.c_void_p()
        assert self.hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        if fill_byte is not None:
            assert self.hip().hipMemset(self.ptr, int(fill_byte), self.nbytes) == 0
            assert self.hip().hipDeviceSynchronize() == 0

    @property
    def addr(self):
        return self.ptr.value

    def to_host(self, dtype, count):
        out = np.zeros(max(int(count), 1), dtype=dtype)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out[:count]

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
