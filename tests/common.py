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
