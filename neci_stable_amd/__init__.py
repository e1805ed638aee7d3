"""MI355X-native FCIQMC iteration engine (drop-in for NECI's PerformFCIMCycPar).

The compute path is the HIP library libneci_gpu.so behind the C ABI of
include/neci_gpu.h; this package holds the host side: system set-up
(lattice, FCIDUMP, reference, RandomHash2), the ctypes binding and the
FciMCPar-style driver with NECI's shift update.
"""
from .config import Spec  # noqa: F401
from .systems import (HubbardChain, MolecularSystem, decode_det, encode_det,  # noqa: F401
                      find_walker_hash, random_hash2, read_fcidump, synthetic_molecule,
                      write_fcidump)


def Engine(*a, **k):
    """GPU engine handle (loads libneci_gpu.so; raises if it is missing)."""
    from .engine import Engine as _E
    return _E(*a, **k)
