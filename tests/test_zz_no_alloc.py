"""ngpu_download_walkers_dev, the allocation-free download, against
ngpu_download_walkers through the C ABI.

Both return the occupied rows (and core rows) in slot order, so the device
compaction must give the host download's rows exactly, in the same order,
with the same det_diagH / det_offdiagH.  Two calls in a row share the one
scratch array of the GPU: a large walker set first, then a smaller one (so
stale counts of the first call lie past the second call's tiles), then the
first engine again with an output capacity below its row count.
"""
import ctypes

import numpy as np
import pytest

from common import DeviceBuffer, hubbard_spec, mol_spec

pytestmark = pytest.mark.gpu


def _grown(spec, iters, tau, n0, core=0):
    from neci_stable_amd.engine import Engine
    eng = Engine(spec)
    eng.start_single_part(n0)
    for it in range(1, iters + 1):
        eng.iterate(it, tau, np.zeros(spec.S), wait=False)
    if core:
        dets, sg, _, _, _ = eng.download()
        o = np.argsort(-np.abs(sg).sum(axis=1), kind="stable")[:core]
        eng.set_core(np.ascontiguousarray(dets[np.sort(o)]), fp32=False)
        eng.iterate(iters + 1, tau, np.zeros(spec.S))
    return eng


def _device_download(eng, cap):
    RL = eng.W + eng.S + 1
    rows = DeviceBuffer(max(cap, 1) * RL * 8, fill_byte=0xA5)
    dh = DeviceBuffer(max(cap, 1) * 8, fill_byte=0xA5)
    oh = DeviceBuffer(max(cap, 1) * 8, fill_byte=0xA5)
    n = DeviceBuffer(8, fill_byte=0xFF)
    eng.download_dev(rows.addr, dh.addr, oh.addr, cap, n.addr)
    assert eng.sync() == 0
    out = (rows.to_host(np.int64, max(cap, 1) * RL).reshape(-1, RL), dh.to_host(np.float64, max(cap, 1)),
           oh.to_host(np.float64, max(cap, 1)), int(n.to_host(np.int64, 1)[0]))
    for b in (rows, dh, oh, n):
        b.free()
    return out


def _host_download(eng):
    n = ctypes.c_int64()
    eng.lib.ngpu_download_walkers(eng.h, None, None, None, 0, ctypes.byref(n))
    m = n.value
    RL = eng.W + eng.S + 1
    rows = np.zeros((max(m, 1), RL), dtype=np.int64)
    dh = np.zeros(max(m, 1))
    oh = np.zeros(max(m, 1))
    from neci_stable_amd._lib import ptr
    assert eng.lib.ngpu_download_walkers(eng.h, ptr(rows), ptr(dh), ptr(oh), m, ctypes.byref(n)) == 0
    return rows[:m], dh[:m], oh[:m]


def _check_same(eng, label):
    hr, hd, ho = _host_download(eng)
    m = len(hr)
    dr, dd, do, n = _device_download(eng, m)
    print(f"{label}: {m} rows, W={eng.W}, S={eng.S}", flush=True)
    assert n == m
    assert np.array_equal(dr[:m], hr)
    assert np.array_equal(dd[:m].view(np.int64), hd.view(np.int64))
    assert np.array_equal(do[:m].view(np.int64), ho.view(np.int64))
    return m


def test_download_dev_matches_host_download_twice():
    # a two-replica two-word molecule with a core space (core flags, NECI's
    # multi-replica flag bits), then a small Hubbard chain
    big = _grown(mol_spec(33, 4, rank=4, mseed=3, n_replicas=2, init_walkers=20000), 60, 0.01, 200.0, core=40)
    small = _grown(hubbard_spec(10, 10, 4.0, init_walkers=500), 8, 0.01, 5.0)
    try:
        m_big = _check_same(big, "first call (2-word molecule, 2 replicas, core)")
        m_small = _check_same(small, "second call (Hubbard chain)")
        assert m_big > 4 * m_small > 0
        # output capacity below the row count: the count is reported, the
        # first cap rows are written, nothing past them
        hr, _, _ = _host_download(big)
        cap = len(hr) // 3
        dr, _, _, n = _device_download(big, cap)
        assert n == len(hr)
        assert np.array_equal(dr[:cap], hr[:cap])
    finally:
        big.close()
        small.close()


def test_download_dev_after_iterations_without_sync():
    """Queued asynchronous iterations are ordered before the download."""
    eng = _grown(hubbard_spec(12, 12, 4.0, init_walkers=3000), 5, 0.01, 300.0)
    try:
        for it in range(6, 26):
            eng.iterate(it, 0.01, np.zeros(1), wait=False)
        RL = eng.W + eng.S + 1
        cap = int(eng.max_walkers)
        rows = DeviceBuffer(cap * RL * 8)
        n = DeviceBuffer(8, fill_byte=0)
        eng.download_dev(rows.addr, None, None, cap, n.addr)
        assert eng.sync() == 0
        hr, _, _ = _host_download(eng)
        m = int(n.to_host(np.int64, 1)[0])
        assert m == len(hr) > 100
        assert np.array_equal(rows.to_host(np.int64, m * RL).reshape(m, RL), hr)
        rows.free()
        n.free()
    finally:
        eng.close()


def test_download_dev_refuses_oversized_tables():
    from neci_stable_amd._lib import NgpuError
    from neci_stable_amd.engine import Engine
    spec = hubbard_spec(10, 10, 4.0, init_walkers=500, max_walkers=(1 << 26) + 1)
    eng = Engine(spec)
    n = DeviceBuffer(8, fill_byte=0xFF)
    try:
        with pytest.raises(NgpuError) as e:
            eng.download_dev(None, None, None, 0, n.addr)
        assert e.value.code == 4
        assert eng.sync() == 0
        assert int(n.to_host(np.int64, 1)[0]) == -1  # nothing was launched
    finally:
        n.free()
        eng.close()
