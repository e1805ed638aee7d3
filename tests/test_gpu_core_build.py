"""The core Hamiltonian built from its connections (ngpu_set_core_build
NGPU_CORE_BUILD_JOIN: a join of the core determinants on their
(nel - 2)-electron subsets, the role of calc_determ_hamil_opt,
fast_determ_hamil.F90:892-1548) against the all-pairs scan (NGPU_CORE_BUILD_SCAN,
the role of calc_determ_hamil_sparse, sparse_arrays.F90:426-572) and the
oracle's core Hamiltonian: the same CSR (rows, columns, and values bit for bit
between the two device builds; within 1e-12 of the oracle's values)."""
import time

import numpy as np
import pytest

from common import mol_spec
from neci_stable_amd._lib import CORE_BUILD_AUTO, CORE_BUILD_JOIN, CORE_BUILD_SCAN
from neci_stable_amd.systems import decode_det, encode_det
from test_gpu_parity import SPECS, engine, oracle

pytestmark = pytest.mark.gpu


def _core_from_growth(spec, ncore, iters, tau, n0):
    eng = engine(spec)
    eng.start_single_part(n0)
    for it in range(1, iters + 1):
        eng.iterate(it, tau, np.full(spec.S, 0.2), wait=False)
    assert eng.sync() == 0
    core = eng.most_populated(ncore)
    eng.close()
    return core


def _built(spec, core, method, fp32):
    eng = engine(spec)
    eng.start_single_part(10.0)
    eng.set_core_build(method)
    t0 = time.perf_counter()
    nnz = eng.set_core(core, fp32=fp32)
    dt = time.perf_counter() - t0
    csr = eng.core_hamil()
    eng.close()
    assert len(csr[1]) == nnz
    return csr, dt


def _check_csr(csr, ncore):
    rp, cols, vals = csr
    assert rp[0] == 0 and rp[-1] == len(cols) and np.all(np.diff(rp) >= 1)
    for i in range(ncore):
        row = cols[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0), i
        assert i in row  # the diagonal is always stored


@pytest.mark.parametrize("name,nrep,ncore,iters,tau,n0", [
    ("mol8o6e_unif_fast", 1, 300, 25, 0.02, 400.0),
    ("mol33o4e_w2", 2, 700, 40, 0.02, 400.0),
    ("mol33o8e_w2", 1, 1500, 40, 0.01, 400.0),
    ("hubbard18", 1, 800, 40, 0.01, 40.0),
])
def test_join_equals_scan_and_oracle(name, nrep, ncore, iters, tau, n0):
    spec = SPECS[name](init_walkers=20000, n_replicas=nrep, seed=53)
    core = _core_from_growth(spec, ncore, iters, tau, n0)
    print(f"{name}: core of {len(core)}", flush=True)
    assert len(core) > ncore // 2
    orc = oracle(spec, tau)
    onnz = orc.set_core(core)
    orp, ocols, ovals = orc.core_csr()
    orc.close()
    for fp32 in (False, True):
        join, _ = _built(spec, core, CORE_BUILD_JOIN, fp32)
        scan, _ = _built(spec, core, CORE_BUILD_SCAN, fp32)
        _check_csr(join, len(core))
        for a, b in zip(join, scan):
            np.testing.assert_array_equal(a, b)
        assert len(join[1]) == onnz
        np.testing.assert_array_equal(join[0], orp)
        np.testing.assert_array_equal(join[1], ocols)
        expect = ovals.astype(np.float32).astype(np.float64) if fp32 else ovals
        np.testing.assert_allclose(join[2], expect, rtol=1e-12 if not fp32 else 1e-6, atol=1e-13)


def _singles_doubles(spec, n, rng):
    """n distinct determinants among the reference and its single and double
    excitations (spin conserved), drawn at random: a core whose
    (nel - 2)-subset groups reach their largest size (every double of one
    electron pair shares the reference less that pair)."""
    nb = spec.nbasis
    ref = list(spec.ref_orbs)
    occ = set(ref)
    virt = [o for o in range(1, nb + 1) if o not in occ]
    seen = {tuple(ref)}
    out = [encode_det(ref, nb)]
    while len(out) < n:
        k = 1 if rng.random() < 0.1 else 2
        el = list(rng.choice(ref, k, replace=False))
        holes = []
        for e in el:  # same spin as the electron
            cand = [v for v in virt if v % 2 == e % 2 and v not in holes]
            holes.append(int(rng.choice(cand)))
        orbs = tuple(sorted((occ - set(int(e) for e in el)) | set(holes)))
        if orbs in seen:
            continue
        seen.add(orbs)
        out.append(encode_det(list(orbs), nb))
    return np.ascontiguousarray(np.stack(out))


def test_join_equals_scan_40k_core_26e_40o():
    """A 40 000-determinant core of a 26-electron, 40-orbital molecule (two
    words per determinant): the join and the scan give the same CSR (both
    build times printed)."""
    spec = mol_spec(40, 26, rank=8, mseed=5, init_walkers=60000)
    core = _singles_doubles(spec, 40000, np.random.default_rng(3))
    assert all(len(decode_det(c.reshape(1, -1))) == 26 for c in core[:10])
    join, tj = _built(spec, core, CORE_BUILD_JOIN, False)
    scan, ts = _built(spec, core, CORE_BUILD_SCAN, False)
    auto, ta = _built(spec, core, CORE_BUILD_AUTO, True)
    print(f"nnz {len(join[1])}: join {tj:.3f} s, scan {ts:.3f} s, default (FP32 values) {ta:.3f} s", flush=True)
    _check_csr(join, len(core))
    for a, b in zip(join, scan):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(auto[0], join[0])
    np.testing.assert_array_equal(auto[1], join[1])
    np.testing.assert_array_equal(auto[2], join[2].astype(np.float32).astype(np.float64))


def test_core_build_rejects_repeated_determinants_and_late_method_change():
    from neci_stable_amd._lib import NgpuError
    spec = SPECS["mol8o6e_unif_fast"](init_walkers=2000)
    eng = engine(spec)
    eng.start_single_part(10.0)
    core = _core_from_growth(spec, 40, 10, 0.02, 400.0)
    with pytest.raises(NgpuError):
        eng.set_core(np.concatenate([core, core[3:4]]))
    eng.set_core(core)
    with pytest.raises(NgpuError):
        eng.set_core_build(CORE_BUILD_SCAN)
    eng.close()


def test_default_build_equals_oracle_50k_core_14e_30o():
    """The 50 000 most populated determinants of a grown 14e/30o population
    (seed 7, rank 60): the default build (the join at this size) gives the
    oracle's core Hamiltonian, rows and columns exactly and values to 1e-12."""
    from neci_stable_amd import Spec, synthetic_molecule
    spec = Spec(synthetic_molecule(30, 14, 60, 7), init_walkers=4e5, seed=71, particles=0, psingles=0.1,
                memory_fac_spawn=6.0)
    eng = engine(spec)
    eng.start_single_part(1000.0)
    it = 0
    while True:
        for _ in range(20):
            it += 1
            eng.iterate(it, 0.001, 0.0, wait=False)
        assert eng.sync() == 0
        if float(np.sum(eng.stats()["tot_parts"])) > 2e5 or it >= 600:
            break
    core = eng.most_populated(50000)
    eng.close()
    assert len(core) == 50000, (it, len(core))
    (rp, cols, vals), dt = _built(spec, core, CORE_BUILD_AUTO, False)
    print(f"50 000 core determinants after {it} iterations: {len(cols)} elements, built in {dt:.3f} s", flush=True)
    orc = oracle(spec, 0.001)
    assert orc.set_core(core) == len(cols)
    orp, ocols, ovals = orc.core_csr()
    orc.close()
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(cols, ocols)
    np.testing.assert_allclose(vals, ovals, rtol=1e-12, atol=1e-13)
