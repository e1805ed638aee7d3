"""Core space chosen on the device (ngpu_most_populated / ngpu_select_core,
generate_space_most_populated, semi_stoch_gen.F90:1031-1226) against the
oracle's selection on the same walker set, then semistochastic iterations
on that core space against the oracle."""
import numpy as np
import pytest

from common import neci_to_internal_flags
from test_gpu_parity import SPECS, engine, oracle

pytestmark = pytest.mark.gpu


def _grown(name, nrep, iters, tau=0.02, n0=400.0, **kw):
    spec = SPECS[name](init_walkers=8000, n_replicas=nrep, seed=41, **kw)
    eng = engine(spec)
    eng.start_single_part(n0)
    shift = np.linspace(0.2, 0.3, nrep)
    for it in range(1, iters + 1):
        eng.iterate(it, tau, shift, wait=False)
    assert eng.sync() == 0
    return spec, eng, shift


@pytest.mark.parametrize("name,nrep", [("mol8o6e_unif_fast", 2), ("mol33o4e_w2", 1), ("hubbard18", 1)])
def test_most_populated_matches_oracle(name, nrep):
    spec, eng, _ = _grown(name, nrep, 25, tau=0.01 if name.startswith("hub") else 0.02,
                          n0=40.0 if name.startswith("hub") else 400.0)
    d, s, f, _, _ = eng.download()
    orc = oracle(spec, 0.02)
    orc.add_walkers_bulk(d, s, neci_to_internal_flags(f, nrep))
    print(f"{name}: {len(d)} determinants", flush=True)
    assert len(d) > 300
    for target in (1, 37, 300, len(d) + 5):
        np.testing.assert_array_equal(eng.most_populated(target), orc.most_populated(target))
    eng.close()


def test_select_core_then_iterations_match_oracle():
    name, nrep, tau = "mol8o6e_unif_fast", 1, 0.02
    spec, eng, shift = _grown(name, nrep, 20)
    d, s, f, _, _ = eng.download()
    orc = oracle(spec, tau)
    orc.add_walkers_bulk(d, s, neci_to_internal_flags(f, nrep))
    core, nnz = eng.select_core(120)
    np.testing.assert_array_equal(core, orc.most_populated(120))
    assert orc.set_core(core) == nnz
    for it in range(21, 41):
        a = eng.iterate(it, tau, shift)
        b = orc.iterate(it, shift)
        assert b["rc"] == 0
        for fld in ("tot_walkers", "n_attempts", "n_spawned"):
            assert a[fld] == b[fld], (fld, it)
        np.testing.assert_allclose(a["tot_parts"][:1], b["tot_parts"][:1], rtol=1e-10)
    eng.close()
