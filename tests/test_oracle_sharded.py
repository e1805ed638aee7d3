"""The hash-sharded, multi-threaded oracle (the CPU baseline's form: P
threads, each owning the determinants DetermineDetNode's role assigns it,
load_balance_calcnodes.F90:25-117) against the one-state oracle (CPU only).

With integer walkers every random number is keyed by determinant and the
merged sums are sums of integers, so the sharded iteration reproduces the
one-state iteration exactly: the same walker set and the same integer
statistics after every iteration, for any shard count.  With a core space
the core amplitudes are real and the projection's rows are split over the
threads in the same summation order, so the two agree to rounding.
"""
import numpy as np
import pytest

import oracle_ctypes as oc
from common import hubbard_spec, mol_spec, sort_walkers

INT_FIELDS = ["nborn", "ndied", "nannihil", "naborted", "nremoved", "tot_parts", "norm_sq", "hf_cyc",
              "noat_doubs", "spawn_from_sing", "inst_no_at_hf", "no_init_dets", "no_init_walk"]


def _pair(spec, tau, nshards, n0):
    a = oc.Oracle(spec, tau)
    b = oc.ShardedOracle(spec, tau, nshards)
    sg = np.full((1, spec.S), n0)
    a.add_walkers(a.ref[None, :], sg, [(1 << spec.S) - 1])
    b.add_walkers_bulk(b.ref[None, :], sg, np.array([(1 << spec.S) - 1], dtype=np.uint32))
    return a, b


@pytest.mark.parametrize("name,nshards,nrep", [("hubbard", 3, 1), ("hubbard", 8, 2), ("mol", 4, 1),
                                               ("mol_unif", 5, 1), ("mol_plain", 2, 1)])
def test_sharded_matches_one_state(name, nshards, nrep):
    if name == "hubbard":
        spec, tau, n0 = hubbard_spec(10, 10, 4.0, init_walkers=3000, n_replicas=nrep, seed=5), 0.01, 40.0
    elif name == "mol":
        spec, tau, n0 = mol_spec(8, 6, init_walkers=3000, seed=5), 0.02, 300.0
    elif name == "mol_unif":
        spec, tau, n0 = mol_spec(8, 6, particles=0, init_walkers=3000, seed=6), 0.02, 300.0
    else:
        spec, tau, n0 = mol_spec(8, 6, particles=0, trunc_initiator=False, init_walkers=3000, seed=7), 0.02, 300.0
    a, b = _pair(spec, tau, nshards, n0)
    shift = np.linspace(0.1, 0.3, nrep)
    for it in range(1, 31):
        x = a.iterate(it, shift)
        y = b.iterate(it, shift)
        assert x["rc"] == 0 and y["rc"] == 0
        for f in INT_FIELDS:
            np.testing.assert_array_equal(x[f][:nrep], y[f][:nrep], err_msg=f"{f} iteration {it}")
        for f in ("tot_walkers", "n_attempts", "n_spawned", "n_new"):
            assert x[f] == y[f], (f, it)
        np.testing.assert_allclose(x["enum_cyc"][:nrep], y["enum_cyc"][:nrep], rtol=1e-12, atol=1e-12)
    d1, s1, f1, _, _ = a.walkers()
    d2, s2, f2 = b.walkers()
    w1, w2 = sort_walkers(d1, s1, f1), sort_walkers(d2, s2, f2)
    for u, v in zip(w1, w2):
        np.testing.assert_array_equal(u, v)
    assert len(w1[0]) > 50
    a.close()
    b.close()


def test_sharded_with_core_space():
    spec = mol_spec(8, 6, init_walkers=3000, seed=8)
    tau = 0.02
    a, b = _pair(spec, tau, 4, 300.0)
    for it in range(1, 11):
        a.iterate(it, 0.2)
        b.iterate(it, 0.2)
    d, s, _, _, _ = a.walkers()
    core = d[np.argsort(-np.abs(s[:, 0]), kind="stable")[:40]]
    assert a.set_core(core) == b.set_core(core)
    for it in range(11, 31):
        x = a.iterate(it, 0.2)
        y = b.iterate(it, 0.2)
        np.testing.assert_allclose(x["tot_parts"][:1], y["tot_parts"][:1], rtol=1e-10)
    d1, s1, _, _, _ = a.walkers()
    d2, s2, _ = b.walkers()
    w1, w2 = sort_walkers(d1, s1), sort_walkers(d2, s2)
    np.testing.assert_array_equal(w1[0], w2[0])
    np.testing.assert_allclose(w1[1], w2[1], rtol=1e-9, atol=1e-9)
    a.close()
    b.close()
