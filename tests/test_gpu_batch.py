"""Captured-graph replay and batches of engines.

* An engine replaying its captured iteration graph (ngpu_set_graph) gives the
  same statistics and walkers as issuing the launches one by one, across a
  recapture after ngpu_set_core.
* A batch of engines of different systems and replica counts, iterated as one
  graph per iteration (ngpu_batch_iterate), gives each engine the statistics
  and walkers of iterating it alone.  Every replay is checked, with the
  host's own synchronising calls (statistics copies) between replays.
"""
import numpy as np
import pytest

from common import hubbard_spec, mol_spec, sort_walkers

pytestmark = pytest.mark.gpu

FIELDS = ["nborn", "ndied", "nannihil", "naborted", "nremoved", "tot_parts", "norm_sq", "hf_cyc",
          "noat_doubs", "spawn_from_sing", "inst_no_at_hf", "no_init_dets", "no_init_walk"]


def same_stats(a, b, S, where, exact=True):
    """exact=False once real core amplitudes enter sums whose order differs
    between launches (block partials): 1e-10 relative then."""
    for f in FIELDS:
        if exact:
            np.testing.assert_array_equal(a[f][:S], b[f][:S], err_msg=f"{f} {where}")
        else:
            np.testing.assert_allclose(a[f][:S], b[f][:S], rtol=1e-10, atol=1e-10, err_msg=f"{f} {where}")
    np.testing.assert_allclose(a["enum_cyc"][:S], b["enum_cyc"][:S], rtol=1e-12, atol=1e-12)
    for f in ("tot_walkers", "n_attempts", "n_spawned", "n_unique", "n_new"):
        assert a[f] == b[f], (f, where)


def same_walkers(e1, e2):
    w1 = sort_walkers(*e1.download()[:3])
    w2 = sort_walkers(*e2.download()[:3])
    for x, y in zip(w1, w2):
        np.testing.assert_array_equal(x, y)
    return len(w1[0])


CASES = {
    "hubbard10": (lambda **k: hubbard_spec(10, 10, 4.0, **k), 0.01, 30.0),
    "hubbard14": (lambda **k: hubbard_spec(14, 14, 2.0, **k), 0.01, 30.0),
    "mol8o6e": (lambda **k: mol_spec(8, 6, **k), 0.02, 400.0),
}


def engine_pair(name, nrep, seed):
    from neci_stable_amd.engine import Engine
    mk, tau, n0 = CASES[name]
    spec = mk(init_walkers=4000, n_replicas=nrep, seed=seed)
    out = []
    for _ in range(2):
        e = Engine(spec)
        e.start_single_part(n0)
        out.append(e)
    return spec, tau, out[0], out[1]


@pytest.mark.parametrize("name,nrep", [("hubbard10", 1), ("mol8o6e", 2)])
def test_graph_replay_matches_launches(name, nrep):
    spec, tau, a, b = engine_pair(name, nrep, 5)
    b.set_graph(True)
    shift = np.linspace(0.2, 0.4, nrep)
    for it in range(1, 41):
        if it == 21:  # the chain changes: the graph is captured again
            d, _, _, _, _ = a.download()
            core = d[:40]
            assert a.set_core(core) == b.set_core(core)
        same_stats(a.iterate(it, tau, shift), b.iterate(it, tau, shift), nrep, f"iteration {it}", exact=it < 21)
    assert same_walkers(a, b) > 50
    a.close()
    b.close()


def test_batch_matches_engines_alone():
    from neci_stable_amd.engine import Batch
    members = [("hubbard10", 1, 3), ("mol8o6e", 2, 4), ("hubbard14", 1, 5), ("mol8o6e", 1, 6)]
    alone, batched, specs = [], [], []
    for name, nrep, seed in members:
        spec, tau, a, b = engine_pair(name, nrep, seed)
        alone.append((a, tau, nrep))
        batched.append(b)
        specs.append(spec)
    batch = Batch(batched)
    shifts = [np.linspace(0.2, 0.4, nrep) for _, _, nrep in alone]
    for it in range(1, 31):
        batch.iterate([it] * len(batched), [t for _, t, _ in alone], shifts)
        for i, (a, tau, nrep) in enumerate(alone):
            same_stats(a.iterate(it, tau, shifts[i]), batched[i].stats(), nrep, f"engine {i} iteration {it}")
        if it % 10 == 0:  # cycle statistics of batched engines, as the host driver reads them
            for i, (a, tau, nrep) in enumerate(alone):
                sb, lb, nb = batched[i].cycle_stats()
                sa, la, na = a.cycle_stats()
                assert na == nb == 10
                same_stats(sa, sb, nrep, f"cycle of engine {i} at {it}")
    for i, (a, _, _) in enumerate(alone):
        assert same_walkers(a, batched[i]) > 20
    batch.close()
    # back to independent use
    for i, (a, tau, nrep) in enumerate(alone):
        for it in range(31, 36):
            same_stats(a.iterate(it, tau, shifts[i]), batched[i].iterate(it, tau, shifts[i]), nrep,
                       f"engine {i} after the batch, iteration {it}")
        a.close()
        batched[i].close()


def test_batched_engine_refuses_own_iterate():
    from neci_stable_amd._lib import NgpuError
    from neci_stable_amd.engine import Batch
    spec, tau, a, b = engine_pair("hubbard10", 1, 7)
    batch = Batch([a, b])
    with pytest.raises(NgpuError) as e:
        a.iterate(1, tau, 0.3)
    assert e.value.code == 7
    batch.close()
    a.iterate(1, tau, 0.3)
    a.close()
    b.close()
