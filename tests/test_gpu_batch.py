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


def test_batch_groups_share_launches():
    """Engines with the same kernel instances form one launch group (the two
    one-replica Hubbard chains; each molecular replica count its own group),
    and direct group launches with stage timing give the graph replay's
    results."""
    from neci_stable_amd.engine import Batch
    members = [("hubbard10", 1, 3), ("mol8o6e", 2, 4), ("hubbard14", 1, 5), ("mol8o6e", 1, 6)]
    alone, batched = [], []
    for name, nrep, seed in members:
        _, tau, a, b = engine_pair(name, nrep, seed)
        alone.append((a, tau, nrep))
        batched.append(b)
    batch = Batch(batched)
    g, n = batch.groups()
    assert n == 3 and g[0] == g[2] and len({int(g[0]), int(g[1]), int(g[3])}) == 3
    shifts = [np.linspace(0.2, 0.4, nrep) for _, _, nrep in alone]
    for it in range(1, 31):
        if it == 11:
            batch.set_graph(False)
            batch.set_profiling(True)
        if it == 21:  # the times are read before timing is switched off (which clears them)
            ms, nit = batch.group_stage_times(int(g[0]))
            assert nit == 10 and ms[2] > 0.0 and ms[6] > 0.0  # spawn and merge-annihilation stages timed
            batch.set_profiling(False)
            batch.set_graph(True)
        batch.iterate([it] * len(batched), [t for _, t, _ in alone], shifts)
        for i, (a, tau, nrep) in enumerate(alone):
            same_stats(a.iterate(it, tau, shifts[i]), batched[i].stats(), nrep, f"engine {i} iteration {it}")
    for i, (a, _, _) in enumerate(alone):
        assert same_walkers(a, batched[i]) > 20
        a.close()
    batch.close()
    for b in batched:
        b.close()


def test_batched_mix_of_molecules_and_chains():
    """A reduced form of the batched workload: two replicas (distinct seeds)
    of the 14e/30o molecular Hamiltonian (seed 7, rank 60; PCHB doubles with
    UNIF-UNIF particles, heat-bath singles, initiator 3, tau 0.002: about
    650 occupied determinants after the 15 growth iterations and 6500 at the
    end, in the oracle) with a 300-determinant
    core space, Hamiltonian and spawned amplitudes stored in FP32, plus
    half-filled Hubbard chains of 12, 13, 17 and 18 sites (colex-rank and
    word-key sorts).  Every batched problem follows the same problem iterated
    alone, iteration by iteration."""
    from neci_stable_amd import HubbardChain, Spec
    from neci_stable_amd.engine import Batch, Engine
    from common import molecule
    mol = molecule(30, 14, 60, 7)
    specs = [(Spec(mol, init_walkers=2e5, seed=700 + r, particles=0, psingles=0.1, memory_fac_spawn=6.0), 0.002,
              200.0) for r in range(2)]
    for L, u in ((12, 3.0), (13, 5.5), (17, 2.5), (18, 7.0)):
        specs.append((Spec(HubbardChain(L, L, u, bhub=-1.0), init_walkers=2e4, seed=800 + L), 0.01, 30.0))
    pairs = []
    for spec, tau, n0 in specs:
        a, b = Engine(spec), Engine(spec)
        a.start_single_part(n0)
        b.start_single_part(n0)
        pairs.append([a, b, tau])
    for it in range(1, 16):  # grow both twins alone
        for a, b, tau in pairs:
            same_stats(a.iterate(it, tau, 0.0), b.iterate(it, tau, 0.0), 1, f"growth {it}")
    for k in range(2):       # core space and FP32 storage for the molecules
        a, b, _ = pairs[k]
        d, sg, _, _, _ = a.download()
        core = np.ascontiguousarray(d[np.sort(np.argsort(-np.abs(sg[:, 0]), kind="stable")[:300])])
        assert a.set_core(core, fp32=True) == b.set_core(core, fp32=True)
        a.set_spawn_precision(True)
        b.set_spawn_precision(True)
    batch = Batch([b for _, b, _ in pairs])
    g, n = batch.groups()
    assert g[0] == g[1] and n >= 3
    print(f"launch groups {list(g)}", flush=True)
    for it in range(16, 46):
        batch.iterate([it] * len(pairs), [t for _, _, t in pairs], [0.0] * len(pairs))
        for i, (a, b, tau) in enumerate(pairs):
            same_stats(a.iterate(it, tau, 0.0), b.stats(), 1, f"problem {i} iteration {it}", exact=i >= 2)
    for i, (a, b, _) in enumerate(pairs):
        w1 = sort_walkers(*a.download()[:2])
        w2 = sort_walkers(*b.download()[:2])
        np.testing.assert_array_equal(w1[0], w2[0])
        np.testing.assert_allclose(w1[1], w2[1], rtol=1e-9, atol=1e-9)
        assert len(w1[0]) > 100
        a.close()
    batch.close()
    for _, b, _ in pairs:
        b.close()


def test_batch_shared_core_space():
    """Replicas of one system with the same core determinants (two engines of
    two replicas and one of three, one launch group) share one core
    projection that reads the core Hamiltonian once (GROUP_SHARED_CORE);
    every engine follows the same engine iterated alone (to 1e-10: the shared
    projection sums in another order).  When one engine's core differs, the
    group projects each engine on its own, with the same results."""
    from neci_stable_amd._lib import GROUP_CORE, GROUP_SHARED_CORE
    from neci_stable_amd.engine import Batch, Engine
    mk, tau, n0 = CASES["mol8o6e"]
    g0 = Engine(mk(init_walkers=4000, n_replicas=1, seed=90))
    g0.start_single_part(n0)
    for it in range(1, 16):
        g0.iterate(it, tau, 0.0)
    core = g0.most_populated(60)
    g0.close()
    assert len(core) == 60
    for differ in (False, True):
        alone, batched, shifts = [], [], []
        for i, nrep in enumerate((2, 2, 3)):
            spec = mk(init_walkers=4000, n_replicas=nrep, seed=91 + i)
            pair = []
            for _ in range(2):
                e = Engine(spec)
                e.start_single_part(n0)
                e.set_core(core[:-1] if (differ and i == 1) else core, fp32=(i == 0))
                pair.append(e)
            alone.append(pair[0])
            batched.append(pair[1])
            shifts.append(np.linspace(0.1, 0.3, nrep))
        # the first engine stores its core Hamiltonian in FP32: another group
        batch = Batch(batched)
        g, n = batch.groups()
        assert n == 2 and g[1] == g[2] and g[0] != g[1]
        flags = batch.group_flags(int(g[1]))
        assert flags & GROUP_CORE
        assert bool(flags & GROUP_SHARED_CORE) == (not differ)
        for it in range(1, 26):
            batch.iterate([it] * 3, [tau] * 3, shifts)
            for i, a in enumerate(alone):
                same_stats(a.iterate(it, tau, shifts[i]), batched[i].stats(), a.S, f"engine {i} iteration {it}",
                           exact=False)
        for a, b in zip(alone, batched):
            w1 = sort_walkers(*a.download()[:3])
            w2 = sort_walkers(*b.download()[:3])
            np.testing.assert_array_equal(w1[0], w2[0])
            np.testing.assert_array_equal(w1[2], w2[2])
            np.testing.assert_allclose(w1[1], w2[1], rtol=1e-9, atol=1e-9)
            assert len(w1[0]) > 20
            a.close()
        batch.close()
        for b in batched:
            b.close()


def test_batch_launch_shape_changes_keep_results():
    """The launch shape of batched engines changed between iterations (the
    batch regroups and recaptures; each knob of the autotuner in turn): every
    batched problem still follows the same problem iterated alone with the
    default shape.  Two replicas of the 14e/30o molecule (64-bit rank sort,
    FP32 spawns, a 300-determinant core) and two Hubbard chains."""
    from neci_stable_amd import HubbardChain, Spec
    from neci_stable_amd.engine import Batch, Engine
    from common import molecule
    mol = molecule(30, 14, 60, 7)
    specs = [(Spec(mol, init_walkers=2e5, seed=710 + r, particles=0, psingles=0.1, memory_fac_spawn=6.0), 0.002,
              200.0) for r in range(2)]
    for L, u in ((12, 3.0), (17, 2.5)):
        specs.append((Spec(HubbardChain(L, L, u, bhub=-1.0), init_walkers=2e4, seed=820 + L), 0.01, 30.0))
    pairs = []
    for spec, tau, n0 in specs:
        a, b = Engine(spec), Engine(spec)
        a.start_single_part(n0)
        b.start_single_part(n0)
        pairs.append([a, b, tau])
    for it in range(1, 16):
        for a, b, tau in pairs:
            same_stats(a.iterate(it, tau, 0.0), b.iterate(it, tau, 0.0), 1, f"growth {it}")
    d, sg, _, _, _ = pairs[0][0].download()
    core = np.ascontiguousarray(d[np.sort(np.argsort(-np.abs(sg[:, 0]), kind="stable")[:300])])
    for k in range(2):
        a, b, _ = pairs[k]
        assert a.set_core(core, fp32=True) == b.set_core(core, fp32=True)
        a.set_spawn_precision(True)
        b.set_spawn_precision(True)
    batch = Batch([b for _, b, _ in pairs])
    cu = pairs[0][1].launch_shape()["n_cu"]
    shapes = [dict(radix_items=2), dict(radix_items=8), dict(chunk_blocks=2 * cu), dict(chunk_blocks=8 * cu),
              dict(grid_blocks=3 * cu), dict(grid_blocks=8 * cu), dict(single_lanes=8), dict(single_lanes=16),
              dict(merge_blocks=2 * cu), dict(merge_blocks=4 * cu)]
    it = 15
    for shape in shapes:
        print(f"shape {shape}", flush=True)
        for _, b, _ in pairs:
            b.set_launch_shape(**shape)
        for _ in range(3):
            it += 1
            batch.iterate([it] * len(pairs), [t for _, _, t in pairs], [0.0] * len(pairs))
            for i, (a, b, tau) in enumerate(pairs):
                same_stats(a.iterate(it, tau, 0.0), b.stats(), 1, f"problem {i} iteration {it} {shape}",
                           exact=i >= 2)
    for i, (a, b, _) in enumerate(pairs):
        w1 = sort_walkers(*a.download()[:2])
        w2 = sort_walkers(*b.download()[:2])
        np.testing.assert_array_equal(w1[0], w2[0])
        np.testing.assert_allclose(w1[1], w2[1], rtol=1e-9, atol=1e-9)
        a.close()
    batch.close()
    for _, b, _ in pairs:
        b.close()
