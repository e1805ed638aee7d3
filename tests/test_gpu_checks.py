"""The device invariant checks (ngpu_set_checks): the walker table's
consistency after every iteration (slots, free queue, hash; the reference's
AnnihilatedDet == HolesInList, load_balancer.fpp:796-804) and pgen against
calc_pgen for one excitation of every occupied determinant (the ASSERT of
gasci_pchb_doubles_spinorb_fastweighted.fpp:191,231).

* Every generator, with and without a core space and through hash rebuilds:
  an engine with the checks on reports no violation and follows the same
  engine without them, iteration by iteration.
* A deliberate mismatch (NGPU_CHECK_SELFTEST scales the recomputed pgen by
  1 + 1e-6, from the host; every kernel still runs normally) halts the engine
  with NGPU_ERR_CHECK at the first iteration, and the counts show it.
"""
import numpy as np
import pytest

from neci_stable_amd._lib import CHECK_PGEN, CHECK_SELFTEST, CHECK_TABLE, NgpuError
from test_gpu_batch import same_stats
from test_gpu_parity import SPECS, engine

pytestmark = pytest.mark.gpu

CASES = [("hubbard10", 2, 0), ("mol8o6e_unif_fast", 1, 40), ("mol33o4e_w2", 2, 30), ("mol8o6e_uniform", 1, 0),
         ("mol8o6e_full_full", 1, 0)]


@pytest.mark.parametrize("name,nrep,ncore", CASES)
def test_checks_pass_and_leave_results_unchanged(name, nrep, ncore):
    spec = SPECS[name](init_walkers=4000, n_replicas=nrep, seed=61)
    tau = 0.01 if name.startswith("hub") else 0.02
    n0 = 40.0 if name.startswith("hub") else 400.0
    a, b = engine(spec), engine(spec)
    for e in (a, b):
        e.start_single_part(n0)
    shift = np.linspace(0.1, 0.2, nrep)
    for it in range(1, 11):
        same_stats(a.iterate(it, tau, shift), b.iterate(it, tau, shift), nrep, f"growth {it}")
    if ncore:
        core = a.most_populated(ncore)
        assert a.set_core(core) == b.set_core(core)
    a.set_checks(CHECK_TABLE | CHECK_PGEN)
    for it in range(11, 41):
        same_stats(a.iterate(it, tau, shift), b.iterate(it, tau, shift), nrep, f"iteration {it}", exact=not ncore)
    c = a.check_counts()
    print(name, c, a.hash_stats(), flush=True)
    assert c["iterations"] == 30 and c["pgen_tests"] > 30
    for k in ("bad_nel", "hash_misses", "hole_mismatches", "live_mismatches", "pgen_mismatches"):
        assert c[k] == 0, (k, c)
    assert a.error() == (0, 0)
    a.close()
    b.close()


def test_checks_through_hash_rebuilds():
    """110 buckets (440 entries) and the shift schedule of the hash-rebuild
    parity test (growth at 0.6, then -30): tombstones, in-place rebuilds, and
    the table checks hold after every iteration."""
    spec = SPECS["hubbard10"](init_walkers=4000, seed=5, n_hash_buckets=110)
    a, b = engine(spec), engine(spec)
    for e in (a, b):
        e.start_single_part(40.0)
    a.set_checks(CHECK_TABLE | CHECK_PGEN)
    for it in range(1, 41):
        shift = 0.6 if it <= 31 else -30.0
        same_stats(a.iterate(it, 0.01, shift), b.iterate(it, 0.01, shift), 1, f"iteration {it}")
    c = a.check_counts()
    assert a.hash_stats()["rebuilds"] >= 1
    assert c["iterations"] == 40 and c["pgen_tests"] > 40
    for k in ("bad_nel", "hash_misses", "hole_mismatches", "live_mismatches", "pgen_mismatches"):
        assert c[k] == 0, (k, c)
    a.close()
    b.close()


def test_check_failure_halts_with_its_status():
    spec = SPECS["mol8o6e_unif_fast"](init_walkers=4000, seed=62)
    e = engine(spec)
    e.start_single_part(400.0)
    for it in range(1, 6):
        e.iterate(it, 0.02, 0.1)
    e.set_checks(CHECK_TABLE | CHECK_PGEN | CHECK_SELFTEST)
    with pytest.raises(NgpuError) as err:
        e.iterate(6, 0.02, 0.1)
    assert err.value.code == 9
    assert e.error() == (9, 6)
    c = e.check_counts()
    assert c["iterations"] == 1 and c["pgen_mismatches"] > 0 and c["pgen_mismatches"] == c["pgen_tests"]
    assert c["hash_misses"] == 0 and c["hole_mismatches"] == 0
    e.close()
