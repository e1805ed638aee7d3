"""The oracle's most-populated core selection (generate_space_most_populated,
semi_stoch_gen.F90:1031-1226) against numpy (CPU only): the target
determinants of largest sum_r |c_r|, weights below 1e-8 left out, equal
weights in determinant (ilut_lt: signed words, word 0 first) order, output
weight descending."""
import numpy as np
import pytest

import oracle_ctypes as oc
from common import mol_spec


def numpy_most_populated(dets, signs, target):
    w = np.abs(signs).sum(axis=1)
    keep = w >= 1e-8
    d, w = dets[keep], w[keep]
    keys = [d[:, q] for q in range(d.shape[1] - 1, -1, -1)] + [-w]
    order = np.lexsort(keys)
    return d[order[:target]]


@pytest.mark.parametrize("nspat,nel,nrep,rank,mseed", [(8, 6, 2, 8, 7), (33, 4, 1, 4, 3)])
def test_oracle_most_populated_matches_numpy(nspat, nel, nrep, rank, mseed):
    spec = mol_spec(nspat, nel, rank=rank, mseed=mseed, init_walkers=5000, n_replicas=nrep)
    o = oc.Oracle(spec, 0.02)
    o.add_walkers(o.ref[None, :], np.full((1, nrep), 300.0), [(1 << nrep) - 1])
    for it in range(1, 25):
        o.iterate(it, np.full(nrep, 0.2))
    d, s, _, _, _ = o.walkers()
    assert len(d) > 200
    for target in (1, 50, len(d) + 10):
        got = o.most_populated(target)
        want = numpy_most_populated(d, s, target)
        np.testing.assert_array_equal(got, want)
    if spec.W == 2:  # word 0 compared as a signed integer: bit 63 set sorts first
        assert (d[:, 0] < 0).any()
    o.close()
