"""The spawn sort at sizes where its per-segment sort leaves LDS.

The colex-rank sorts make one stable pass on the top 9 bits of rank*odd
constant (32-bit keys, or 64-bit keys modulo 2^bits), then sort each of the
512 segments by the bytes below: in LDS up to 2048 keys, through global
memory beyond.  Through the
explicit-list entry point (ngpu_annihilate, the same sort and merge as an
iteration) on an empty walker list:
* 1.5e6 spawns over 3e5 determinants of 10 electrons in 40 spin orbitals
  (segments of ~2900 keys, all past the LDS size);
* a skewed list in which one determinant takes 1e5 of 3e5 spawns (its
  segment alone passes the LDS size, the others stay inside it);
* the same two lists, and one that stays inside LDS, over determinants of
  14 electrons in 60 spin orbitals (the 64-bit rank sort).
The merged list must equal a numpy group-by of the spawns (integer values,
so every sum is exact) and the oracle's merge of the same list; initiator
flags and aborts follow from the group-by as in the oracle pinning test.
"""
import numpy as np
import pytest

from common import mol_spec, sort_walkers
from neci_stable_amd.systems import encode_det

pytestmark = pytest.mark.gpu


def _random_dets(rng, n, nspat=20, nel=10):
    out = np.zeros((n, 1), dtype=np.int64)
    alpha = np.arange(2, 2 * nspat + 1, 2)
    beta = np.arange(1, 2 * nspat + 1, 2)
    for i in range(n):
        orbs = list(rng.choice(alpha, nel // 2, replace=False)) + list(rng.choice(beta, nel // 2, replace=False))
        out[i] = encode_det(sorted(int(o) for o in orbs), 2 * nspat)
    return out


def _check(eng, orc, sdet, sval, sinfo):
    a, md, ms, mf = eng.annihilate(3, sdet, sval, sinfo)
    b, md2, ms2, mf2 = orc.annihilate(3, sdet, sval, sinfo)
    assert b["rc"] == 0
    m1 = sort_walkers(md, ms, mf)
    m2 = sort_walkers(md2, ms2, mf2)
    for x, y in zip(m1, m2):
        np.testing.assert_array_equal(x, y)
    # group-by of the spawns: merged sums over nonzero totals
    keys, inv = np.unique(sdet[:, 0], return_inverse=True)
    sums = np.zeros(len(keys))
    np.add.at(sums, inv.reshape(-1), sval)
    keep = sums != 0
    assert len(m1[0]) == int(keep.sum())
    got = dict(zip(m1[0][:, 0].tolist(), m1[1][:, 0].tolist()))
    for k, v in zip(keys[keep].tolist(), sums[keep].tolist()):
        assert got[k] == v
    assert a["n_unique"] == b["n_unique"] == int(keep.sum())
    return len(m1[0])


@pytest.mark.parametrize("case,nspat,nel", [("large", 20, 10), ("skewed", 20, 10), ("large", 30, 14),
                                            ("skewed", 30, 14), ("lds", 30, 14)])
def test_partition_sort_segments_past_lds(case, nspat, nel):
    """10e/20o: the 32-bit colex-rank sort; 14e/30o: the 64-bit one (46-bit
    keys, five byte passes per segment; "lds": 4e5 spawns, ~780 per segment,
    every segment inside LDS)."""
    from neci_stable_amd.engine import Engine
    from oracle_ctypes import Oracle
    rng = np.random.default_rng(17)
    if case == "large":
        pool = np.unique(_random_dets(rng, 300_000, nspat, nel), axis=0)
        sdet = pool[rng.integers(0, len(pool), 1_500_000)]
    elif case == "lds":
        pool = np.unique(_random_dets(rng, 200_000, nspat, nel), axis=0)
        sdet = pool[rng.integers(0, len(pool), 400_000)]
    else:
        pool = np.unique(_random_dets(rng, 100_000, nspat, nel), axis=0)
        sdet = pool[rng.integers(0, len(pool), 300_000)]
        sdet[rng.permutation(len(sdet))[:100_000]] = pool[0]
    n = len(sdet)
    sval = rng.integers(-3, 4, n).astype(float)
    sval[sval == 0] = 1.0
    sinfo = ((rng.random(n) < 0.3).astype(np.uint32) << 8)
    spec = mol_spec(nspat, nel, rank=8, init_walkers=n, memory_fac_spawn=1.2, memory_fac_part=1.2,
                    memory_fac_hash=1.5)
    eng = Engine(spec)
    orc = Oracle(spec, 0.01)
    nm = _check(eng, orc, sdet, sval, sinfo)
    print(f"{case}: {n} spawns -> {nm} merged determinants", flush=True)
    eng.close()
