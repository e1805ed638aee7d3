"""The launch shape (ngpu_set_launch_shape) changes how work is split over
blocks, never the results: engines with different shapes, started from the
same walkers, keep identical statistics and walker sets."""
import numpy as np
import pytest

from common import mol_spec, sort_walkers

pytestmark = pytest.mark.gpu


def test_launch_shape_does_not_change_results():
    from neci_stable_amd.engine import Engine
    spec = mol_spec(8, 6, init_walkers=4000, seed=12)
    shapes = [dict(chunk_blocks=0, grid_blocks=0, radix_items=4),
              dict(chunk_blocks=64, grid_blocks=64, radix_items=2, single_lanes=16, merge_blocks=512),
              dict(chunk_blocks=256, grid_blocks=4096, radix_items=8, single_lanes=8, merge_blocks=64)]
    engines = []
    for sh in shapes:
        e = Engine(spec)
        e.set_launch_shape(**sh)
        e.start_single_part(400.0)
        engines.append(e)
    assert engines[1].launch_shape()["radix_items"] == 2 and engines[1].launch_shape()["single_lanes"] == 16
    for it in range(1, 41):
        st = [e.iterate(it, 0.02, 0.3) for e in engines]
        for f in ("nborn", "ndied", "nannihil", "naborted", "tot_parts", "no_init_dets"):
            for o in st[1:]:
                np.testing.assert_array_equal(st[0][f], o[f], err_msg=f"{f} iteration {it}")
        for o in st[1:]:
            assert st[0]["tot_walkers"] == o["tot_walkers"] and st[0]["n_spawned"] == o["n_spawned"]
    ref = sort_walkers(*engines[0].download()[:3])
    assert len(ref[0]) > 50
    for e in engines[1:]:
        for x, y in zip(ref, sort_walkers(*e.download()[:3])):
            np.testing.assert_array_equal(x, y)
    for e in engines:
        e.close()


def test_launch_shape_rejects_out_of_range():
    from neci_stable_amd._lib import NgpuError
    from neci_stable_amd.engine import Engine
    e = Engine(mol_spec(8, 6, init_walkers=1000))
    for bad in (dict(radix_items=3), dict(chunk_blocks=10), dict(grid_blocks=-1), dict(chunk_blocks=10 ** 7),
                dict(single_lanes=3), dict(merge_blocks=10)):
        with pytest.raises(NgpuError):
            e.set_launch_shape(**bad)
    e.close()
