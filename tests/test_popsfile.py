"""POPSFILE v4 reader/writer (neci_stable_amd/popsfile.py) against a POPSFILE
the reference wrote (tests/golden/beh_pops, copied by
tests/golden/make_beh_pops_fixture.py from the reference test
BeH_open_shell_dyn_core: 588 determinants, lenof_sign 2, 3 electrons)."""
import os

import numpy as np

from neci_stable_amd.popsfile import read_popsfile, write_popsfile

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "beh_pops")


def test_reference_binary_popsfile_reads_and_rewrites_byte_for_byte(tmp_path):
    h, rows = read_popsfile(os.path.join(GOLD, "POPSFILE"))
    assert h["PopNEl"] == 3 and h["PopLensign"] == 2 and h["PopNIfD"] == 0
    W, S = 1, 2
    # every record is an occupied determinant of PopNEl electrons; the count
    # is the header's PopTotwalk and PopWalkersOnNodes
    assert len(rows) == h["PopTotwalk"] == h["PopWalkersOnNodes"] == 588
    pop = [bin(int(x) & (2 ** 64 - 1)).count("1") for x in rows[:, 0]]
    assert set(pop) == {3}
    sg = rows[:, W:W + S].copy().view(np.float64)
    assert np.all(np.isfinite(sg)) and np.all(np.abs(sg).sum(axis=1) > 0)
    assert len(np.unique(rows[:, 0])) == len(rows)
    # writing the rows back reproduces the reference's record stream exactly
    stem = str(tmp_path / "POPSFILE")
    n = write_popsfile(stem, rows, W, S, 3, PopSft=h["PopSft"], PopSumNoatHF=h["PopSumNoatHF"],
                       PopSumENum=h["PopSumENum"], PopCyc=h["PopCyc"], PopTau=h["PopTau"],
                       PopRandomHash=h["PopRandomHash"])
    assert n == 588
    assert open(stem + "BIN", "rb").read() == open(os.path.join(GOLD, "POPSFILEBIN"), "rb").read()
    h2, rows2 = read_popsfile(stem)
    np.testing.assert_array_equal(rows2, rows)
    for k in ("PopNEl", "PopLensign", "PopTotwalk", "PopCyc", "PopNIfD", "PopNIfSgn", "PopNIfTot",
              "PopRandomHash"):
        assert h2[k] == h[k], k
    for k in ("PopSft", "PopTau", "PopSumENum"):
        assert abs(h2[k] - h[k]) <= 1e-12 * max(1.0, abs(h[k])), k


def test_text_popsfile_round_trip(tmp_path):
    """Text POPSFILE (i24 det words, f30.8 signs, i24 flags): integer signs
    and flags come back exactly; all-zero rows are not written."""
    rng = np.random.default_rng(4)
    W, S, n = 2, 3, 300
    rows = np.zeros((n, W + S + 1), dtype=np.int64)
    rows[:, :W] = rng.integers(-2 ** 62, 2 ** 62, (n, W))
    sg = rng.integers(-20, 21, (n, S)).astype(np.float64)
    sg[:10] = 0.0
    rows[:, W:W + S] = sg.view(np.int64)
    rows[:, W + S] = rng.integers(0, 2 ** 20, n)
    stem = str(tmp_path / "POPSFILE")
    assert write_popsfile(stem, rows, W, S, 10, binary=False, PopSft=[0.1, 0.2, 0.3],
                          PopSumNoatHF=[1.0, 2.0, 3.0], PopSumENum=[4.0, 5.0, 6.0]) == n - 10
    h, back = read_popsfile(stem)
    np.testing.assert_array_equal(back, rows[10:])
    assert h["PopMultiSft"] == [0.1, 0.2, 0.3]
