"""FCIDUMP reader of the engine library (ngpu_read_fcidump, host code) against
the CPU restatement of READFCIINT, the reference's own Ne integral file and
the energy NECI printed for it.

Golden data: tests/golden/ne_fcidump/ holds the FCIDUMP of the reference test
Ne_FciMCPar_pchb and the values its benchmark output records (number of
two-electron integrals stored, ECORE, the Fermi determinant and <D0|H|D0>),
written by tests/golden/make_ne_fcidump_fixture.py.
"""
import json
import os

import numpy as np
import pytest

from fcidump_ref import read_fcidump_ref
from neci_stable_amd import Spec, synthetic_molecule
from neci_stable_amd._lib import NgpuError
from neci_stable_amd.systems import read_fcidump, write_fcidump

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ne_fcidump")


def test_ne_fcidump_matches_restatement_and_reference_output():
    path = os.path.join(GOLD, "FCIDUMP")
    exp = json.load(open(os.path.join(GOLD, "expected.json")))
    hdr, tmat, umat, ecore, n_two = read_fcidump_ref(path)
    mol = read_fcidump(path)
    assert (mol.nspat, mol.nel, mol.ms2, mol.isym) == (hdr["norb"], hdr["nelec"], hdr["ms2"], hdr["isym"])
    np.testing.assert_array_equal(mol.orbsym, hdr["orbsym"])
    np.testing.assert_array_equal(mol.tmat, tmat)
    np.testing.assert_array_equal(mol.umat, umat)
    assert mol.ecore == ecore == exp["ecore"]
    # "Number of non-zero integrals" in the reference's output for this file
    assert mol.n_two == n_two == exp["n_two"]
    # <D0|H|D0> of the reference's Fermi determinant (sltcnd_0 in the oracle);
    # the printed value has 14 decimals
    from oracle_ctypes import Oracle
    orc = Oracle(Spec(mol, init_walkers=100, ref_orbs=exp["d0_orbs"]), 0.01)
    assert abs(orc.diag(exp["d0_orbs"]) - exp["e_d0"]) < 1e-11


def test_synthetic_round_trip_is_exact(tmp_path):
    """write_fcidump (17 significant digits) then the library reader: the
    arrays of the synthetic molecule come back bit for bit."""
    mol = synthetic_molecule(9, 6, 12, 4)
    p = str(tmp_path / "FCIDUMP")
    write_fcidump(p, mol, thresh=0.0)
    back = read_fcidump(p, umat_eps=0.0)
    np.testing.assert_array_equal(back.tmat, mol.tmat)
    np.testing.assert_array_equal(back.umat, mol.umat)
    assert back.ecore == mol.ecore


def _write(path, header, lines):
    with open(path, "w") as f:
        f.write(header)
        for z, i, j, k, l in lines:
            f.write(f" {z} {i} {j} {k} {l}\n")


def test_format_variants(tmp_path):
    """Fortran D exponents, a '/' namelist end, ORBSYM repeat counts and the
    UMatEps screen (tiny two-electron values skipped, the core energy kept)."""
    p = str(tmp_path / "FCIDUMP")
    lines = [("0.5D+00", 1, 1, 1, 1), ("2.5d-1", 2, 1, 2, 1), ("1.0E-09", 2, 2, 1, 1),
             ("-1.25D0", 1, 1, 0, 0), ("0.1", 2, 1, 0, 0), ("-0.75", 2, 2, 0, 0),
             ("3.0", 1, 0, 0, 0), ("1.0D-12", 0, 0, 0, 0)]
    _write(p, "&FCI NORB=2, NELEC=2, MS2=0, ORBSYM=2*1, ISYM=1 /\n", lines)
    mol = read_fcidump(p)
    hdr, tmat, umat, ecore, n_two = read_fcidump_ref(p)
    np.testing.assert_array_equal(mol.orbsym, [1, 1])
    np.testing.assert_array_equal(mol.tmat, tmat)
    np.testing.assert_array_equal(mol.umat, umat)
    assert n_two == mol.n_two == 2
    assert mol.ecore == 1.0e-12
    assert mol.tmat[0, 0] == mol.tmat[1, 1] == -1.25
    assert mol.tmat[2, 0] == mol.tmat[0, 2] == mol.tmat[3, 1] == 0.1
    assert mol.tmat[0, 1] == 0.0  # no coupling between spins


@pytest.mark.parametrize("case", ["uhf", "tmat_conflict", "missing", "bad_index"])
def test_refused_inputs(tmp_path, case):
    p = str(tmp_path / "FCIDUMP")
    head = "&FCI NORB=2, NELEC=2, MS2=0, ORBSYM=1,1, ISYM=1 &END\n"
    if case == "uhf":
        _write(p, "&FCI NORB=2, NELEC=2, MS2=0, ORBSYM=1,1, ISYM=1, UHF=.TRUE. &END\n",
               [("0.5", 1, 1, 1, 1)])
    elif case == "tmat_conflict":  # readint.F90:838-845 stops on two different values
        _write(p, head, [("0.5", 2, 1, 0, 0), ("0.6", 1, 2, 0, 0)])
    elif case == "bad_index":
        _write(p, head, [("0.5", 3, 1, 1, 1)])
    else:
        p = str(tmp_path / "absent")
    with pytest.raises(NgpuError) as e:
        read_fcidump(p)
    assert e.value.code == 4
