"""The Fortran host (ISO_C_BINDING module + FciMCPar-style driver) over the GPU engine."""
import os
import subprocess

import pytest

import ed_reference as ed
from neci_stable_amd import HubbardChain

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fortran", "neci_gpu_driver")


@pytest.mark.skipif(not os.path.exists(EXE), reason="amdflang driver not built")
def test_fortran_driver_hubbard6_energy():
    r = subprocess.run([EXE, "6", "6", "4.0", "0.02", "3000", "3000", "7", "0"], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    line = [x for x in r.stdout.splitlines() if x.startswith("mean shift energy")][-1]
    e = float(line.split(":")[1])
    sysm = HubbardChain(6, 6, 4.0, bhub=-1.0)
    e_ed = ed.ground_state(ed.hubbard_hamiltonian(6, sysm.nbr, 4.0, -1.0, ed.hubbard_dets(6, 3, 3)))
    assert abs(e - e_ed) < 0.1, (e, e_ed)
