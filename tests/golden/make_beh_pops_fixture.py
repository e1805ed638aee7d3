"""Copies the reference test's binary POPSFILE pair (data) used as a format
fixture: test_suite/dneci/rdm/BeH_open_shell_dyn_core/POPSFILEHEAD and
POPSFILEBIN (NECI output, 588 determinants, two sign components).

Run in a container where the reference tree is present; the copies are
committed and the tests read only them.
"""
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "test_suite/dneci/rdm/BeH_open_shell_dyn_core")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "beh_pops")

if __name__ == "__main__":
    os.makedirs(OUT, exist_ok=True)
    for f in ("POPSFILEHEAD", "POPSFILEBIN"):
        shutil.copyfile(os.path.join(SRC, f), os.path.join(OUT, f))
        print("copied", f)
