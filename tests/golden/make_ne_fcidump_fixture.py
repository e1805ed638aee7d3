"""Writes the Ne FCIDUMP fixture: the reference test's integral file (data)
and the values its benchmark output records for it.

Source: test_suite/neci/parallel/Ne_FciMCPar_pchb/ (FCIDUMP, benchmark.out*).
Run in a container where the reference tree is present; the outputs are
committed and the tests read only them.
"""
import glob
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "test_suite/neci/parallel/Ne_FciMCPar_pchb")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ne_fcidump")


def main():
    os.makedirs(OUT, exist_ok=True)
    shutil.copyfile(os.path.join(SRC, "FCIDUMP"), os.path.join(OUT, "FCIDUMP"))
    bench = open(glob.glob(os.path.join(SRC, "benchmark.out*"))[0]).read()

    def num(pat):
        return float(re.search(pat, bench).group(1))

    d0 = re.search(r"Fermi det \(D0\):\(([\d,\s]+)\)", bench).group(1)
    exp = {
        "source": "test_suite/neci/parallel/Ne_FciMCPar_pchb/benchmark.out (NECI output for this FCIDUMP)",
        "n_two": int(num(r"Number of non-zero integrals:\s+(\d+)")),
        "ecore": num(r"ECORE=\s+([-\d.Ee+]+)"),
        "d0_orbs": [int(x) for x in d0.replace(",", " ").split()],
        "e_d0": num(r"<D0\|H\|D0>=\s+([-\d.Ee+]+)"),
    }
    with open(os.path.join(OUT, "expected.json"), "w") as f:
        json.dump(exp, f, indent=1)
    print(json.dumps(exp))


if __name__ == "__main__":
    main()
