"""CPU restatement of READFCIINT (src/readint.F90:585-950), test infrastructure.

Only the tests use this module: it is the checker for the engine library's
FCIDUMP reader (ngpu_read_fcidump), never a product path.  Spin-restricted,
real integrals, free-format "Z I J K L" lines:
  * &FCI namelist: NORB, NELEC, MS2, ORBSYM, ISYM (readint.F90:614-636);
  * |Z| < UMatEps with I > 0: skipped (:727-741);
  * I = 0: ECORE (:818-826); J = 0: orbital energy, not stored (:827-833);
  * K = 0: TMAT2D(2I-s, 2J-s) = TMAT2D(2J-s, 2I-s) = Z for s = 1, 2 (:834-852);
  * otherwise UMAT(UMatInd(I, K, J, L)) = Z and NonZeroInt += 1 (:853-898).
UMatInd_base (UMatCache.F90:257-303) is restated here on its own.
"""
import re

import numpy as np


def _pair(a, b):
    return a * (a - 1) // 2 + b if a > b else b * (b - 1) // 2 + a


def umat_ind(i, j, k, l):
    return _pair(_pair(i, k), _pair(j, l))


def read_fcidump_ref(path, umat_eps=1e-8):
    text = open(path).read()
    m = re.search(r"&FCI(.*?)(&END|/)", text, re.S | re.I)
    if not m:
        raise ValueError("no &FCI namelist")
    nl = m.group(1)

    def ints(key):
        mm = re.search(r"(?<![A-Z0-9_])" + key + r"\s*=\s*([-0-9,\s*]+?)(?=[A-Z][A-Z0-9_]*\s*=|$)", nl, re.I | re.S)
        if not mm:
            return []
        out = []
        for tok in re.split(r"[,\s]+", mm.group(1).strip()):
            if not tok:
                continue
            if "*" in tok:
                r, v = tok.split("*")
                out += [int(v)] * int(r)
            else:
                out.append(int(tok))
        return out

    norb = ints("NORB")[0]
    hdr = {"norb": norb, "nelec": ints("NELEC")[0], "ms2": (ints("MS2") or [0])[0],
           "isym": (ints("ISYM") or [1])[0], "orbsym": ints("ORBSYM")}
    nb = 2 * norb
    npair = norb * (norb + 1) // 2
    tmat = np.zeros((nb, nb))
    umat = np.zeros(npair * (npair + 1) // 2)
    ecore = 0.0
    n_two = 0
    for line in text[m.end():].splitlines():
        f = line.split()
        if not f:
            continue
        z = float(f[0].replace("D", "E").replace("d", "e"))
        i, j, k, l = (int(x) for x in f[1:5])
        if abs(z) < umat_eps and i > 0:
            continue
        if i == 0:
            ecore = z
        elif j == 0:
            pass
        elif k == 0:
            for s in (1, 2):
                tmat[2 * i - s, 2 * j - s] = z
                tmat[2 * j - s, 2 * i - s] = z
        else:
            umat[umat_ind(i, k, j, l) - 1] = z
            n_two += 1
    return hdr, tmat, umat, ecore, n_two
