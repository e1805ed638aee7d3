"""POPSFILE version 4: checkpoint and restart of the walker list.

The engine's side of checkpointing is ngpu_download_walkers (CurrentDets rows:
det words | signs as FP64 bit patterns | flag word) and ngpu_upload_walkers.
This module writes and reads those rows in the reference's POPSFILE v4 format,
so a run can restart from a POPSFILE of either program:

* header (`write_popsfile_header`, Popsfile.F90:1951-2052): '# POPSFILE VERSION 4',
  then the POPSHEAD namelist -- Pop64Bit, PopHPHF, PopLz, PopLensign, PopNEl,
  PopTotwalk, PopSft (or PopMultiSft/PopMultiSumNoatHF/PopMultiSumENum with
  several replicas), PopSumNoatHF, PopSumENum, PopCyc, PopNIfD, PopNIfY,
  PopNIfSgn, PopNIfFlag, PopNIfTot, PopTau, PopiBlockingIter, PopPSingles,
  PopPParallel, ..., PopTotImagTime, PopNNodes, PopWalkersOnNodes,
  PopRandomHash -- closed by '&END';
* binary body (POPSFILEBIN next to POPSFILEHEAD, `write_pops_det`
  Popsfile.F90:2054-2110): one Fortran sequential unformatted record per
  occupied determinant, det(0:NIfD) int64, real_sgn(lenof_sign) FP64, flag as
  int64, each record framed by its 4-byte length before and after;
* text body (in the POPSFILE itself, after the header): per determinant the
  det words as i24, the signs as f30.8 and the flag as i24.

Reading follows `read_popsfile_det` (Popsfile.F90:662-765): records whose
signs are all zero are skipped.
"""
import os
import re
import struct

import numpy as np

__all__ = ["write_popsfile", "read_popsfile", "PopsHeader"]


class PopsHeader(dict):
    """POPSHEAD values by name (PopNEl, PopLensign, PopTotwalk, PopSft, ...)."""


def _fmt_l(b):
    return "T" if b else "F"


def _f18_12(x):
    return f"{x:18.12f}"


def _g18_12(x):
    s = f"{x:18.12G}"
    return s if len(s) <= 18 else f"{x:18.11E}"


def _list_directed(x):
    return f"{x: .16E}"


def _header_text(h):
    """write_popsfile_header, Popsfile.F90:1951-2052 (including the format
    reversions that start some lines with a comma)."""
    S = int(h["PopLensign"])
    L = []
    L.append("# POPSFILE VERSION 4")
    L.append("&POPSHEAD Pop64Bit=T")
    L.append(f"PopHPHF={_fmt_l(h.get('PopHPHF', False))},PopLz={_fmt_l(h.get('PopLz', False))},"
             f"PopLensign={S:2d},PopNEl={int(h['PopNEl']):4d},")
    L.append(f"PopTotwalk={int(h['PopTotwalk']):15d}")
    sft = np.atleast_1d(np.asarray(h.get("PopSft", 0.0), dtype=float))
    noat = np.atleast_1d(np.asarray(h.get("PopSumNoatHF", np.zeros(S)), dtype=float))
    enum = np.atleast_1d(np.asarray(h.get("PopSumENum", 0.0), dtype=float))
    if len(sft) == 1:
        L.append(" PopSft= " + _list_directed(sft[0]))
        L.append(" PopSumNoatHF= " + " ".join(_list_directed(x) for x in noat))
        L.append(" PopSumENum= " + _list_directed(enum[0]))
    else:
        L.append(" PopMultiSft= " + " ".join(_list_directed(x) for x in sft))
        L.append(" PopMultiSumNoatHF= " + " ".join(_list_directed(x) for x in noat))
        L.append(" PopMultiSumENum= " + " ".join(_list_directed(x) for x in enum))
    L.append(f"PopCyc={int(h.get('PopCyc', 0)):16d},PopNIfD={int(h['PopNIfD']):2d},PopNifY={0:2d},"
             f"PopNIfSgn={S:2d},")
    L.append(f"PopNIfFlag={1:2d},PopNIfTot={int(h['PopNIfD']) + S + 1:2d},"
             f"PopTau={_f18_12(float(h.get('PopTau', 0.0)))},")
    L.append(f"PopiBlockingIter={int(h.get('PopiBlockingIter', 0)):16d}")
    L.append(f"PopPSingles={_f18_12(float(h.get('PopPSingles', 0.0)))},"
             f"PopPParallel={_f18_12(float(h.get('PopPParallel', 0.0)))}")
    L.append(f",PopPDoubles={_f18_12(float(h.get('PopPDoubles', 0.0)))},"
             f"PopPTriples={_f18_12(0.0)}")
    L.append(f",PopPSing_spindiff1={_f18_12(0.0)},PopPDoub_spindiff1={_f18_12(0.0)}")
    L.append(f",PopPDoub_spindiff2={_f18_12(0.0)}")
    L.append(f"PopTotImagTime={_f18_12(float(h.get('PopTotImagTime', 0.0)))}")
    L.append(f"PopGammaSing={_g18_12(0.0)},PopGammaDoub={_g18_12(0.0)},PopGammaTrip={_g18_12(0.0)},"
             f"PopGammaOpp={_g18_12(0.0)},PopGammaPar={_g18_12(0.0)},PopMaxDeathCpt={_g18_12(0.0)}")
    L.append(f",PopGammaSing_spindiff1={_g18_12(0.0)},PopGammaDoub_spindiff1={_g18_12(0.0)},"
             f"PopGammaDoub_spindiff2={_g18_12(0.0)}")
    L.append(f"PopNNodes={1:6d}")
    L.append(f"PopWalkersOnNodes={int(h['PopTotwalk']):16d},")
    rh = h.get("PopRandomHash")
    if rh is not None:
        L.append("PopRandomHash= " + "".join(f"{int(x):12d}," for x in rh))
    L.append(" &END")
    return "\n".join(L) + "\n"


_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[EeDd][-+]?\d+)?"


def _parse_value(v):
    v = v.strip().rstrip(",").strip()
    if not v:
        return None
    if v in ("T", ".TRUE.", ".T.", "t"):
        return True
    if v in ("F", ".FALSE.", ".F.", "f"):
        return False
    if v.startswith("("):  # complex: keep the real part (real-valued runs)
        nums = re.findall(_NUM, v)
        return float(nums[0].replace("D", "E").replace("d", "e"))
    toks = [t for t in re.split(r"[,\s]+", v) if t]
    out = []
    for t in toks:
        t2 = t.replace("D", "E").replace("d", "e")
        if re.fullmatch(r"[-+]?\d+", t2):
            out.append(int(t2))
        else:
            out.append(float(t2))
    return out[0] if len(out) == 1 else out


def _parse_header(text):
    """The POPSHEAD namelist as name -> value (numbers, logicals, lists)."""
    if not text.lstrip().startswith("# POPSFILE VERSION 4"):
        raise ValueError("not a version 4 POPSFILE")
    m = re.search(r"&POPSHEAD(.*?)&END", text, re.S | re.I)
    if not m:
        raise ValueError("POPSHEAD namelist not closed by &END")
    body = m.group(1)
    keys = list(re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*=", body))
    h = PopsHeader()
    for a, b in zip(keys, keys[1:] + [None]):
        end = b.start() if b is not None else len(body)
        h[a.group(1)] = _parse_value(body[a.end():end])
    return h, m.end()


def write_popsfile(stem, rows, W, S, nel, binary=True, **head):
    """Write occupied rows (n, W+S+1) int64, signs as FP64 bit patterns.

    binary=True writes stem+'HEAD' and stem+'BIN', else one text file stem.
    head: further POPSHEAD values (PopSft, PopSumNoatHF, PopSumENum, PopCyc,
    PopTau, PopPSingles, PopPParallel, PopPDoubles, PopTotImagTime,
    PopRandomHash, ...).  Rows whose signs are all zero are not written
    (write_pops_det, Popsfile.F90:2072-2078).  Returns the number written."""
    rows = np.ascontiguousarray(rows, dtype=np.int64).reshape(-1, W + S + 1)
    sg = rows[:, W:W + S].copy().view(np.float64)
    keep = np.abs(sg).sum(axis=1) > 0.0
    rows = rows[keep]
    h = PopsHeader(head)
    h.update(PopLensign=S, PopNEl=nel, PopTotwalk=len(rows), PopNIfD=W - 1)
    text = _header_text(h)
    if binary:
        with open(stem + "HEAD", "w") as f:
            f.write(text)
        rec = 8 * (W + S + 1)
        mark = struct.pack("<i", rec)
        with open(stem + "BIN", "wb") as f:
            for r in rows:
                f.write(mark + r.astype("<i8").tobytes() + mark)
    else:
        with open(stem, "w") as f:
            f.write(text)
            s = rows[:, W:W + S].copy().view(np.float64)
            for i in range(len(rows)):
                line = "".join(f"{int(x):24d}" for x in rows[i, :W])
                line += "".join(f"{x:30.8f}" for x in s[i])
                line += f"{int(rows[i, W + S]):24d}"
                f.write(line + "\n")
    return len(rows)


def read_popsfile(path):
    """Read a POPSFILE v4.  path is either a text POPSFILE or the stem of a
    POPSFILEHEAD/POPSFILEBIN pair (the stem, or either file name).  Returns
    (header, rows) with rows (n, W+S+1) int64 in CurrentDets layout."""
    stem = path
    for suf in ("HEAD", "BIN"):
        if path.endswith(suf) and os.path.exists(path[:-len(suf)] + "HEAD"):
            stem = path[:-len(suf)]
    binary = os.path.exists(stem + "HEAD") and os.path.exists(stem + "BIN")
    text = open(stem + "HEAD" if binary else stem).read()
    h, end = _parse_header(text)
    W = int(h["PopNIfD"]) + 1
    S = int(h.get("PopNIfSgn", h.get("PopLensign")))
    width = W + S + 1
    rows = []
    if binary:
        data = open(stem + "BIN", "rb").read()
        p = 0
        while p < len(data):
            (n,) = struct.unpack_from("<i", data, p)
            if n < 8 * width or p + 8 + n > len(data):
                raise ValueError(f"POPSFILEBIN: bad record length {n} at byte {p}")
            (n2,) = struct.unpack_from("<i", data, p + 4 + n)
            if n2 != n:
                raise ValueError(f"POPSFILEBIN: record markers differ at byte {p}")
            rows.append(np.frombuffer(data, dtype="<i8", count=width, offset=p + 4))
            p += 8 + n
    else:
        for line in text[end:].splitlines():
            f = line.split()
            if len(f) < width:
                continue
            r = np.zeros(width, dtype=np.int64)
            r[:W] = [int(x) for x in f[:W]]
            r[W:W + S] = np.array([float(x) for x in f[W:W + S]]).view(np.int64)
            r[W + S] = int(f[W + S])
            rows.append(r)
    rows = np.array(rows, dtype=np.int64).reshape(-1, width)
    sg = rows[:, W:W + S].copy().view(np.float64)
    rows = rows[np.abs(sg).sum(axis=1) > 0.0]  # read_popsfile_det skips empty records
    return h, rows
