"""Host logic of bench.py (CPU only): the spawn-sort key and the compulsory
byte model behind the roofline figures, the kernel-trace summary, and the
refusal of more than one GPU."""
import csv
import json
import math
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

COUNTS = dict(Nd=5.45e5, Na=9.97e5, Ns=5.46e5, Nu=5.0e5, Nn=1.0e5)


@pytest.mark.parametrize("nbasis,nel,expect", [
    (40, 10, (4, 2)),    # C(40,10) = 847 660 528 < 2^32: colex-rank key, partition sort
    (20, 10, (4, 2)),    # C(20,10) = 184 756
    (36, 18, (8, 5)),    # C(36,18) > 2^32: determinant word, 36 bits in 5 passes
    (80, 26, (16, 10)),  # two words: 16 bits of the second (2 passes) + 8 passes of the first
    (60, 14, (8, 8)),    # C(60,14) > 2^32, one word
])
def test_sort_key_matches_the_engine_rule(nbasis, nel, expect):
    assert bench.sort_key(nbasis, nel) == expect
    assert (expect[0] == 4) == (math.comb(nbasis, nel) <= 2 ** 32)


def test_sort_and_iteration_bytes():
    kb, passes = bench.sort_key(40, 10)
    assert bench.algorithmic_bytes("sort", COUNTS, nel=10) == passes * COUNTS["Ns"] * (kb + 2 * (kb + 4))
    stages = sum(bench.algorithmic_bytes(s, COUNTS, nel=10) for s in bench.STAGES if s != "sort")
    assert bench.iteration_bytes(COUNTS, nel=10) == pytest.approx(stages + 2 * COUNTS["Ns"] * (kb + 4))
    # the merge is fused into the annihilation stage
    assert bench.algorithmic_bytes("merge", COUNTS) == 0.0
    # prepass: 53 bytes per determinant for one word and one replica
    assert bench.algorithmic_bytes("prepass", COUNTS) == pytest.approx(53 * COUNTS["Nd"])
    # spawn: parents (17 B), one 20-byte alias entry per double attempt, the
    # queued singles (16 B) and the double spawns (20 B) written
    nq, nsq = 0.1 * COUNTS["Na"], 0.1 * COUNTS["Ns"]
    assert bench.algorithmic_bytes("spawn", COUNTS) == pytest.approx(
        17 * COUNTS["Nd"] + 20 * (COUNTS["Na"] - nq) + 16 * nq + 20 * (COUNTS["Ns"] - nsq))


def test_summarize_trace(tmp_path, capsys):
    """Per-iteration windows end at k_end_iter; averages over the last ones."""
    path = tmp_path / "trace.csv"
    rows = []
    t = 0
    for it in range(4):
        for name, dur in (("void ngpu::k_spawn<1, 1, 1>(ngpu::DevSys)", 50_000 + it),
                          ("void ngpu::k_end_iter(ngpu_iter_out_t*)", 5_000)):
            rows.append({"Kernel_Name": name, "Start_Timestamp": t, "End_Timestamp": t + dur})
            t += dur + 1000
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        w.writerows(rows)
    bench.summarize_trace(str(path), 2)
    out = capsys.readouterr().out
    assert "`k_spawn<1, 1, 1>` | 1 | 50.0 |" in out  # (50002 + 50003) / 2 ns
    assert "`k_end_iter` | 1 | 5.0 |" in out


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_window_counts_cover_every_timed_step():
    """The bench's per-step counts come from the driver's running totals, so
    a timed window of 20 steps after 5 warm-up steps (not aligned to the
    10-step update cycle) counts all 20 iterations."""
    from test_driver_host import ScriptedBackend
    from neci_stable_amd.fcimc import FciMCPar
    d = FciMCPar(ScriptedBackend(), tau=0.01, target_walkers=10.0, steps_shift=10, shift_damp=0.1)
    d.run(5)
    t0 = dict(d.totals)
    d.run(20)
    c = bench.window_counts(t0, d.totals)
    its = list(range(6, 26))
    assert c["Na"] == pytest.approx(sum(1000 + i for i in its) / 20)
    assert c["walkers"] == pytest.approx(c["Na"])   # one attempt per walker in the script
    assert c["Ns"] == pytest.approx(sum(its) / 20)


def test_fp32_values_and_core_stage_bytes():
    """FP32 storage takes 4 bytes off every spawned amplitude written or
    gathered and off every core Hamiltonian element; the core stage reads
    each stored element (value + int32 column) once and the core vector."""
    c = dict(COUNTS, Nc=1.18e7, Ncore=5e4)
    nsd = 0.9 * COUNTS["Ns"]
    assert (bench.algorithmic_bytes("spawn", c) - bench.algorithmic_bytes("spawn", c, val_bytes=4)
            == pytest.approx(4 * nsd))
    assert (bench.algorithmic_bytes("annihilate", c) - bench.algorithmic_bytes("annihilate", c, val_bytes=4)
            == pytest.approx(4 * COUNTS["Ns"]))
    assert bench.algorithmic_bytes("core", c) == pytest.approx(1.18e7 * 12 + 2 * 8 * 5e4)
    assert bench.algorithmic_bytes("core", c, val_bytes=4) == pytest.approx(1.18e7 * 8 + 2 * 8 * 5e4)
    assert bench.algorithmic_bytes("core", COUNTS) == 0.0   # no core space
