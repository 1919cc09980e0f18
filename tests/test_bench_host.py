"""bench.py's host-side parts (no GPU): the one-GPU guard, the rocprofv3 table
summaries, and the algorithmic byte counts of the roofline."""
import csv
import os
import sqlite3
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_summarize_keeps_the_last_dispatches_of_every_kernel(tmp_path, capsys):
    db = tmp_path / "run_results.db"
    con = sqlite3.connect(db)
    con.execute("create table kernels (name text, start integer, end integer)")
    t = 0
    for step in range(5):
        for name, dur in (("qs_pass_0", 100_000 + step * 1000), ("qs_pass_1", 50_000)):
            con.execute("insert into kernels values (?, ?, ?)", (name, t, t + dur))
            t += dur + 10
    con.commit()
    con.close()
    bench.summarize(str(db), 2)
    out = capsys.readouterr().out
    rows = {line.split("|")[1].strip(" `"): line.split("|")[2:-1] for line in out.splitlines()[2:]}
    calls, total, avg = int(rows["qs_pass_0"][0]), float(rows["qs_pass_0"][1]), float(rows["qs_pass_0"][2])
    assert calls == 2 and total == pytest.approx(103.0 + 104.0) and avg == pytest.approx(103.5)
    assert float(rows["qs_pass_1"][2]) == pytest.approx(50.0)


def test_counter_summary_averages_per_dispatch(tmp_path, capsys):
    path = tmp_path / "run_counter_collection.csv"
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for d in range(4):
            # two rows per dispatch (per-XCC values) are summed first
            w.writerow([d, "qs_pass_1", "FETCH_SIZE", 10.0 * d])
            w.writerow([d, "qs_pass_1", "FETCH_SIZE", 1.0])
    bench.summarize(str(path), 2)
    out = capsys.readouterr().out
    line = [x for x in out.splitlines() if "qs_pass_1" in x][0]
    assert float(line.split("|")[4]) == pytest.approx((21.0 + 31.0) / 2)


def test_pass_bytes_follow_the_algorithmic_units():
    B, n = 64, 20
    amps = B * 2 ** n
    b = bench.pass_bytes(B, n, 6, stores_last=False)
    assert b == [8 * amps] + [16 * amps] * 4 + [8 * amps]
    assert sum(b) / 6 == pytest.approx(amps * 80 / 6)
    assert bench.pass_bytes(B, n, 1, stores_last=True) == [8 * amps]
