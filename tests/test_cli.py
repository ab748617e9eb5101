"""The drop-in command line `mqw n m filename [start problems]` (MultiQW.cpp:241-458)."""
import os
import subprocess

import numpy as np
import pytest

import multistageqw_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "sk_n10_m2_seed1.bin")


def _run(args, cwd, env=None):
    e = dict(os.environ, **(env or {}))
    return subprocess.run([M.CLI_PATH, *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=300, env=e)


def test_cli_usage_error(tmp_path):
    r = _run([], tmp_path)
    assert r.returncode == 255  # the reference returns -1 when argc < 4 (MultiQW.cpp:244)


@pytest.mark.parametrize("args,msg", [
    ([10, 2, "/nonexistent/file"], "cannot read"),
    ([10, 2, FIX, 60, 10], "holds 64 records"),        # short file: the reference would repeat a record
    ([3, 2, FIX], "n <="),
    ([31, 2, FIX], "n <="),
    ([10, 0, FIX], "m <="),
])
def test_cli_rejects_bad_input(tmp_path, args, msg):
    r = _run(args, tmp_path)
    assert r.returncode == 1 and msg in r.stderr
    assert not os.path.exists(tmp_path / "output_10_2")


@pytest.mark.gpu
def test_cli_end_to_end_matches_oracle(tmp_path, oracle):
    """Default problems = 2000 needs a 2000-record file; with start/problems the CLI runs a slice."""
    print("[case] cli n=10 m=2 slice 8..20", flush=True)
    r = _run([10, 2, FIX, 8, 12], tmp_path)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 12
    recs = np.fromfile(FIX, "<f8").reshape(64, 55)[8:20]
    po, tto = oracle.solve(10, 2, recs, threads=4)
    out = np.fromfile(str(tmp_path / "output_10_2"), "<f4")
    assert out.size == 12 and np.max(np.abs(out - po)) < 1e-5
    for line, t, p in zip(lines, tto, out):
        assert line == M.format_line(t, p)
    # a second run appends another block (MultiQW.cpp:266)
    r2 = _run([10, 2, FIX, 0, 2], tmp_path)
    assert r2.returncode == 0
    assert np.fromfile(str(tmp_path / "output_10_2"), "<f4").size == 14


@pytest.mark.gpu
def test_cli_incremental_flush(tmp_path):
    """MQW_FLUSH=1 appends per batch; the file holds the same floats as one block at exit."""
    print("[case] cli flush n=10 m=2 slice 0..12", flush=True)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    r1 = _run([10, 2, FIX, 0, 12], a)
    r2 = _run([10, 2, FIX, 0, 12], b, env={"MQW_FLUSH": "1", "MQW_BATCH": "5"})
    assert r1.returncode == 0 and r2.returncode == 0, r2.stderr
    assert r1.stdout == r2.stdout
    assert (a / "output_10_2").read_bytes() == (b / "output_10_2").read_bytes()
