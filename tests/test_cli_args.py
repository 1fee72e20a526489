"""The command line's argument contract (reference main.cu:175-185): the
paths that end before any device call, so they run without a GPU."""
import os
import subprocess

import gpuothello2_amd as go2
import oracle_lib as ol


def run(args, env=None, timeout=60):
    e = dict(os.environ)
    e.pop("GO2_WEIGHTS", None)
    if env:
        e.update(env)
    return subprocess.run([go2.SOLVER_BIN, *args], capture_output=True, text=True, timeout=timeout, env=e)


def board_file(tmp_path, empties=10, n=3):
    b = ol.generate(91, 13, n, empties)
    p = tmp_path / "in.txt"
    p.write_text(f"{n}\n" + "".join(ol.encode(int(x), int(y)) + "\n" for x, y in b))
    return p


def test_usage_exits_1():
    r = run([])
    assert r.returncode == 1 and "usage" in r.stderr
    r = run(["a", "b"])
    assert r.returncode == 1


def test_bad_depths_exit_2(tmp_path):
    inp = board_file(tmp_path)
    out = tmp_path / "out.txt"
    for depth in ("0", "33", "x"):
        r = run([str(inp), str(out), depth])
        assert r.returncode == 2 and "DEPTH" in r.stderr, depth
    for think in ("0", "31", "y"):
        r = run([str(inp), str(out), "60", think])
        assert r.returncode == 2 and "THINK_DEPTH" in r.stderr, think


def test_input_problems_exit_2(tmp_path):
    out = tmp_path / "out.txt"
    r = run([str(tmp_path / "missing.txt"), str(out), "10"])
    assert r.returncode == 2 and "cannot open input" in r.stderr
    inp = board_file(tmp_path, empties=14)
    r = run([str(inp), str(out), "12"])
    assert r.returncode == 2 and "more than DEPTH" in r.stderr


def test_think_mode_needs_weights(tmp_path):
    inp = board_file(tmp_path, empties=30)
    r = run([str(inp), str(tmp_path / "out.txt"), "60", "4"])
    assert r.returncode == 2 and "GO2_WEIGHTS" in r.stderr
