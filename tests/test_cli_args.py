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
    # DEPTH takes any value a game can have, 1..60 (main.cu:127 reads any
    # integer; below 1 or above 60 no board fits).
    inp = board_file(tmp_path)
    out = tmp_path / "out.txt"
    for depth in ("0", "61", "x", "-3"):
        r = run([str(inp), str(out), depth])
        assert r.returncode == 2 and "DEPTH" in r.stderr, depth
    for think in ("0", "31", "y"):
        r = run([str(inp), str(out), "60", think])
        assert r.returncode == 2 and "THINK_DEPTH" in r.stderr, think


def test_deep_depth_passes_argument_checks(tmp_path):
    # DEPTH 33..60 on 10-empty boards passes every argument and board check;
    # without a GPU the run then stops at the first device call.
    inp = board_file(tmp_path)
    out = tmp_path / "out.txt"
    for depth in ("33", "40", "60"):
        r = run([str(inp), str(out), depth])
        assert "DEPTH" not in r.stderr and "empties" not in r.stderr, (depth, r.stderr)
        assert r.returncode in (0, 2), r.stderr


def test_board_beyond_solver_limit_exits_2(tmp_path):
    # A 34-empty board with DEPTH 40: more empties than the exact solver takes.
    inp = board_file(tmp_path, empties=34)
    r = run([str(inp), str(tmp_path / "out.txt"), "40"])
    assert r.returncode == 2 and "at most 32" in r.stderr


def test_six_arguments_solve_exactly(tmp_path):
    # Think mode is argc == 5 exactly (main.cu:180); more arguments solve.
    inp = board_file(tmp_path)
    r = run([str(inp), str(tmp_path / "out.txt"), "10", "0", "extra"])
    assert "THINK_DEPTH" not in r.stderr and "GO2_WEIGHTS" not in r.stderr, r.stderr
    assert r.returncode in (0, 2)


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
