"""The command line's batch pipeline (reference main.cu:130-155): more than
2,000,000 boards go through two tasks whose uploads are queued from
page-locked memory on their own streams, so the second batch copies while
the first solves.  The output must equal the single-batch run byte for byte,
and a sample is checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol

pytestmark = pytest.mark.gpu
N = 2_100_000


def base81_lines(boards: np.ndarray) -> bytes:
    """Board text of every board (base81, reference test.cpp:7-19), one per line."""
    p = boards["player"]
    o = boards["opponent"]
    out = np.empty((len(boards), 17), dtype=np.uint8)
    w = (1, 3, 9, 32)
    for i in range(16):
        v = np.full(len(boards), 33, dtype=np.int64)
        for j in range(4):
            sq = np.uint64(4 * i + j)
            v += ((((p >> sq) & np.uint64(1)) + 2 * ((o >> sq) & np.uint64(1))).astype(np.int64)) * w[j]
        out[:, i] = v.astype(np.uint8)
    out[:, 16] = ord("\n")
    return out.tobytes()


def run_cli(inp, out, env_extra):
    env = {"PATH": "/usr/bin:/bin", **env_extra}
    r = subprocess.run([go2.SOLVER_BIN, str(inp), str(out), "10"], capture_output=True, text=True, timeout=240,
                       env=env)
    assert r.returncode == 0, r.stderr
    return r.stderr


def test_two_batches_equal_single_batch(tmp_path):
    b = go2.generate(2024, 11, N, 10)
    text = base81_lines(b)
    for k in (0, 1, N - 1):  # the vectorised encoder agrees with the oracle's
        assert text[17 * k:17 * k + 16].decode() == ol.encode(int(b["player"][k]), int(b["opponent"][k]))
    inp = tmp_path / "in.txt"
    with open(inp, "wb") as f:
        f.write(f"{N}\n".encode())
        f.write(text)
    out2, out1 = tmp_path / "two.txt", tmp_path / "one.txt"
    err2 = run_cli(inp, out2, {})
    err1 = run_cli(inp, out1, {"GO2_CLI_BATCH": str(N + 1)})
    print(err2.strip().splitlines()[-2], "|", err1.strip().splitlines()[-2])
    uniq = len(np.unique(np.stack([b["player"], b["opponent"]], 1), axis=0))
    assert f"n = {uniq}" in err2 and uniq > 2_000_000
    assert os.path.getsize(out2) == os.path.getsize(out1)
    assert out2.read_bytes() == out1.read_bytes()
    # a sample of lines against the oracle (first, middle, last)
    lines = out2.read_text().splitlines()
    assert len(lines) == uniq
    sample = lines[:500] + lines[uniq // 2:uniq // 2 + 500] + lines[-500:]
    boards = np.array([ol.decode(ln.split()[0]) for ln in sample], dtype=np.uint64)
    want, _, _ = ol.solve_batch(boards, 10, threads=16)
    assert [int(ln.split()[1]) for ln in sample] == [int(x) for x in want]
