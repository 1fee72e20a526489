"""The solver's board-file path (read, validate, sort, de-duplicate, write),
run through bin/board_io_check without a GPU.  The reading rules are the
reference's (main.cu:49-67): fscanf("%lld\\n") for the count, fgets(buf, 32)
for each line, the first 16 bytes decoded.  CPU only."""
import os
import random
import subprocess
import time

import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol

CHECK = os.path.join(os.path.dirname(go2.SOLVER_BIN), "board_io_check")


def run(tmp_path, text, *extra):
    inp = tmp_path / "in.txt"
    inp.write_bytes(text.encode("latin-1"))
    out = tmp_path / "out.txt"
    r = subprocess.run([CHECK, str(inp), str(out), *extra], capture_output=True, text=True, timeout=300)
    return r, (out.read_text() if out.exists() else None)


def expected(pairs, with_moves=False):
    uniq = sorted(set(pairs))
    tail = " 0 ps\n" if with_moves else " 0\n"
    return "".join(ol.encode(p, o) + tail for p, o in uniq)


def pairs_of(b):
    return [(int(p), int(o)) for p, o in b]


def test_sorted_unique_round_trip(tmp_path):
    b = pairs_of(ol.generate(81, 12, 3000, 20))
    rows = b + b[:700]
    random.Random(5).shuffle(rows)
    text = f"{len(rows)}\n" + "".join(ol.encode(p, o) + "\n" for p, o in rows)
    r, out = run(tmp_path, text)
    assert r.returncode == 0, r.stderr
    assert out == expected(rows)
    r, out = run(tmp_path, text, "moves")
    assert out == expected(rows, with_moves=True)


def test_reference_reading_rules(tmp_path):
    b = pairs_of(ol.generate(82, 12, 3, 20))
    e = [ol.encode(p, o) for p, o in b]
    # blank lines and spaces after the count are skipped; CRLF endings and
    # trailing text after the 16 characters are ignored; a 31-byte fgets cut
    # turns one long physical line into two board lines.
    text = "3 \n\n  " + e[0] + "\r\n" + e[1] + " " * 15 + e[2] + "\n"
    r, out = run(tmp_path, text)
    assert r.returncode == 0, r.stderr
    assert out == expected(b)


def test_errors(tmp_path):
    e = ol.encode(*pairs_of(ol.generate(83, 12, 1, 20))[0])
    r, _ = run(tmp_path, "2\n" + e + "\n")
    assert r.returncode == 1 and "Failed to load boards" in r.stderr
    r, _ = run(tmp_path, "1\n" + e[:10] + "\n")
    assert r.returncode == 2 and "shorter" in r.stderr
    r, _ = run(tmp_path, "1\n" + e[:5] + "~" + e[6:] + "\n")
    assert r.returncode == 2 and "base81" in r.stderr
    r, _ = run(tmp_path, "x\n" + e + "\n")
    assert r.returncode == 2 and "count" in r.stderr
    r, out = run(tmp_path, "0\n")
    assert r.returncode == 0 and out == ""
    r, _ = run(tmp_path, "1000000000000000\n" + e + "\n")  # a count far beyond the file
    assert r.returncode == 1 and "Failed to load boards" in r.stderr


def test_two_million_lines(tmp_path):
    # Throughput of the host path on a 2,000,000-line file (one solver batch).
    b = ol.generate(84, 12, 20000, 24)
    rng = np.random.default_rng(3)
    idx = rng.integers(0, len(b), 2000000)
    enc = [ol.encode(int(p), int(o)) for p, o in b]
    text = "2000000\n" + "\n".join(enc[i] for i in idx) + "\n"
    t0 = time.perf_counter()
    r, out = run(tmp_path, text)
    secs = time.perf_counter() - t0
    assert r.returncode == 0, r.stderr
    assert out == expected(pairs_of(b[np.unique(idx)]))
    print(f"2,000,000 lines: {secs:.2f} s end to end; {r.stderr.strip()}")
    assert secs < 30
