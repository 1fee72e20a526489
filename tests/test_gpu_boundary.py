"""The drop-in boundary on the GPU: any DEPTH a game allows (reference
main.cu:127 reads any integer), boards beyond the solver's limit rejected per
board, the export level capped to the stack, and the command line's table
counters (main.cu:158-162)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cfg(**kw):
    kw.setdefault("time_limit_ms", 100000)
    return go2.default_config(**kw)


def test_depth_above_32_on_shallow_boards():
    d = np.load(os.path.join(GOLD, "solve_small.npz"))
    b = go2.as_boards(d["player"], d["opponent"])
    for depth in (33, 40, 60):
        s, _ = go2.solve(b, depth, cfg())
        assert np.array_equal(s, d["score"]), depth
        t = go2.Task(len(b), depth, cfg())
        t.upload(b)
        t.launch()
        s2, _ = t.download()
        t.close()
        assert np.array_equal(s2, d["score"]), depth


def test_board_beyond_limit_rejected_at_any_depth():
    # one 34-empty board among shallow ones: GO2_E_BOARD naming it, at DEPTH
    # 40 (above the board's empties but beyond the exact solver's 32)
    good = ol.generate(3, 7, 20, 10)
    deep = ol.generate(4, 7, 1, 34)
    b = go2.as_boards(np.concatenate([good[:5], deep, good[5:]]))
    for depth in (34, 40, 60):
        with pytest.raises(go2.Go2Error) as e:
            go2.solve(b, depth, cfg())
        assert e.value.code == go2.GO2_E_BOARD and "board 5 " in str(e.value)


def test_board_deeper_than_depth_rejected():
    b = go2.as_boards(ol.generate(5, 7, 8, 14))
    with pytest.raises(go2.Go2Error) as e:
        go2.solve(b, 12, cfg())
    assert e.value.code == go2.GO2_E_BOARD


@pytest.mark.parametrize("export_min", [13, 15, 40, 60])
def test_export_level_above_stack(export_min):
    # Export levels above the deepest stack level (14 here) must leave every
    # index in range and every score exact.
    d = np.load(os.path.join(GOLD, "solve_mid.npz"))
    b = go2.as_boards(d["player"], d["opponent"])
    s, st = go2.solve(b, 14, cfg(export_min_empties=export_min, export_low_empties=10))
    assert np.array_equal(s, d["score"]), export_min


def test_cli_table_counters_and_depth_40(tmp_path):
    b = np.concatenate([ol.generate(61, 2, 300, 14), ol.generate(62, 2, 100, 12)])
    lines = [ol.encode(int(p), int(o)) for p, o in b]
    inp = tmp_path / "in.txt"
    inp.write_text(f"{len(lines)}\n" + "\n".join(lines) + "\n")
    conf = tmp_path / "cfg.json"
    conf.write_text(json.dumps({"tt_log2_buckets": 20, "tt_min_empties": 8}))
    out = tmp_path / "out.txt"
    r = subprocess.run([go2.SOLVER_BIN, str(inp), str(out), "40"], capture_output=True, text=True, timeout=120,
                       env={"GO2_CONFIG": str(conf), "PATH": "/usr/bin:/bin"})
    assert r.returncode == 0, r.stderr
    m = re.search(r"table update count: (\d+), table hit: (\d+), table find: (\d+)", r.stderr)
    assert m, r.stderr
    updates, hits, finds = (int(x) for x in m.groups())
    assert updates > 0 and finds > 0 and 0 < hits <= finds
    uniq = sorted({(int(p), int(o)) for p, o in b})
    want, _, _ = ol.solve_batch(np.array(uniq, dtype=np.uint64), 14, threads=16)
    assert out.read_text() == "".join(f"{ol.encode(p, o)} {int(s)}\n" for (p, o), s in zip(uniq, want))
