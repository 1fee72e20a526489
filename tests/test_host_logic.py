"""Host-side logic that needs no GPU: the autotuner's configuration plumbing
and the command line's JSON config fields."""
import json
import os
import re

import gpuothello2_amd as go2
from gpuothello2_amd import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_apply_copies_and_sets_fields():
    base = go2.default_config(ordering=go2.ORDER_MOBILITY)
    out = autotune.apply(base, {"positions_per_wave": 4, "tt_log2_buckets": 29, "tt_min_empties": 8})
    assert (out.positions_per_wave, out.tt_log2_buckets, out.tt_min_empties) == (4, 29, 8)
    assert out.ordering == go2.ORDER_MOBILITY
    assert base.positions_per_wave != 4 or base.tt_log2_buckets != 29  # the base is untouched


def test_save_load_round_trip(tmp_path):
    p = tmp_path / "tuned.json"
    autotune.save(str(p), {"params": {"block_size": 128, "split_min_empties": 7}})
    assert json.loads(p.read_text()) == {"block_size": 128, "split_min_empties": 7}
    assert autotune.load(str(p)) == {"block_size": 128, "split_min_empties": 7}


def test_search_space_names_real_fields():
    fields = {name for name, _ in go2.Config._fields_}
    for space in (autotune.SPACE, autotune.DEEP_SPACE):
        assert set(space) <= fields


def test_cli_reads_every_config_field():
    # GO2_CONFIG accepts every go2_config field the tuner can write; think
    # mode is chosen by the fifth argument, as in the reference, not by JSON.
    src = open(os.path.join(ROOT, "gpuothello2_amd", "cli", "solver.cpp")).read()
    names = set(re.findall(r'\{"([a-z0-9_]+)", &c->', src))
    fields = {name for name, _ in go2.Config._fields_} - {"reserved", "think_depth"}
    assert fields <= names, fields - names
    assert "think_depth" not in names


def test_missing_library_fails_loudly():
    # No CPU fallback: without the built HIP library every entry point raises.
    import subprocess
    import sys
    code = ("import gpuothello2_amd as go2, numpy as np\n"
            "go2.LIB_PATH = '/nonexistent/libgo2.so'\n"
            "try:\n"
            "    go2.solve(go2.as_boards(np.zeros((1, 2), dtype=np.uint64)), 10)\n"
            "except ImportError as e:\n"
            "    print('raised', e)\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0 and "raised" in r.stdout and "not built" in r.stdout, r.stderr
