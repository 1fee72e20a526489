"""The C-ABI library: loads, exports every declared symbol, and its host-only
entry points (text codec, argument checks) behave.  CPU only: no call here
reaches a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(go2_[a-z_0-9]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = go2.lib()
    decl = declared_symbols()
    assert decl, "no declarations found"
    for name in decl:
        assert hasattr(L, name), name
    assert set(go2.EXPORTED_SYMBOLS) <= set(decl)


def test_exports_listed_by_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", go2.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (go2_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_text_codec_matches_reference_vectors():
    d = np.load(os.path.join(GOLD, "ref_bitboard.npz"))
    for p, o, t in zip(d["player"], d["opponent"], d["text"]):
        assert go2.board_to_text(int(p), int(o)) == str(t)
        assert go2.board_from_text(str(t)) == (int(p), int(o))


def test_text_codec_rejects_garbage():
    with pytest.raises(go2.Go2Error) as e:
        go2.board_from_text("!" * 15 + "~")
    assert e.value.code == go2.GO2_E_PARSE


def test_argument_errors_need_no_device():
    b = go2.as_boards([0x0000000810000000], [0x0000001008000000])
    for depth in (0, 61, -5):  # DEPTH is 1..60, as a game allows
        with pytest.raises(go2.Go2Error) as e:
            go2.solve(b, depth)
        assert e.value.code == go2.GO2_E_ARG
    with pytest.raises(go2.Go2Error) as e:
        go2.solve(b, 12, go2.default_config(block_size=100))
    assert e.value.code == go2.GO2_E_ARG
    with pytest.raises(go2.Go2Error) as e:
        go2.solve(b, 12, go2.default_config(block_size=512))
    assert e.value.code == go2.GO2_E_ARG
    with pytest.raises(go2.Go2Error) as e:
        go2.solve(b, 12, go2.default_config(ordering=7))
    assert e.value.code == go2.GO2_E_ARG
    L = go2.lib()
    assert L.go2_generate(1, 1, 4, 10, 5, None) == go2.GO2_E_ARG


def test_config_default_is_valid():
    c = go2.default_config()
    assert c.ordering in (go2.ORDER_STATIC, go2.ORDER_MOBILITY, go2.ORDER_LEARNED)
    assert c.block_size % 64 == 0


def _run_cli(args, stdin_text=None, tmp_path=None):
    return subprocess.run([go2.SOLVER_BIN] + args, capture_output=True, text=True, timeout=60)


def test_cli_usage_and_errors(tmp_path):
    r = _run_cli([])
    assert r.returncode == 1 and "usage:" in r.stderr
    inp = tmp_path / "in.txt"
    inp.write_text("2\n!!%[nNNNNNnNNNNN\n")
    r = _run_cli([str(inp), str(tmp_path / "out.txt"), "10"])
    assert r.returncode == 1 and "Failed to load boards" in r.stderr
    inp.write_text("1\n!!%[nNNNNNnNNNNN\n")
    r = _run_cli([str(inp), str(tmp_path / "out.txt"), "9"])
    assert r.returncode == 2 and "more than DEPTH" in r.stderr
    r = _run_cli([str(inp), str(tmp_path / "out.txt"), "ten"])
    assert r.returncode == 2
    r = _run_cli([str(inp), str(tmp_path / "out.txt"), "10", "4"])
    assert r.returncode == 2 and "think mode" in r.stderr
    r = _run_cli([str(tmp_path / "missing.txt"), str(tmp_path / "out.txt"), "10"])
    assert r.returncode == 2
    inp.write_text("1\n!!%[nNNNNNnNNNN~\n")
    r = _run_cli([str(inp), str(tmp_path / "out.txt"), "10"])
    assert r.returncode == 2 and "not base81" in r.stderr
