"""Autotuner and the uneven-depth batch (one persistent launch over boards of
10..14 empties, captured in a HIP graph): the tuned configuration must keep
every score exact."""
import json

import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol
from gpuothello2_amd import autotune

pytestmark = pytest.mark.gpu


def test_autotune_then_uneven_batch_exact(tmp_path):
    b = go2.generate(777, 3, 6000, 10, 14)
    res = autotune.autotune(b, 14, sample=3000)
    assert np.isfinite(res["kernel_ms"]) and len(res["history"]) > 3
    path = tmp_path / "tuned.json"
    autotune.save(str(path), res)
    params = autotune.load(str(path))
    assert params == {k: int(v) for k, v in res["params"].items()}
    cfg = autotune.apply(go2.default_config(graph_launch=1), params)
    t = go2.Task(len(b), 14, cfg)
    t.upload(b)
    t.launch()
    s1, st = t.download()
    t.launch()
    s2, _ = t.download()
    t.close()
    assert np.array_equal(s1, s2)
    k = 1500
    want, _, _ = ol.solve_batch(np.stack([b["player"][:k], b["opponent"][:k]], 1), 14, threads=16)
    assert np.array_equal(s1[:k], want)
    e = 64 - np.array([bin(int(p) | int(o)).count("1") for p, o in zip(b["player"], b["opponent"])])
    assert set(np.unique(e)) == set(range(10, 15))


def test_cli_reads_tuned_config(tmp_path):
    import subprocess
    b = ol.generate(17, 2, 200, 10)
    lines = [ol.encode(int(p), int(o)) for p, o in b]
    inp = tmp_path / "in.txt"
    inp.write_text(f"{len(lines)}\n" + "\n".join(lines) + "\n")
    cfg = tmp_path / "cfg.json"
    cfg.write_text(json.dumps({"positions_per_wave": 4, "split_min_empties": 6, "lds_levels": 7}))
    out = tmp_path / "out.txt"
    r = subprocess.run([go2.SOLVER_BIN, str(inp), str(out), "10"], capture_output=True, text=True, timeout=120,
                       env={"GO2_CONFIG": str(cfg), "PATH": "/usr/bin:/bin"})
    assert r.returncode == 0, r.stderr
    uniq = sorted({(int(p), int(o)) for p, o in b})
    want, _, _ = ol.solve_batch(np.array(uniq, dtype=np.uint64), 10, threads=16)
    assert out.read_text() == "".join(f"{ol.encode(p, o)} {int(s)}\n" for (p, o), s in zip(uniq, want))
