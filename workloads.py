"""The five workload configurations of this solver, run end to end on one GPU.

  C1  10,000 positions at 10 empties, plain reference method on the host CPU
      (the oracle's reference-method search), and the same set on the GPU.
  C2  1,000,000 positions at 12 empties, plain batched GPU kernel (static
      corner / X-square ordering, no table); bit-exact against the CPU on a
      sample.
  C3  200,000 positions with empties drawn uniformly from 10..18, one
      persistent launch per solve replayed from a captured HIP graph, launch
      configuration picked by the autotuner on a sample.
  C4  learned ordering: the trained int16 evaluator orders shallow nodes of
      100,000 positions at 16 empties; the same positions with mobility-only
      ordering; node counts and positions/s for both.
  C5  8,192 positions at 20 empties with a 2^26-bucket transposition table
      and learned ordering; hit rate and table traffic.

Every configuration prints one JSON line.  Positions come from the GPU
generator (random legal playouts, SplitMix64 streams).  Usage:

    python workloads.py [c1 c2 c3 c4 c5] [--scale 0.1]

`--scale` shrinks every batch (for quick checks); the configurations above
are scale 1.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from gpuothello2_amd import (ORDER_LEARNED, ORDER_MOBILITY, ORDER_STATIC, Task,  # noqa: E402
                             default_config, generate)
from gpuothello2_amd import autotune, learn  # noqa: E402


def log(msg: str) -> None:
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib  # test infrastructure: the CPU checker and C1's CPU leg

    return oracle_lib


def _solve(boards, depth, cfg, label):
    t = Task(len(boards), depth, cfg)
    t.upload(boards)
    t0 = time.perf_counter()
    t.launch()
    t.wait()
    wall = time.perf_counter() - t0
    s, st = t.download()
    t.close()
    log(f"[{label}] {len(boards)} positions: kernel {st['kernel_ms']:.0f} ms, nodes {st['nodes']}")
    return s, st, wall


def _pairs(b):
    return np.stack([b["player"], b["opponent"]], 1)


def c1(scale: float) -> dict:
    ol = _oracle()
    n = max(100, int(10000 * scale))
    b = generate(101, 1, n, 10)
    cpu, nodes, secs = ol.solve_batch(_pairs(b), 10, threads=min(16, os.cpu_count() or 1))
    g, st, _ = _solve(b, 10, default_config(ordering=ORDER_STATIC), "c1")
    return {"config": "C1", "positions": n, "empties": 10, "cpu_positions_per_s": n / secs,
            "cpu_threads": min(16, os.cpu_count() or 1), "gpu_positions_per_s": n / (st["kernel_ms"] / 1e3),
            "match": bool(np.array_equal(cpu, g))}


def c2(scale: float) -> dict:
    ol = _oracle()
    n = max(1000, int(1000000 * scale))
    b = generate(202, 2, n, 12)
    # the plain batched kernel: one lane per board, static order, no table, no help, no leaf batches
    plain = default_config(ordering=ORDER_STATIC, positions_per_wave=64, split_min_empties=99,
                           export_min_empties=99, leaf_batch_min=65)
    s, st, _ = _solve(b, 12, plain, "c2")
    k = min(n, 20000)
    cpu, _, _ = ol.solve_batch(_pairs(b[:k]), 12, threads=min(16, os.cpu_count() or 1))
    d, std, _ = _solve(b, 12, default_config(ordering=ORDER_STATIC), "c2 default shape")
    return {"config": "C2", "positions": n, "empties": 12, "ordering": "static",
            "default_shape_positions_per_s": n / (std["kernel_ms"] / 1e3), "default_shape_match": bool(np.array_equal(d, s)),
            "positions_per_s": n / (st["kernel_ms"] / 1e3), "nodes_per_position": st["nodes"] / n,
            "checked_vs_cpu": k, "match": bool(np.array_equal(cpu, s[:k]))}


def c3(scale: float) -> dict:
    ol = _oracle()
    n = max(2000, int(200000 * scale))
    b = generate(303, 3, n, 10, 18)
    base = default_config(ordering=ORDER_LEARNED if os.path.exists(learn.WEIGHTS_PATH) else ORDER_MOBILITY)
    if base.ordering == ORDER_LEARNED:
        learn.install()
    tuned = autotune.autotune(b, 18, base, sample=max(1000, n // 10), log=log)
    cfg = autotune.apply(base, tuned["params"])
    cfg.graph_launch = 1
    t = Task(n, 18, cfg)
    t.upload(b)
    times = []
    for _ in range(2):  # first launch captures the graph, the second replays it
        t.launch()
        t.wait()
        s, st = t.download()
        times.append(st["kernel_ms"])
    t.close()
    e = 64 - np.array([bin(int(p) | int(o)).count("1") for p, o in zip(b["player"], b["opponent"])])
    low = np.nonzero(e <= 14)[0][:5000]
    cpu, _, _ = ol.solve_batch(_pairs(b[low]), 14, threads=min(16, os.cpu_count() or 1))
    return {"config": "C3", "positions": n, "empties": "10..18",
            "histogram": {int(k): int(v) for k, v in zip(*np.unique(e, return_counts=True))},
            "tuned": tuned["params"], "graph_replay_ms": times[1], "capture_launch_ms": times[0],
            "positions_per_s": n / (times[1] / 1e3), "checked_vs_cpu": int(len(low)),
            "match": bool(np.array_equal(cpu, s[low]))}


def c4(scale: float) -> dict:
    n = max(2000, int(100000 * scale))
    meta = learn.install()
    b = generate(20261203, 4, n, 16)
    sl, stl, _ = _solve(b, 16, default_config(ordering=ORDER_LEARNED), "c4 learned")
    sm, stm, _ = _solve(b, 16, default_config(ordering=ORDER_MOBILITY), "c4 mobility")
    return {"config": "C4", "positions": n, "empties": 16,
            "learned": {"positions_per_s": n / (stl["kernel_ms"] / 1e3), "nodes_per_position": stl["nodes"] / n},
            "mobility_only": {"positions_per_s": n / (stm["kernel_ms"] / 1e3), "nodes_per_position": stm["nodes"] / n},
            "evaluator_valid_mse": meta["valid_mse"], "match": bool(np.array_equal(sl, sm))}


def c5(scale: float) -> dict:
    n = max(256, int(8192 * scale))
    learn.install()
    b = generate(505, 5, n, 20)
    cfg = default_config(ordering=ORDER_LEARNED, tt_log2_buckets=26, tt_min_empties=10, time_limit_ms=900000)
    s, st, _ = _solve(b, 20, cfg, "c5")
    secs = st["kernel_ms"] / 1e3
    k = min(n, 256)
    cfg2 = default_config(ordering=ORDER_LEARNED, time_limit_ms=300000)
    s2, st2, _ = _solve(b[:k], 20, cfg2, "c5 no table (prefix)")
    return {"config": "C5", "positions": n, "empties": 20, "table_buckets": 2 ** 26,
            "positions_per_s": n / secs, "nodes_per_position": st["nodes"] / n,
            "table_probes": st["tt_probes"], "table_hits": st["tt_hits"], "table_cuts": st["tt_cuts"],
            "table_stores": st["tt_stores"], "hit_rate": st["tt_hits"] / max(1, st["tt_probes"]),
            "table_gb_per_s_algorithmic": (st["tt_probes"] + st["tt_stores"]) * 16 / secs / 1e9,
            "no_table_prefix_match": bool(np.array_equal(s[:k], s2))}


def readme(scale: float) -> dict:
    """The reference README's run (README.md:22-24): 1,091,780 unique boards at
    10 empties through the command line, DEPTH 10.  Its boards are not
    published, so random-playout boards of the same count and depth stand in;
    the reference's GPU is not named."""
    import subprocess
    import tempfile

    from gpuothello2_amd import SOLVER_BIN

    n = max(1000, int(1091780 * scale))
    b = generate(1091780, 6, int(n * 1.02) + 64, 10)
    pairs = np.unique(_pairs(b), axis=0)[:n]
    assert len(pairs) == n, "not enough unique boards"
    rng = np.random.default_rng(1)
    pairs = pairs[rng.permutation(n)]  # input order is not sorted; the CLI sorts
    out = {"config": "README scale", "positions": n, "empties": 10, "reference_elapsed_s": 4.623645,
           "reference_positions_per_s": 1091780 / 4.623645}
    with tempfile.TemporaryDirectory() as d:
        inp, res = os.path.join(d, "in.txt"), os.path.join(d, "out.txt")
        chars = _base81(pairs)
        with open(inp, "wb") as f:
            f.write(f"{n}\n".encode())
            f.write(chars)
        for label, env in (("default", {}), ("static", {"GO2_ORDERING": "static"})):
            t0 = time.perf_counter()
            r = subprocess.run([SOLVER_BIN, inp, res, "10"], capture_output=True, text=True,
                               env={"PATH": "/usr/bin:/bin", **env})
            wall = time.perf_counter() - t0
            assert r.returncode == 0, r.stderr
            line = next(ln for ln in r.stderr.splitlines() if "elapsed:" in ln)
            ms = int(line.split("elapsed: ")[1].split("ms")[0])
            out[label] = {"elapsed_ms": ms, "positions_per_s": n / (ms / 1e3), "process_wall_s": wall,
                          "status": line}
            log(f"[readme] {label}: {line}")
    return out


def _base81(pairs: np.ndarray) -> bytes:
    """Base81 lines (reference test.cpp:7-19) of (n, 2) boards."""
    p, o = pairs[:, 0], pairs[:, 1]
    out = np.empty((len(pairs), 17), dtype=np.uint8)
    for i in range(16):
        v = np.full(len(pairs), 33, dtype=np.int64)
        for j, w in enumerate((1, 3, 9, 32)):
            sq = np.uint64(4 * i + j)
            v += ((((p >> sq) & np.uint64(1)) + 2 * ((o >> sq) & np.uint64(1))).astype(np.int64)) * w
        out[:, i] = v.astype(np.uint8)
    out[:, 16] = ord("\n")
    return out.tobytes()


RUNNERS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c5": c5, "readme": readme}

if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    scale = 1.0
    if "--scale" in sys.argv:
        scale = float(sys.argv[sys.argv.index("--scale") + 1])
        args = [a for a in args if a != sys.argv[sys.argv.index("--scale") + 1]]
    for name in args or ["c1", "c2", "c3", "c4", "c5"]:
        log(f"[workloads] {name} (scale {scale})")
        print(json.dumps(RUNNERS[name](scale)), flush=True)
