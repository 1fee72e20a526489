"""The five workload configurations, each run with its real settings on the
GPU and checked bit for bit against oracle scores committed under
tests/golden/ (written by tests/golden/make_config_fixtures.py with
orc_solve_ref, the oracle's restatement of solver.cu:54-283).

The fixtures hold no boards: the GPU generator reproduces the oracle's
boards, which each test first checks through the boards' SHA-256.

  C1  10,000 x 10 empties: the plain kernel and the default solver
  C2  1,000,000 x 12 empties: the plain batched kernel (one lane per board,
      static corner / X-square order, no table, no help)
  C3  the first 20,000 of the 10..18-empties mix: autotuned configuration,
      one persistent launch per solve captured in a HIP graph and replayed
  C4  100,000 x 16 empties: the headline solver (learned order, 2^31-bucket
      table from 8 empties, 40 boards per wavefront); learned order without
      a table on the first 8,192
  C5  20 empties with a 2^26-bucket table from 10 empties, learned order
"""
import hashlib
import os
import sys

import numpy as np
import pytest

import gpuothello2_amd as go2
from gpuothello2_amd import autotune, learn

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TIME_LIMIT_MS = 100000  # every launch here ends in seconds; a slower one fails instead of hanging


def log(msg):
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


def fixture(name):
    path = os.path.join(GOLD, name)
    assert os.path.exists(path), f"{name} missing: run tests/golden/make_config_fixtures.py"
    return np.load(path)


def boards_of(fx):
    """The fixture's boards from the GPU generator, checked against its digest."""
    b = go2.generate(int(fx["seed"]), int(fx["config_id"]), int(fx["n"]), int(fx["empties_min"]),
                     int(fx["empties_max"]))
    digest = hashlib.sha256(np.ascontiguousarray(b).view("<u8").tobytes()).hexdigest()
    assert digest == str(fx["boards_sha256"]), "GPU generator differs from the oracle's"
    return b


def plain_config(**kw):
    """The plain batched kernel: one lane per board, static order, no table,
    no intra- or cross-wavefront help, no leaf batches (each lane scores its
    own leaves)."""
    return go2.default_config(ordering=go2.ORDER_STATIC, positions_per_wave=64, split_min_empties=99,
                              export_min_empties=99, leaf_batch_min=65, tt_log2_buckets=0,
                              time_limit_ms=TIME_LIMIT_MS, **kw)


def solve_task(b, depth, cfg, launches=1):
    t = go2.Task(len(b), depth, cfg)
    try:
        t.upload(b)
        out = []
        for _ in range(launches):
            t.launch()
            out.append(t.download())
        return out
    finally:
        t.close()


def assert_equal_scores(got, want, label):
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, f"{label}: {len(bad)} of {len(want)} scores differ, first at index {bad[:5]}"


@pytest.fixture(scope="module")
def evaluator():
    return learn.install()


def test_c1_ten_empties():
    fx = fixture("c1_10e.npz")
    b = boards_of(fx)
    for label, cfg in (("plain", plain_config()), ("default", go2.default_config(time_limit_ms=TIME_LIMIT_MS))):
        s, st = go2.solve(b, 10, cfg)
        log(f"C1 {label}: {len(b)} boards, {st['kernel_ms']:.1f} ms")
        assert_equal_scores(s, fx["score"], f"C1 {label}")


def test_c2_plain_batched_kernel_full_size():
    fx = fixture("c2_12e.npz")
    b = boards_of(fx)
    assert len(b) == 1_000_000
    [(s, st)] = solve_task(b, 12, plain_config())
    log(f"C2 plain: {len(b)} boards, {st['kernel_ms']:.0f} ms, {st['nodes'] / len(b):.0f} nodes/board")
    assert st["steals"] == 0 and st["exports"] == 0 and st["tt_probes"] == 0  # really the plain kernel
    assert_equal_scores(s, fx["score"], "C2")


def test_c3_uneven_mix_autotuned_graph(evaluator):
    fx = fixture("c3_mix.npz")
    b = boards_of(fx)
    idx = fx["index"]
    e = fx["empties"]
    for lv in range(10, 19):
        assert (e == lv).sum() >= 200, lv  # every level of the mix is checked
    base = go2.default_config(ordering=go2.ORDER_LEARNED, time_limit_ms=TIME_LIMIT_MS)
    tuned = autotune.autotune(b, 18, base, sample=3000, seed=1)
    log(f"C3 tuned: {tuned['params']}")
    cfg = autotune.apply(base, tuned["params"])
    cfg.graph_launch = 1
    runs = solve_task(b, 18, cfg, launches=2)  # capture, then replay
    for k, (s, st) in enumerate(runs):
        log(f"C3 launch {k}: {len(b)} boards, {st['kernel_ms']:.0f} ms")
        assert_equal_scores(s[idx], fx["score"], f"C3 launch {k}")


def test_c4_headline_solver_full_size(evaluator):
    fx = fixture("c4_16e.npz")
    b = boards_of(fx)
    assert len(b) == 100_000
    cfg = go2.default_config(ordering=go2.ORDER_LEARNED, tt_log2_buckets=31, tt_min_empties=8,
                             positions_per_wave=40, time_limit_ms=TIME_LIMIT_MS)
    [(s, st)] = solve_task(b, 16, cfg)
    log(f"C4 headline: {st['kernel_ms']:.0f} ms, hits {st['tt_hits']} of {st['tt_probes']}")
    assert st["tt_probes"] > 0 and st["tt_hits"] > 0
    assert_equal_scores(s, fx["score"], "C4 headline")
    k = 8192
    s2, st2 = go2.solve(b[:k], 16, go2.default_config(ordering=go2.ORDER_LEARNED, time_limit_ms=TIME_LIMIT_MS))
    log(f"C4 learned, no table: {k} boards, {st2['kernel_ms']:.0f} ms")
    assert_equal_scores(s2, fx["score"][:k], "C4 learned no table")


def test_c5_twenty_empties_with_table(evaluator):
    fx = fixture("c5_20e.npz")
    b = boards_of(fx)
    for label, extra in (("auto lanes", {}), ("16 lanes per board", {"positions_per_wave": 4})):
        cfg = go2.default_config(ordering=go2.ORDER_LEARNED, tt_log2_buckets=26, tt_min_empties=10,
                                 time_limit_ms=TIME_LIMIT_MS, **extra)
        s, st = go2.solve(b, 20, cfg)
        log(f"C5 {label}: {len(b)} boards, {st['kernel_ms']:.0f} ms, hit rate "
            f"{st['tt_hits'] / max(1, st['tt_probes']):.3f}")
        assert st["tt_hits"] > 0
        assert_equal_scores(s, fx["score"], f"C5 {label}")
