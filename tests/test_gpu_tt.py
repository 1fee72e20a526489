"""Transposition table: any table size, threshold or generation must leave
every exact score unchanged; counters report real use."""
import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def cfg(**kw):
    kw.setdefault("time_limit_ms", 100000)
    return go2.default_config(**kw)


@pytest.fixture(scope="module")
def boards():
    b = np.concatenate([ol.generate(901, 2, 400, 12), ol.generate(902, 2, 80, 14)])
    b = np.concatenate([b, b[:40]])  # duplicates share table entries
    want, _, _ = ol.solve_batch(b, 14, threads=16)
    return b, want


@pytest.fixture(scope="module")
def evaluator():
    from gpuothello2_amd import learn
    return learn.install()


@pytest.mark.parametrize("log2,tmin", [(10, 6), (16, 8), (22, 10), (24, 12)])
def test_scores_exact_with_table(boards, evaluator, log2, tmin):
    b, want = boards
    for ordering in (go2.ORDER_STATIC, go2.ORDER_MOBILITY, go2.ORDER_LEARNED):
        got, st = go2.solve(go2.as_boards(b), 14, cfg(ordering=ordering, tt_log2_buckets=log2,
                                                      tt_min_empties=tmin, split_min_empties=tmin))
        assert np.array_equal(got, want), (log2, tmin, ordering)
        assert st["tt_probes"] > 0 and st["tt_stores"] > 0
        assert st["tt_hits"] <= st["tt_probes"] and st["tt_cuts"] <= st["tt_hits"]


def test_generations_and_reuse(boards):
    b, want = boards
    t = go2.Task(len(b), 14, cfg(tt_log2_buckets=20, tt_min_empties=8, split_min_empties=8))
    t.upload(go2.as_boards(b))
    hits = []
    for _ in range(3):  # later launches must not read earlier launches' entries
        t.launch()
        s, st = t.download()
        assert np.array_equal(s, want)
        hits.append(st["tt_hits"])
    t.close()
    assert max(hits) > 0


def test_table_saves_nodes_at_depth():
    b = go2.generate(4040, 4, 256, 16)
    base, st0 = go2.solve(b, 16, cfg(ordering=go2.ORDER_MOBILITY))
    tt, st1 = go2.solve(b, 16, cfg(ordering=go2.ORDER_MOBILITY, tt_log2_buckets=22, tt_min_empties=10))
    assert np.array_equal(base, tt)
    assert st1["tt_cuts"] > 0
    assert st1["nodes"] < st0["nodes"]
