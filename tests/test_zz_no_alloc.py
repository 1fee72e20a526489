"""The no-allocation entry point (go2_solve_device_noalloc) against the
allocating one it mirrors (go2_solve_device), through the C ABI: equal
scores on two calls in a row, the second with a smaller input, and each
best move among the oracle's optimal ones."""
import ctypes

import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol
from gpuothello2_amd.hip import DeviceBuffer

pytestmark = pytest.mark.gpu


def device_boards(b):
    arr = go2.as_boards(b)
    d = DeviceBuffer(arr.nbytes)
    d.upload(arr)
    return d


def run_noalloc(d_b, n, depth, cfg=None, moves=False):
    d_s = DeviceBuffer(n)
    d_m = DeviceBuffer(n) if moves else None
    st = go2.solve_device_noalloc(d_b.addr, n, depth, d_s.addr, d_m.addr if moves else None, cfg)
    s = np.zeros(n, dtype=np.int8)
    d_s.download(s)
    m = None
    if moves:
        m = np.zeros(n, dtype=np.uint8)
        d_m.download(m)
    return s, m, st


def run_alloc(d_b, n, depth, cfg=None):
    d_s = DeviceBuffer(n)
    go2.solve_device(d_b.addr, n, depth, d_s.addr, cfg)
    s = np.zeros(n, dtype=np.int8)
    d_s.download(s)
    return s


def test_noalloc_matches_allocating_entry_point_two_calls():
    big = np.concatenate([ol.generate(501, 8, 2000, 12), ol.generate(502, 8, 600, 14)])
    small = ol.generate(503, 8, 700, 11)
    want_big, _, _ = ol.solve_batch(big, 14, threads=16)
    want_small, _, _ = ol.solve_batch(small, 11, threads=16)
    cfg = go2.default_config(time_limit_ms=100000)
    d_big, d_small = device_boards(big), device_boards(small)
    # first call, then a smaller one: the second reuses the array the first left behind
    s1, _, st1 = run_noalloc(d_big, len(big), 14, cfg)
    s2, _, st2 = run_noalloc(d_small, len(small), 11, cfg)
    assert np.array_equal(s1, run_alloc(d_big, len(big), 14, cfg))
    assert np.array_equal(s2, run_alloc(d_small, len(small), 11, cfg))
    assert np.array_equal(s1, want_big) and np.array_equal(s2, want_small)
    assert st1["positions"] == len(big) and st2["positions"] == len(small)
    assert st1["kernel_ms"] > 0 and st2["kernel_ms"] > 0


def test_noalloc_best_moves_and_orderings():
    from gpuothello2_amd import learn
    learn.install()
    b = ol.generate(504, 8, 1500, 13)
    want, masks, kinds = ol.best_moves_batch(b, 13, threads=16)
    d_b = device_boards(b)
    for ordering in (go2.ORDER_STATIC, go2.ORDER_MOBILITY, go2.ORDER_LEARNED):
        cfg = go2.default_config(ordering=ordering, time_limit_ms=100000)
        s, m, _ = run_noalloc(d_b, len(b), 13, cfg, moves=True)
        assert np.array_equal(s, want), ordering
        assert np.array_equal(s, run_alloc(d_b, len(b), 13, cfg)), ordering
        for i in range(len(b)):
            if kinds[i] == ol.MOVES_PLAY:
                assert (int(masks[i]) >> int(m[i])) & 1, (ordering, i)
            else:
                assert m[i] == (go2.MOVE_PASS if kinds[i] == ol.MOVES_PASS else go2.MOVE_NONE)


def test_noalloc_deep_depth_shrinks_grid_and_rejects_table():
    b = ol.generate(505, 8, 500, 10)
    want, _, _ = ol.solve_batch(b, 10, threads=16)
    d_b = device_boards(b)
    for depth in (24, 32, 60):  # stacks beyond the full grid's share of the scratch
        s, _, _ = run_noalloc(d_b, len(b), depth, go2.default_config(time_limit_ms=100000))
        assert np.array_equal(s, want), depth
    d_s = DeviceBuffer(len(b))
    with pytest.raises(go2.Go2Error) as e:
        go2.solve_device_noalloc(d_b.addr, len(b), 10, d_s.addr, None, go2.default_config(tt_log2_buckets=20))
    assert e.value.code == go2.GO2_E_ARG
    assert ctypes.c_int(go2.lib().go2_solve_device_noalloc(None, 0, 10, None, None, None, None, None)).value == 0
