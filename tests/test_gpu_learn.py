"""GPU pattern evaluator and its training kernels against the CPU oracle,
and learned move ordering: exact scores, and what it saves over ordering
without the trained table."""
import ctypes
import os

import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol
from gpuothello2_amd.hip import DeviceBuffer

pytestmark = pytest.mark.gpu
NI, NW, NB = 46, 167265, 64


def sample_boards(n=3000):
    rng = np.random.default_rng(11)
    x = rng.integers(0, 2**63, n, dtype=np.int64).astype(np.uint64) << np.uint64(1) | np.uint64(1)
    y = rng.integers(0, 2**63, n, dtype=np.int64).astype(np.uint64) << np.uint64(1)
    rnd = np.stack([x & y, x & ~y], 1)
    play = np.concatenate([ol.generate(21, 5, n // 3, e) for e in (8, 14, 24)])
    return np.concatenate([rnd, play]).astype(np.uint64)


def gpu_features(b):
    n = len(b)
    db = DeviceBuffer(n * 16)
    db.upload(np.ascontiguousarray(b))
    feat, emp = DeviceBuffer(n * NI * 4), DeviceBuffer(n)
    assert go2.lib().go2_pattern_features_device(db.addr, n, feat.addr, emp.addr, None) == 0
    f = np.zeros((n, NI), dtype=np.uint32)
    e = np.zeros(n, dtype=np.uint8)
    feat.download(f)
    emp.download(e)
    return f, e, (db, feat, emp)


def oracle_features(b):
    L = ol.oracle()
    out = np.zeros((len(b), NI), dtype=np.uint32)
    for i, (p, o) in enumerate(b):
        L.orc_pattern_indices(int(p), int(o), out[i].ctypes.data)
    return out


def test_pattern_info():
    assert go2.pattern_info() == (NI, NW, NB)


def test_features_match_oracle():
    b = sample_boards()
    f, e, _ = gpu_features(b)
    assert np.array_equal(f, oracle_features(b))
    assert np.array_equal(e.astype(int), [64 - bin(int(p) | int(o)).count("1") for p, o in b])


def test_int16_eval_matches_oracle():
    b = sample_boards(1500)
    rng = np.random.default_rng(4)
    w = rng.integers(-2000, 2000, NW).astype(np.int16)
    bias = rng.integers(-2000, 2000, NB).astype(np.int16)
    dw, dbias, db, out = DeviceBuffer(NW * 2), DeviceBuffer(NB * 2), DeviceBuffer(len(b) * 16), DeviceBuffer(len(b) * 4)
    dw.upload(w)
    dbias.upload(bias)
    db.upload(np.ascontiguousarray(b))
    assert go2.lib().go2_pattern_eval_device(dw.addr, dbias.addr, db.addr, len(b), out.addr, None) == 0
    got = np.zeros(len(b), dtype=np.int32)
    out.download(got)
    L = ol.oracle()
    want = [L.orc_eval_int(w.ctypes.data, bias.ctypes.data, int(p), int(o)) for p, o in b]
    assert np.array_equal(got, want)


def test_sgd_gradient_and_step_match_oracle():
    b = sample_boards(2400)
    n = len(b)
    f, e, keep = gpu_features(b)
    rng = np.random.default_rng(5)
    y = rng.normal(0, 25, n).astype(np.float32)
    w0 = rng.normal(0, 2, NW).astype(np.float32)
    b0 = rng.normal(0, 2, NB).astype(np.float32)
    dw, dbias, dy = DeviceBuffer(NW * 4), DeviceBuffer(NB * 4), DeviceBuffer(n * 4)
    grad, cnt, sse = DeviceBuffer((NW + NB) * 4), DeviceBuffer((NW + NB) * 4), DeviceBuffer(8)
    dw.upload(w0)
    dbias.upload(b0)
    dy.upload(y)
    grad.upload(np.zeros(NW + NB, dtype=np.float32))
    cnt.upload(np.zeros(NW + NB, dtype=np.uint32))
    sse.upload(np.zeros(1))
    L = go2.lib()
    _, feat, emp = keep
    assert L.go2_sgd_gradient_device(dw.addr, dbias.addr, feat.addr, emp.addr, dy.addr, n, grad.addr,
                                     cnt.addr, sse.addr, None) == 0
    g_gpu = np.zeros(NW + NB, dtype=np.float32)
    c_gpu = np.zeros(NW + NB, dtype=np.uint32)
    s_gpu = np.zeros(1)
    grad.download(g_gpu)
    cnt.download(c_gpu)
    sse.download(s_gpu)
    # oracle: same batch, same start weights
    wo, bo = w0.copy(), b0.copy()
    g_orc = np.zeros(NW + NB, dtype=np.float32)
    c_orc = np.zeros(NW + NB, dtype=np.uint32)
    mse = ol.oracle().orc_sgd_step(wo.ctypes.data, bo.ctypes.data, f.ctypes.data, e.ctypes.data, y.ctypes.data,
                                   n, ctypes.c_float(0.3), g_orc.ctypes.data, c_orc.ctypes.data)
    assert np.array_equal(c_gpu, c_orc)
    # float atomics sum in a different order: tolerance relative to the gradient scale
    np.testing.assert_allclose(g_gpu, g_orc, rtol=1e-4, atol=1e-4 * float(np.abs(g_orc).max()))
    assert abs(s_gpu[0] / n - mse) < 1e-4 * mse
    assert L.go2_sgd_apply_device(dw.addr, dbias.addr, grad.addr, cnt.addr, ctypes.c_float(0.3), None) == 0
    w1 = np.zeros(NW, dtype=np.float32)
    b1 = np.zeros(NB, dtype=np.float32)
    dw.download(w1)
    dbias.download(b1)
    np.testing.assert_allclose(w1, wo, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(b1, bo, rtol=1e-4, atol=1e-4)
    grad.download(g_gpu)
    cnt.download(c_gpu)
    assert not g_gpu.any() and not c_gpu.any()  # cleared for the next batch


def test_learned_ordering_requires_an_evaluator_then_keeps_scores_exact():
    rng = np.random.default_rng(6)
    # any table at all must give the exact scores
    go2.set_evaluator(rng.integers(-3000, 3000, NW).astype(np.int16),
                      rng.integers(-3000, 3000, NB).astype(np.int16), 64.0)
    b = np.concatenate([ol.generate(71, 2, 400, 12), ol.generate(72, 2, 60, 14)])
    want, _, _ = ol.solve_batch(b, 14, threads=16)
    for lm in (8, 11, 14):
        got, _ = go2.solve(go2.as_boards(b), 14, go2.default_config(ordering=go2.ORDER_LEARNED,
                                                                     learned_min_empties=lm,
                                                                     time_limit_ms=100000))
        assert np.array_equal(got, want), lm


def test_trained_evaluator_gains_over_mobility_ordering():
    from gpuothello2_amd import learn
    if not os.path.exists(learn.WEIGHTS_PATH):
        pytest.skip("no trained evaluator file")
    meta = learn.install()
    assert meta["valid_mse"] < meta["valid_mse_constant_predictor"]
    b = go2.generate(555, 4, 2000, 16)
    cfg_m = go2.default_config(ordering=go2.ORDER_MOBILITY, time_limit_ms=100000)
    cfg_l = go2.default_config(ordering=go2.ORDER_LEARNED, time_limit_ms=100000)
    s_m, st_m = go2.solve(b, 16, cfg_m)
    s_l, st_l = go2.solve(b, 16, cfg_l)
    assert np.array_equal(s_m, s_l)
    assert st_l["nodes"] < st_m["nodes"]
