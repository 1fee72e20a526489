"""The oracle's pattern evaluator pinned against the reference's index rule
(eval.cu:113-121: base3[pext(me, mask)] + 2*base3[pext(op, mask)]), the
reference's own PEXT (eval_host.cpp, when its sources are here) and the
calculus of the SGD step.  CPU only."""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol

MASK = (1 << 64) - 1


def sym_map(s, sq):
    r, c = sq >> 3, sq & 7
    if s & 4:
        r, c = c, r
    if s & 1:
        c = 7 - c
    if s & 2:
        r = 7 - r
    return r * 8 + c


def sym_board(x, s):
    y = 0
    for q in range(64):
        if (x >> sym_map(s, q)) & 1:
            y |= 1 << q
    return y


@pytest.fixture(scope="module")
def layout():
    L = ol.oracle()
    cls = (ctypes.c_int * 46)()
    nc = (ctypes.c_int * 46)()
    cells = (ctypes.c_int * 460)()
    off = (ctypes.c_int * 12)()
    nw = L.orc_pattern_layout(cls, nc, cells, off)
    inst = [(cls[i], [cells[i * 10 + j] for j in range(nc[i])]) for i in range(46)]
    return nw, list(off), inst


def boards(n, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        x, y = (int(v) for v in rng.integers(0, 2**63, 2))
        x = (x << 1 | 1) & MASK
        y = (y << 1) & MASK
        out.append((x & y, x & ~y & MASK))
    return out + [tuple(int(v) for v in b) for b in ol.generate(3, 9, n, 20)]


def test_layout(layout):
    nw, off, inst = layout
    assert nw == 167265 and off[-1] == nw
    assert len({tuple(sorted(c)) for _, c in inst}) == 46  # distinct cell sets
    sizes = [len(c) for _, c in inst]
    assert max(sizes) == 10 and min(sizes) == 4
    for k in range(11):
        n = [len(c) for kk, c in inst if kk == k][0]
        assert off[k + 1] - off[k] == 3 ** n


def test_index_equals_reference_rule(layout):
    """feature - class offset == base3(pext(me)) + 2*base3(pext(op)) over the
    instance's cells, with the board mapped so the cells come in base order."""
    L = ol.oracle()
    pext = ol.reference().ref_pext if ol.ref_available() else L.orc_pext
    nw, off, inst = layout
    base = {}
    for k, c in inst:
        base.setdefault(k, c)
    feat = (ctypes.c_uint32 * 46)()
    for p, o in boards(150):
        L.orc_pattern_indices(p, o, feat)
        for i, (k, cells) in enumerate(inst):
            s = [s for s in range(8) if [sym_map(s, q) for q in base[k]] == cells][0]
            mask = sum(1 << q for q in base[k])
            assert L.orc_ref_pattern_index(sym_board(p, s), sym_board(o, s), mask) == feat[i] - off[k]
            a, b = pext(sym_board(p, s), mask), pext(sym_board(o, s), mask)
            t3 = lambda v: sum(((v >> j) & 1) * 3 ** j for j in range(10))  # noqa: E731
            assert t3(a) + 2 * t3(b) == feat[i] - off[k]


@pytest.mark.skipif(not ol.ref_available(), reason="reference host build (oracle/_ref) not present")
def test_pext_matches_reference():
    L, R = ol.oracle(), ol.reference()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x, m = (int(v) for v in rng.integers(0, 2**63, 2))
        assert L.orc_pext(x, m) == R.ref_pext(x, m)


def test_eval_int_is_sum_of_tables(layout):
    L = ol.oracle()
    nw = layout[0]
    rng = np.random.default_rng(2)
    w = rng.integers(-300, 300, nw).astype(np.int16)
    b = rng.integers(-300, 300, 64).astype(np.int16)
    feat = (ctypes.c_uint32 * 46)()
    for p, o in boards(40):
        L.orc_pattern_indices(p, o, feat)
        want = int(b[64 - bin(p | o).count("1")]) + sum(int(w[f]) for f in feat)
        assert L.orc_eval_int(w.ctypes.data, b.ctypes.data, p, o) == want


def test_sgd_gradient_is_derivative_of_squared_error(layout):
    """grad_f = d/dw_f sum_i (pred_i - y_i)^2, checked by finite differences."""
    L = ol.oracle()
    nw = layout[0]
    bs = boards(60)
    n = len(bs)
    feat = np.zeros((n, 46), dtype=np.uint32)
    f1 = (ctypes.c_uint32 * 46)()
    emp = np.zeros(n, dtype=np.uint8)
    for i, (p, o) in enumerate(bs):
        L.orc_pattern_indices(p, o, f1)
        feat[i] = list(f1)
        emp[i] = 64 - bin(p | o).count("1")
    rng = np.random.default_rng(3)
    y = rng.normal(0, 20, n).astype(np.float32)
    w0 = rng.normal(0, 1, nw).astype(np.float32)
    b0 = rng.normal(0, 1, 64).astype(np.float32)
    grad = np.zeros(nw + 64, dtype=np.float32)
    cnt = np.zeros(nw + 64, dtype=np.uint32)
    w, b = w0.copy(), b0.copy()
    mse = L.orc_sgd_step(w.ctypes.data, b.ctypes.data, feat.ctypes.data, emp.ctypes.data, y.ctypes.data, n,
                         ctypes.c_float(0.0), grad.ctypes.data, cnt.ctypes.data)

    def sse(wv, bv):
        pred = wv[feat].astype(np.float64).sum(1) + bv[emp]
        return float(((pred - y) ** 2).sum())

    assert abs(mse - sse(w0, b0) / n) < 1e-3 * max(1.0, mse)
    for f in list(np.unique(feat[:3]))[:10] + [nw + int(emp[0])]:
        h = 1e-2
        wp, bp = w0.astype(np.float64).copy(), b0.astype(np.float64).copy()
        if f < nw:
            wp[f] += h
        else:
            bp[f - nw] += h
        num = (sse(wp, bp) - sse(w0, b0)) / h
        assert abs(num - grad[f]) < 0.05 * max(1.0, abs(num)), (f, num, grad[f])
        occ = int((feat == f).sum()) if f < nw else int((emp == f - nw).sum())
        assert cnt[f] == occ
    # the applied step moves each seen weight by lr * grad / count
    w2, b2 = w0.copy(), b0.copy()
    L.orc_sgd_step(w2.ctypes.data, b2.ctypes.data, feat.ctypes.data, emp.ctypes.data, y.ctypes.data, n,
                   ctypes.c_float(0.25), None, None)
    seen = cnt[:nw] > 0
    np.testing.assert_allclose(w2[seen], w0[seen] - 0.25 * grad[:nw][seen] / cnt[:nw][seen], rtol=1e-5, atol=1e-5)
    assert np.array_equal(w2[~seen], w0[~seen])
