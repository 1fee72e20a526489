"""Pins the oracle's depth-limited search (the reference's think mode,
thinker.cu:120-320) by what fixes it without the reference: exactness once
the horizon passes the end of the game, the leaf rounding of the evaluator,
depth 1 in closed form, and passes that cost no depth.  CPU only."""
import math
import os

import numpy as np
import pytest

import oracle_lib as ol

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpuothello2_amd",
                       "weights", "pattern_int16.npz")


@pytest.fixture(scope="module")
def ev():
    d = np.load(WEIGHTS, allow_pickle=False)
    return (np.ascontiguousarray(d["w16"], dtype=np.int16), np.ascontiguousarray(d["b16"], dtype=np.int16),
            float(d["scale"]))


def children(O, p, o):
    m = O.orc_mobility(p, o)
    while m:
        bit = m & -m
        m ^= bit
        sq = bit.bit_length() - 1
        f = O.orc_flip(p, o, sq)
        yield sq, o ^ f, (p ^ f) | bit


def test_horizon_past_the_end_is_the_exact_solve(ev):
    # More plies than empty squares: the horizon is never reached (the
    # position after the depth-th move is scored by the evaluator even when
    # the board is full, thinker.cu:302-306, so depth == empties is not exact).
    w, b, s = ev
    boards = np.concatenate([ol.generate(41, 9, 60, 8), ol.generate(42, 9, 20, 10)])
    want, masks, kinds = ol.best_moves_batch(boards, 10)
    for depth in (11, 14):
        got, gm, gk = ol.think_batch(w, b, s, boards, depth)
        assert np.array_equal(got, want)
        assert np.array_equal(gm, masks) and np.array_equal(gk, kinds)


def test_depth_equal_to_empties_scores_full_boards_by_the_evaluator(ev):
    w, b, s = ev
    # Two empties, depth 2: the position after the second move is scored by
    # the evaluator (not by the final count), depth 3 is past the end.
    boards = ol.generate(47, 9, 40, 2)
    differ = 0
    for p, o in boards:
        p, o = int(p), int(o)
        v, _, _ = ol.think(w, b, s, p, o, 2)
        e, _, _ = ol.think(w, b, s, p, o, 3)
        x, _, _ = ol.best_moves(p, o, 2)
        assert e == x
        differ += v != x
    assert differ > 0


def test_leaf_is_the_rounded_clamped_evaluation(ev):
    w, b, s = ev
    O = ol.oracle()
    boards = ol.generate(43, 9, 300, 30)
    for p, o in boards:
        p, o = int(p), int(o)
        x = O.orc_eval_int(w.ctypes.data, b.ctypes.data, p, o)
        r = math.floor(abs(x) / s + 0.5) * (1 if x >= 0 else -1)  # half away from zero
        assert ol.think_leaf(w, b, s, p, o) == max(-64, min(64, r))


def test_depth_one_closed_form(ev):
    w, b, s = ev
    O = ol.oracle()
    boards = ol.generate(44, 9, 200, 36)
    for p, o in boards:
        p, o = int(p), int(o)
        v, mask, kind = ol.think(w, b, s, p, o, 1)
        vals = {sq: -ol.think_leaf(w, b, s, cp, co) for sq, cp, co in children(O, p, o)}
        assert kind == ol.MOVES_PLAY and v == max(vals.values())
        assert mask == sum(1 << sq for sq, x in vals.items() if x == v)


def test_pass_costs_no_depth(ev):
    # A side without a move hands the same depth to the opponent: the value
    # is the opponent's value at that depth, negated.
    w, b, s = ev
    O = ol.oracle()
    found = 0
    for i in range(4000):
        bd = ol.generate(45, 9, 1, 14, start=i)[0]
        p, o = int(bd[0]), int(bd[1])
        for _, cp, co in children(O, p, o):
            if not O.orc_mobility(cp, co) and O.orc_mobility(co, cp):
                for depth in (1, 2, 3):
                    v, mask, kind = ol.think(w, b, s, cp, co, depth)
                    u, _, _ = ol.think(w, b, s, co, cp, depth)
                    assert kind == ol.MOVES_PASS and mask == 0 and v == -u
                found += 1
        if found >= 5:
            break
    assert found >= 5


def test_planted_sign_error_is_caught(ev):
    # Depth 2 computed without negating the grandchildren's leaf scores
    # must disagree with the oracle.
    w, b, s = ev
    O = ol.oracle()
    wrong = 0
    for p, o in ol.generate(46, 9, 40, 30):
        p, o = int(p), int(o)
        v, _, _ = ol.think(w, b, s, p, o, 2)
        best = -65
        for _, cp, co in children(O, p, o):
            inner = [ol.think_leaf(w, b, s, gp, go) for _, gp, go in children(O, cp, co)]  # sign dropped
            if inner:
                best = max(best, -max(inner))
        wrong += best != v
    assert wrong > 0


def test_alpha_beta_value_equals_the_plain_recursion(ev):
    w, b, s = ev
    for depth, n, e in ((1, 100, 40), (2, 100, 36), (3, 80, 30), (4, 40, 24), (5, 20, 20), (9, 40, 8), (12, 20, 10)):
        for p, o in ol.generate(48 + depth, 9, n, e):
            p, o = int(p), int(o)
            assert ol.think_value(w, b, s, p, o, depth) == ol.think(w, b, s, p, o, depth)[0]
