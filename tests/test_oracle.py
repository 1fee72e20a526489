"""Pins the CPU oracle: against the reference's own host codec (golden
vectors written by to_board.cpp compiled unchanged), against brute force, and
against what the game's mathematics fixes.  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MASK = (1 << 64) - 1


@pytest.fixture(scope="module")
def O():
    return ol.oracle()


@pytest.fixture(scope="module")
def refvec():
    return np.load(os.path.join(GOLD, "ref_bitboard.npz"))


def _boards_for_bitboard_checks():
    rng = np.random.default_rng(20261203)
    out = []
    for i in range(300):
        x = int(rng.integers(0, 2**63)) << 1 | (i & 1)
        y = int(rng.integers(0, 2**63)) << 1
        p, o = x & y & MASK, x & ~y & MASK
        out.append((p, o))
    for e in (4, 10, 16, 24, 36, 50):
        out += [(int(b[0]), int(b[1])) for b in ol.generate(77, 9, 40, e)]
    return out


def test_flip_equals_ray_walk(O):
    # orc_flip (ray tables) against orc_flip_walk (a square-by-square walk
    # along each of the eight directions) on every empty square.
    bad = 0
    for p, o in _boards_for_bitboard_checks():
        for sq in range(64):
            if ((p | o) >> sq) & 1:
                continue
            bad += O.orc_flip(p, o, sq) != O.orc_flip_walk(p, o, sq)
    assert bad == 0


def test_mobility_equals_ray_walk(O):
    for p, o in _boards_for_bitboard_checks():
        assert O.orc_mobility(p, o) == O.orc_mobility_walk(p, o)


def test_bitboard_known_answers(O):
    # Bit i = square i, rows of 8.  Standard start position, first player:
    # four legal moves, and the move at bit 19 flips the one disc at bit 27.
    p, o = 0x0000000810000000, 0x0000001008000000
    for f in (O.orc_mobility, O.orc_mobility_walk):
        assert f(p, o) == 0x0000102004080000
    for f in (O.orc_flip, O.orc_flip_walk):
        assert f(p, o, 19) == 1 << 27
        # a full row: own disc at bit 0, six opponents, move at bit 7
        assert f(1, 0x7E, 7) == 0x7E
        # no own disc at the end of the run: nothing flips
        assert f(0, 0x7E, 7) == 0
        # a run that would wrap from bit 7 to bit 8 is not a line
        assert f(1 << 8, 1 << 7, 6) == 0
        # diagonal: own disc at bit 0, opponents at 9 and 18, move at 27
        assert f(1, (1 << 9) | (1 << 18), 27) == (1 << 9) | (1 << 18)


def test_base81_matches_reference_vectors(refvec):
    for p, o, t in zip(refvec["player"], refvec["opponent"], refvec["text"]):
        assert ol.encode(int(p), int(o)) == str(t)
        assert ol.decode(str(t)) == (int(p), int(o))


def test_base81_closed_form():
    # test.cpp:7-19 of the reference: char i = 33 + sum_j w_j * digit(4i+j), w = 1,3,9,32.
    rng = np.random.default_rng(3)
    w = [1, 3, 9, 32]
    for _ in range(300):
        x, y = (int(v) for v in rng.integers(0, 2**63, 2)), None
        x = int(rng.integers(0, 2**63)) << 1 | 1
        y = int(rng.integers(0, 2**63)) << 1
        p, o = x & y & MASK, x & ~y & MASK
        s = "".join(chr(33 + sum(w[j] * (((p >> (4 * i + j)) & 1) + 2 * ((o >> (4 * i + j)) & 1))
                                 for j in range(4))) for i in range(16))
        assert ol.encode(p, o) == s
        assert ol.decode(s) == (p, o)


def test_base81_rejects_bad_chars():
    assert ol.decode(" " * 16) is None          # below '!'
    assert ol.decode("!" * 15 + "|") is None    # 91 > 90
    assert ol.decode("!" * 15 + "=") is None    # low part 28 > 26


def test_final_score_closed_form(O):
    # Winner takes the empties; exhaustive over disc counts.
    for pc in range(0, 65):
        for oc in range(0, 65 - pc):
            p = (1 << pc) - 1
            o = ((1 << oc) - 1) << pc
            want = 0 if pc == oc else (64 - 2 * oc if pc > oc else 2 * pc - 64)
            assert O.orc_final_score(p, o) == want


def test_next_bit_order(O):
    # Reference node.cuh:21-28: corners first, then ascending square.
    assert O.orc_next_bit(0x8100000000000081 | 0x10) == 0x1
    assert O.orc_next_bit(0x8000000000000000 | 0x10) == 0x8000000000000000
    assert O.orc_next_bit(0x30) == 0x10


def test_iteration_order_closed_form(O):
    # Reference node.cuh:21-28 (MobilityGenerator::next_bit): the unchecked
    # squares come out as the corners in ascending order, then every other
    # square in ascending order.  Checked as a sort over the squares.
    corners = {0, 7, 56, 63}
    for b in ol.generate(5, 3, 50, 20):
        p, o = int(b[0]), int(b[1])
        untried = ~(p | o) & MASK
        want = sorted((sq for sq in range(64) if (untried >> sq) & 1), key=lambda q: (q not in corners, q))
        ours = []
        while untried:
            bit = O.orc_next_bit(untried)
            untried ^= bit
            ours.append(bit.bit_length() - 1)
        assert ours == want


def test_small_scores_equal_minimax(O):
    d = np.load(os.path.join(GOLD, "solve_small.npz"))
    for p, o, s in zip(d["player"], d["opponent"], d["score"]):
        p, o = int(p), int(o)
        assert ol.minimax(p, o) == int(s)
    sc, _, _ = ol.solve_batch(np.stack([d["player"], d["opponent"]], 1), 10)
    assert np.array_equal(sc, d["score"])


def test_mid_fixture_scores(O):
    d = np.load(os.path.join(GOLD, "solve_mid.npz"))
    k = 60  # a slice keeps the CPU suite fast
    sc, _, _ = ol.solve_batch(np.stack([d["player"][:k], d["opponent"][:k]], 1), 14)
    assert np.array_equal(sc, d["score"][:k])


def test_scores_are_even_and_in_range():
    d = np.load(os.path.join(GOLD, "solve_mid.npz"))
    s = d["score"].astype(int)
    assert np.all(np.abs(s) <= 64) and np.all(s % 2 == 0)


def test_one_empty_closed_form(O):
    # One empty: the mover plays it if legal, else the other side, else game over.
    rng = np.random.default_rng(4)
    for b in ol.generate(8, 5, 60, 1):
        p, o = int(b[0]), int(b[1])
        sq = (~(p | o) & MASK).bit_length() - 1
        f = O.orc_flip(p, o, sq)
        if f:
            want = O.orc_final_score(o ^ f, p ^ f | 1 << sq)
            want = -want
        else:
            g = O.orc_flip(o, p, sq)
            want = O.orc_final_score(p ^ g, o ^ g | 1 << sq) if g else O.orc_final_score(p, o)
        assert ol.minimax(p, o) == want
        assert ol.solve_batch(np.array([[p, o]], dtype=np.uint64), 10)[0][0] == want


def test_readme_boards_decode_and_score(O):
    with open(os.path.join(GOLD, "readme_boards.json")) as f:
        rows = json.load(f)
    assert len(rows) == 9
    for r in rows:
        p, o = int(r["player"]), int(r["opponent"])
        assert p & o == 0
        assert ol.decode(r["text"]) == (p, o)
        assert ol.empties(p, o) == r["empties"]
        assert ol.minimax(p, o) == r["score"]


def test_generator_is_deterministic_and_exact():
    a = ol.generate(1, 4, 30, 14)
    b = ol.generate(1, 4, 30, 14)
    assert np.array_equal(a, b)
    for p, o in a:
        assert int(p) & int(o) == 0
        assert ol.empties(int(p), int(o)) == 14
    c = ol.generate(2, 4, 30, 14)
    assert not np.array_equal(a, c)


def test_planted_mistake_is_caught(O):
    # A search that forgets to negate a child's value must disagree with the fixtures.
    d = np.load(os.path.join(GOLD, "solve_small.npz"))
    wrong = 0
    for p, o, s in zip(d["player"][:80], d["opponent"][:80], d["score"][:80]):
        p, o = int(p), int(o)
        m = O.orc_mobility(p, o)
        if not m:
            continue
        best = -65
        while m:
            bit = m & -m
            m ^= bit
            f = O.orc_flip(p, o, bit.bit_length() - 1)
            best = max(best, ol.minimax(o ^ f, (p ^ f) | bit))  # sign dropped on purpose
        wrong += best != int(s)
    assert wrong > 0


def test_best_moves_equal_brute_force(O):
    # Optimal squares by definition: legal moves whose child's no-pruning
    # minimax value, negated, equals the position's value.
    d = np.load(os.path.join(GOLD, "solve_small.npz"))
    checked = {ol.MOVES_PLAY: 0, ol.MOVES_PASS: 0, ol.MOVES_GAME_OVER: 0}
    keep = d["empties"] <= 8
    for p, o, s in zip(d["player"][keep], d["opponent"][keep], d["score"][keep]):
        p, o = int(p), int(o)
        score, mask, kind = ol.best_moves(p, o)
        assert score == int(s)
        m = O.orc_mobility(p, o)
        if m:
            assert kind == ol.MOVES_PLAY
            want = 0
            while m:
                bit = m & -m
                m ^= bit
                f = O.orc_flip(p, o, bit.bit_length() - 1)
                if -ol.minimax(o ^ f, (p ^ f) | bit) == score:
                    want |= bit
            assert mask == want and want != 0
        else:
            assert mask == 0
            assert kind == (ol.MOVES_PASS if O.orc_mobility(o, p) else ol.MOVES_GAME_OVER)
        checked[kind] += 1
    assert checked[ol.MOVES_PLAY] > 100 and checked[ol.MOVES_PASS] > 0 and checked[ol.MOVES_GAME_OVER] > 0


def test_best_moves_batch_matches_single():
    b = ol.generate(31, 7, 40, 11)
    sc, masks, kinds = ol.best_moves_batch(b, 11, threads=4)
    ref, _, _ = ol.solve_batch(b, 11)
    assert np.array_equal(sc, ref)
    for i in range(len(b)):
        s, m, k = ol.best_moves(int(b[i, 0]), int(b[i, 1]), 11)
        assert (s, m, k) == (int(sc[i]), int(masks[i]), int(kinds[i]))
