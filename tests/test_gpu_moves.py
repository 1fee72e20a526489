"""Best moves of the exact solve, through the C ABI.  A returned square must
be one of the oracle's optimal squares (its child solves to the board's
exact score); a side without a move gets MOVE_PASS, a finished game
MOVE_NONE.  Scores stay bit-exact."""
import os

import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def check_moves(boards: np.ndarray, depth: int, scores, moves):
    want, masks, kinds = ol.best_moves_batch(boards, depth, threads=16)
    assert np.array_equal(scores, want)
    for i in range(len(boards)):
        k, m = int(kinds[i]), int(moves[i])
        if k == ol.MOVES_PASS:
            assert m == go2.MOVE_PASS, i
        elif k == ol.MOVES_GAME_OVER:
            assert m == go2.MOVE_NONE, i
        else:
            assert m < 64 and (int(masks[i]) >> m) & 1, (i, m, hex(int(masks[i])))
    return kinds


def mixed_boards():
    d = np.load(os.path.join(GOLD, "solve_small.npz"))
    small = np.stack([d["player"], d["opponent"]], 1)  # 0..10 empties, passes, finished games
    return np.concatenate([small, ol.generate(909, 5, 300, 12), ol.generate(910, 5, 40, 14)])


@pytest.mark.parametrize("ordering", [go2.ORDER_STATIC, go2.ORDER_MOBILITY])
def test_best_moves_are_optimal(ordering):
    b = mixed_boards()
    s, m, _ = go2.solve_moves(go2.as_boards(b), 14, go2.default_config(ordering=ordering))
    kinds = check_moves(b, 14, s, m)
    assert (kinds == ol.MOVES_PASS).any() and (kinds == ol.MOVES_GAME_OVER).any()


@pytest.mark.parametrize("kw", [
    dict(positions_per_wave=1, split_min_empties=5),                     # 63 helper lanes per board
    dict(positions_per_wave=2, split_min_empties=6, export_min_empties=10,
         export_low_empties=6),                                          # helpers in other wavefronts
    dict(tt_log2_buckets=20, tt_min_empties=6, split_min_empties=6),     # table cuts and narrowed windows
])
def test_best_moves_with_helpers_and_table(kw):
    b = np.concatenate([ol.generate(911, 5, 200, 12), ol.generate(912, 5, 60, 14)])
    s, m, st = go2.solve_moves(go2.as_boards(b), 14, go2.default_config(ordering=go2.ORDER_MOBILITY, **kw))
    check_moves(b, 14, s, m)
    assert st["steals"] > 0


def test_best_moves_deep_batch_property():
    # 16 empties, learned ordering when the table is present, whole-GPU help:
    # the child of each returned move, solved again, has the negated score.
    b = go2.generate(913, 5, 512, 16)
    cfg = go2.default_config(ordering=go2.ORDER_MOBILITY)
    s, m, st = go2.solve_moves(b, 16, cfg)
    assert st["exports"] > 0
    P, O = b["player"].astype(np.uint64), b["opponent"].astype(np.uint64)
    L = ol.oracle()
    child = np.zeros(len(b), dtype=go2.BOARD_DTYPE)
    for i in range(len(b)):
        p, o, sq = int(P[i]), int(O[i]), int(m[i])
        assert sq < 64 and (L.orc_mobility(p, o) >> sq) & 1, i
        f = L.orc_flip(p, o, sq)
        child[i] = (o ^ f, (p ^ f) | (1 << sq))
    cs, _ = go2.solve(child, 15, cfg)
    assert np.array_equal(-cs.astype(np.int32), s.astype(np.int32))


def test_task_moves_match_solve_moves():
    b = ol.generate(914, 5, 500, 11)
    t = go2.Task(len(b), 11)
    t.upload(go2.as_boards(b))
    t.launch()
    t.wait()
    s, _ = t.download()
    m = t.download_moves()
    t.close()
    check_moves(b, 11, s, m)


def test_cli_best_move_column(tmp_path):
    import subprocess
    b = ol.generate(915, 5, 200, 10)
    lines = [ol.encode(int(p), int(o)) for p, o in b]
    inp = tmp_path / "in.txt"
    inp.write_text(f"{len(lines)}\n" + "\n".join(lines) + "\n")
    out = tmp_path / "out.txt"
    env = dict(os.environ, GO2_BEST_MOVE="1")
    r = subprocess.run([go2.SOLVER_BIN, str(inp), str(out), "10"], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    rows = [ln.split() for ln in out.read_text().splitlines()]
    assert len(rows) == len({(int(p), int(o)) for p, o in b})
    for text, score, mv in rows:
        p, o = ol.decode(text)
        want, mask, kind = ol.best_moves(p, o, 10)
        assert int(score) == want
        if kind == ol.MOVES_PLAY:
            sq = (ord(mv[0]) - ord("A")) + 8 * (ord(mv[1]) - ord("1"))
            assert (mask >> sq) & 1 and go2.move_text(sq) == mv
        else:
            assert mv == ("ps" if kind == ol.MOVES_PASS else "--")


@pytest.mark.parametrize("width", [2, 4, 8, 30])
def test_aspiration_windows_exact(width):
    # Learned ordering searches each board first with the window estimate
    # +- width and once more with one bound open when the value falls outside;
    # narrow windows force that re-search on most boards.  Scores and best
    # moves must stay exact.
    from gpuothello2_amd import learn
    learn.install()
    b = np.concatenate([ol.generate(811, 12, 300, 12), ol.generate(812, 12, 120, 14)])
    want, masks, kinds = ol.best_moves_batch(b, 14, threads=16)
    cfg = go2.default_config(ordering=go2.ORDER_LEARNED, aspiration_width=width, time_limit_ms=100000)
    s, m, _ = go2.solve_moves(go2.as_boards(b), 14, cfg)
    assert np.array_equal(s, want), width
    for i in range(len(b)):
        if kinds[i] == ol.MOVES_PLAY:
            assert (int(masks[i]) >> int(m[i])) & 1, (width, i)


@pytest.mark.parametrize("pvs,tt,ordering", [(5, 0, "static"), (5, 20, "learned"), (8, 22, "learned"),
                                             (6, 0, "mobility")])
def test_null_window_search_exact(pvs, tt, ordering):
    # Principal variation search: later children get a null window first and
    # a full re-search when they fail high inside the window (also when the
    # table settles them with a bound), with helpers and the table on.
    from gpuothello2_amd import learn
    learn.install()
    b = np.concatenate([ol.generate(821, 12, 300, 12), ol.generate(822, 12, 120, 14)])
    want, masks, kinds = ol.best_moves_batch(b, 14, threads=16)
    order = {"static": go2.ORDER_STATIC, "mobility": go2.ORDER_MOBILITY, "learned": go2.ORDER_LEARNED}[ordering]
    cfg = go2.default_config(ordering=order, pvs_min_empties=pvs, tt_log2_buckets=tt, tt_min_empties=6,
                             split_min_empties=6, time_limit_ms=100000)
    s, m, st = go2.solve_moves(go2.as_boards(b), 14, cfg)
    assert np.array_equal(s, want), (pvs, tt, ordering)
    for i in range(len(b)):
        if kinds[i] == ol.MOVES_PLAY:
            assert (int(masks[i]) >> int(m[i])) & 1, (pvs, i)
