"""Think mode (the reference's depth-limited search, thinker.cu:120-320)
through the C ABI: every value equals the oracle's orc_think with the same
int16 evaluator, and every move is one of the oracle's optimal squares."""
import os
import subprocess

import numpy as np
import pytest

import gpuothello2_amd as go2
import oracle_lib as ol
from gpuothello2_amd import learn

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ev():
    w16, b16, scale, _ = learn.load()
    go2.set_evaluator(w16, b16, scale)
    return np.ascontiguousarray(w16, dtype=np.int16), np.ascontiguousarray(b16, dtype=np.int16), scale


def check(ev, boards, depth, s, m):
    w, b, scale = ev
    want, masks, kinds = ol.think_batch(w, b, scale, boards, depth, threads=16)
    assert np.array_equal(s, want)
    for i in range(len(boards)):
        k, mv = int(kinds[i]), int(m[i])
        if k == ol.MOVES_PASS:
            assert mv == go2.MOVE_PASS, i
        elif k == ol.MOVES_GAME_OVER:
            assert mv == go2.MOVE_NONE, i
        else:
            assert mv < 64 and (int(masks[i]) >> mv) & 1, (i, mv)
    return kinds


@pytest.mark.parametrize("depth,n", [(1, 400), (2, 400), (3, 300), (4, 200), (5, 40)])
def test_think_midgame_matches_oracle(ev, depth, n):
    b = np.concatenate([ol.generate(71, 11, n // 2, 30), ol.generate(72, 11, n - n // 2, 44)])
    s, m, _ = go2.think(go2.as_boards(b), depth)
    check(ev, b, depth, s, m)


@pytest.mark.parametrize("depth", [1, 2, 3, 4, 9, 10, 12])
def test_think_near_the_end(ev, depth):
    # 0..10 empties: passes, finished games, horizons before and past the end.
    d = np.load(os.path.join(GOLD, "solve_small.npz"))
    b = np.stack([d["player"], d["opponent"]], 1)
    s, m, _ = go2.think(go2.as_boards(b), depth)
    kinds = check(ev, b, depth, s, m)
    assert (kinds == ol.MOVES_PASS).any() and (kinds == ol.MOVES_GAME_OVER).any()
    if depth >= 11:
        exact, _, _ = ol.best_moves_batch(b, 10)
        assert np.array_equal(s, exact)


def test_think_rejects_bad_depth(ev):
    b = go2.as_boards(ol.generate(73, 11, 4, 30))
    for depth in (0, 31):
        with pytest.raises(go2.Go2Error) as e:
            go2.think(b, depth)
        assert e.value.code == go2.GO2_E_ARG


def test_cli_think_mode(ev, tmp_path):
    w, b, scale = ev
    boards = ol.generate(74, 11, 150, 34)
    lines = [ol.encode(int(p), int(o)) for p, o in boards]
    inp = tmp_path / "in.txt"
    inp.write_text(f"{len(lines)}\n" + "\n".join(lines) + "\n")
    out = tmp_path / "out.txt"
    wpath = tmp_path / "w.bin"
    learn.export_raw(str(wpath))
    env = dict(os.environ, GO2_WEIGHTS=str(wpath))
    r = subprocess.run([go2.SOLVER_BIN, str(inp), str(out), "60", "3"], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    rows = [ln.split() for ln in out.read_text().splitlines()]
    assert len(rows) == len({(int(p), int(o)) for p, o in boards})
    for text, score, mv in rows:
        p, o = ol.decode(text)
        want, mask, kind = ol.think(w, b, scale, p, o, 3)
        assert int(score) == want
        if kind == ol.MOVES_PLAY:
            sq = (ord(mv[0]) - ord("A")) + 8 * (ord(mv[1]) - ord("1"))
            assert (mask >> sq) & 1
        else:
            # the reference passes a finished game's root before scoring it
            # (thinker.cu:151-154,166-169): a pass either way
            assert mv == "ps"
