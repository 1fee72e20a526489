"""Writes the golden fixtures under tests/golden/.

Run in the build container (where the reference sources exist) after
`make -C oracle && make -C oracle ref`:

    python tests/golden/make_golden.py

Fixtures:
  ref_bitboard.npz  base81 text of random and playout boards written by the
                    reference's own fromBoard (to_board.cpp, compiled unchanged
                    from the reference sources by oracle/Makefile).
  solve_small.npz   exact scores of boards with 0..10 empties: the oracle's
                    reference-method search, asserted equal to a no-pruning
                    minimax when written.
  solve_mid.npz     exact scores at 12 and 14 empties (reference-method search,
                    asserted invariant under the 8 board symmetries).
  readme_boards.json the nine input boards of the reference README
                    (README.md:11-19) decoded by the reference decoder, with
                    their exact scores (no-pruning minimax).
"""
from __future__ import annotations

import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402

MASK = (1 << 64) - 1


def sym_images(p: int, o: int):
    """The 8 images of a board under the square's symmetry group."""
    def tr(x, s):
        r = 0
        for sq in range(64):
            if (x >> sq) & 1:
                row, col = sq >> 3, sq & 7
                if s & 4:
                    row, col = col, row
                if s & 1:
                    col = 7 - col
                if s & 2:
                    row = 7 - row
                r |= 1 << (row * 8 + col)
        return r
    return [(tr(p, s), tr(o, s)) for s in range(8)]


def random_boards(rng, n):
    out = []
    for i in range(n):
        x, y, z = (int(v) for v in rng.integers(0, 2**63, size=3, dtype=np.int64))
        x = (x << 1 | (i & 1)) & MASK
        y = (y << 1) & MASK
        p, o = x & y, x & ~y & MASK
        if i % 3 == 1:
            p &= z
        if i % 3 == 2:
            o &= z
        out.append((p, o))
    return out


def main():
    if not ol.ref_available():
        raise SystemExit("oracle/_ref/libref.so missing: run `make -C oracle ref` here first")
    R = ol.reference()
    rng = np.random.default_rng(20261203)

    # ---- reference base81 vectors -----------------------------------------------
    boards = random_boards(rng, 400)
    for e in (4, 10, 16, 24, 36, 50):
        boards += [tuple(int(v) for v in b) for b in ol.generate(77, 9, 50, e)]
    n = len(boards)
    P = np.array([b[0] for b in boards], dtype=np.uint64)
    O = np.array([b[1] for b in boards], dtype=np.uint64)
    text = []
    buf = ctypes.create_string_buffer(17)
    for p, o in boards:
        R.ref_from_board(p, o, buf)
        text.append(buf.raw[:16].decode("latin-1"))
    np.savez_compressed(os.path.join(HERE, "ref_bitboard.npz"), player=P, opponent=O, text=np.array(text))

    # ---- exact scores, small -------------------------------------------------------
    sb = []
    for e in range(0, 11):
        cnt = 40 if e <= 8 else 16
        for b in ol.generate(11, 1, cnt, e, start=1000 * e):
            sb.append((int(b[0]), int(b[1])))
    # hand-made corner cases: both sides blocked with empties left, root pass.
    sb.append((0xFFFFFFFFFFFFFF00, 0x0))          # O wiped out, 8 empties
    sb.append((0x0, 0x00FFFFFFFFFFFFFF))          # P wiped out
    sb.append((0x00000000FFFFFFFF, 0x7FFFFFFF00000000))  # one empty: H8
    P = np.array([b[0] for b in sb], dtype=np.uint64)
    O = np.array([b[1] for b in sb], dtype=np.uint64)
    sc, _, _ = ol.solve_batch(np.stack([P, O], 1), 10)
    for i, (p, o) in enumerate(sb):
        mm = ol.minimax(p, o)
        assert mm == int(sc[i]), (i, hex(p), hex(o), mm, int(sc[i]))
    np.savez_compressed(os.path.join(HERE, "solve_small.npz"), player=P, opponent=O, score=sc,
                        empties=np.array([ol.empties(p, o) for p, o in sb], dtype=np.int8))

    # ---- exact scores, 12 and 14 empties --------------------------------------------
    mid = np.concatenate([ol.generate(12, 2, 200, 12), ol.generate(14, 2, 60, 14, start=5000)])
    sc, _, _ = ol.solve_batch(mid, 14)
    imgs = []
    for b in mid:
        imgs += sym_images(int(b[0]), int(b[1]))
    imgs = np.array(imgs, dtype=np.uint64)
    sc_img, _, _ = ol.solve_batch(imgs, 14)
    assert np.array_equal(np.repeat(sc, 8), sc_img), "symmetry invariance failed"
    np.savez_compressed(os.path.join(HERE, "solve_mid.npz"), player=mid[:, 0], opponent=mid[:, 1],
                        score=sc, empties=np.array([ol.empties(int(a), int(b)) for a, b in mid], dtype=np.int8))

    # ---- README boards ------------------------------------------------------------------
    readme = ["!!%[nNNNNNnNNNNN", "!!%zN{Nrn{w{Z{{{", "!!&OrNXR{OoO{R{[", "!!&WOZOXrZ[X{[{[",
              "!!&WOZX[rZXx{z{[", "!!&WOZxXxX{[{W{{", "!!&wRzXroqxnRn{{", "!!&zO{xRxQ{X{[o[",
              "!!(RzOqQQRQQwNNN"]
    rows = []
    p64, o64 = ctypes.c_uint64(), ctypes.c_uint64()
    for t in readme:
        R.ref_to_board(t.encode("latin-1"), ctypes.byref(p64), ctypes.byref(o64))
        p, o = p64.value, o64.value
        rows.append({"text": t, "player": str(p), "opponent": str(o), "empties": ol.empties(p, o),
                     "score": ol.minimax(p, o)})
    with open(os.path.join(HERE, "readme_boards.json"), "w") as f:
        json.dump(rows, f, indent=1)
    print("wrote fixtures:", n, "bitboard boards,", len(sb), "small,", len(mid), "mid,", len(rows), "readme")


if __name__ == "__main__":
    main()
