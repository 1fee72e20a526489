"""Writes oracle scores for the five workload configurations under tests/golden/.

The boards are not stored: the GPU generator (go2_generate) reproduces the
oracle's generator (orc_generate) bit for bit, and each fixture carries the
SHA-256 of the (player, opponent) uint64 pairs it was computed on, so a test
first checks that it is solving the same boards.  Scores come from
orc_solve_ref, the oracle's restatement of the reference search
(solver.cu:54-283: fail-soft negamax, window (-64, 64), fastest-first in the
top DEPTH-9 plies, corner/LSB order below, no table), whose exactness is
pinned by tests/test_oracle.py (no-pruning minimax, symmetry invariance,
closed forms).

    python tests/golden/make_config_fixtures.py [c1 c2 c3 c4 c5] [--threads 8]

Fixtures (one .npz each):
  c1_10e.npz   10,000 positions at 10 empties (seed 101, config 1): all scores
  c2_12e.npz   1,000,000 positions at 12 empties (seed 202, config 2): all scores
  c3_mix.npz   the first 20,000 positions of the 10..18-empties mix (seed 303,
               config 3): every board with <= 16 empties and the first 250
               at each of 17 and 18 empties (indices + scores)
  c4_16e.npz   100,000 positions at 16 empties (seed 20261203, config 4): all
               scores (the benchmark's batch)
  c5_20e.npz   the first 128 of the 20-empties positions (seed 505, config 5)
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402

# name -> (seed, config id, positions, empties min, empties max)
CONFIGS = {
    "c1": (101, 1, 10_000, 10, 10),
    "c2": (202, 2, 1_000_000, 12, 12),
    "c3": (303, 3, 20_000, 10, 18),
    "c4": (20261203, 4, 100_000, 16, 16),
    "c5": (505, 5, 128, 20, 20),
}
FILES = {"c1": "c1_10e.npz", "c2": "c2_12e.npz", "c3": "c3_mix.npz", "c4": "c4_16e.npz", "c5": "c5_20e.npz"}
C3_PER_DEEP_LEVEL = 250  # boards checked at 17 and at 18 empties


def boards_sha256(boards: np.ndarray) -> str:
    """SHA-256 of the little-endian (player, opponent) uint64 pairs."""
    return hashlib.sha256(np.ascontiguousarray(boards, dtype="<u8").tobytes()).hexdigest()


def oracle_boards(seed: int, config: int, n: int, emin: int, emax: int) -> np.ndarray:
    """(n, 2) boards of orc_generate, empties per index from orc_empties_for."""
    if emin == emax:
        return ol.generate(seed, config, n, emin)
    L = ol.oracle()
    out = np.zeros((n, 2), dtype=np.uint64)
    p, o = ol.u64(), ol.u64()
    for i in range(n):
        e = L.orc_empties_for(seed, config, i, emin, emax)
        assert L.orc_generate(seed, config, i, e, ol.ctypes.byref(p), ol.ctypes.byref(o)) >= 1
        out[i] = (p.value, o.value)
    return out


def empties_of(boards: np.ndarray) -> np.ndarray:
    return (64 - np.bitwise_count(boards[:, 0] | boards[:, 1])).astype(np.int8)


PARTIAL = os.path.join(HERE, "_partial")  # resumable chunks of a long run (git-ignored)
CHUNK = 10_000


def solve_chunked(name: str, boards: np.ndarray, depth: int, threads: int):
    """orc_solve_batch over chunks of CHUNK boards, each chunk's scores kept
    in PARTIAL so that an interrupted run resumes where it stopped."""
    os.makedirs(PARTIAL, exist_ok=True)
    digest = boards_sha256(boards)[:16]
    parts, secs = [], 0.0
    for lo in range(0, len(boards), CHUNK):
        path = os.path.join(PARTIAL, f"{name}_{digest}_{lo}.npy")
        if os.path.exists(path):
            parts.append(np.load(path))
            continue
        sc, _, t = ol.solve_batch(boards[lo:lo + CHUNK], depth, threads=threads)
        secs += t
        np.save(path, sc)
        parts.append(sc)
        print(f"  {name}: {lo + len(sc)}/{len(boards)} solved", flush=True)
    return np.concatenate(parts).astype(np.int8), secs


def make(name: str, threads: int) -> None:
    seed, cid, n, emin, emax = CONFIGS[name]
    t0 = time.time()
    b = oracle_boards(seed, cid, n, emin, emax)
    e = empties_of(b)
    meta = dict(seed=seed, config_id=cid, n=n, empties_min=emin, empties_max=emax,
                boards_sha256=boards_sha256(b))
    if name == "c3":
        idx = [i for i in range(n) if e[i] <= 16]
        for deep in (17, 18):
            idx += [int(i) for i in np.nonzero(e == deep)[0][:C3_PER_DEEP_LEVEL]]
        idx = np.array(sorted(idx), dtype=np.int64)
    else:
        idx = np.arange(n, dtype=np.int64)
    sc, secs = solve_chunked(name, b[idx], emax, threads)
    path = os.path.join(HERE, FILES[name])
    arrays = dict(score=sc, empties=e[idx])
    if name == "c3":
        arrays["index"] = idx.astype(np.int32)
    np.savez_compressed(path, **arrays, **{k: np.array(v) for k, v in meta.items()})
    print(f"{name}: {len(idx)} scores ({secs:.0f} s of solve on {threads} threads, "
          f"{time.time() - t0:.0f} s total) -> {os.path.basename(path)} "
          f"({os.path.getsize(path)} bytes)", flush=True)


def main() -> None:
    args = sys.argv[1:]
    threads = os.cpu_count() or 8
    if "--threads" in args:
        i = args.index("--threads")
        threads = int(args[i + 1])
        del args[i:i + 2]
    for name in args or list(CONFIGS):
        make(name, threads)


if __name__ == "__main__":
    main()
