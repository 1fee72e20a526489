"""Autotuner: picks the launch configuration of a batch by timing short
trial solves on a sample of it.

Searched knobs (coordinate descent from the library defaults):
  positions_per_wave  boards a wavefront owns: 64 (one lane per board),
                      16 (four lanes per board), 4 (sixteen), 0 (from n)
  block_size          lanes per workgroup
  lds_levels          search-stack levels kept in LDS (the rest in HBM; 0 = as many as fit)
  split_min_empties   smallest node helper lanes may split
  tt_log2_buckets     transposition table size (deep batches only)
  tt_min_empties      smallest node the table serves (deep batches only)
The result is a plain JSON object of go2_config fields; `apply()` turns it
into a Config, and the CLI reads the same file through GO2_CONFIG.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

from . import Config, Go2Error, Task, default_config

SPACE = {
    "positions_per_wave": [0, 64, 16, 4],
    "block_size": [256, 128],
    "lds_levels": [0, 6, 8],
    "split_min_empties": [8, 7, 10],
}
DEEP_SPACE = {"tt_log2_buckets": [0, 26, 29, 31], "tt_min_empties": [12, 10, 8]}


def apply(base: Config | None, params: dict) -> Config:
    c = base if base is not None else default_config()
    out = Config.from_buffer_copy(c)
    for k, v in params.items():
        setattr(out, k, int(v))
    return out


def trial(boards: np.ndarray, max_depth: int, cfg: Config, limit_ms: int = 60000) -> float:
    """Kernel milliseconds of one solve of `boards` with `cfg` (inf on timeout)."""
    c = Config.from_buffer_copy(cfg)
    c.time_limit_ms = limit_ms
    t = Task(len(boards), max_depth, c)
    try:
        t.upload(boards)
        t.launch()
        t.wait()
        _, st = t.download()
        return float(st["kernel_ms"])
    except Go2Error:
        return float("inf")
    finally:
        t.close()


def autotune(boards: np.ndarray, max_depth: int, base: Config | None = None, sample: int = 20000,
             seed: int = 0, log=None) -> dict:
    """Returns the best parameter dict found on a random sample of `boards`."""
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(boards))[: min(sample, len(boards))]
    smp = boards[np.sort(idx)]
    base = base if base is not None else default_config()
    space = dict(SPACE)
    empties = 64 - np.array([bin(int(p) | int(o)).count("1") for p, o in zip(smp["player"], smp["opponent"])])
    if empties.max() >= 18:
        space.update(DEEP_SPACE)
    best = {}
    best_ms = trial(smp, max_depth, apply(base, best))
    history = [{"params": dict(best), "ms": best_ms}]
    for knob, values in space.items():
        for v in values:
            cand = dict(best)
            cand[knob] = v
            if cand == best:
                continue
            ms = trial(smp, max_depth, apply(base, cand))
            history.append({"params": dict(cand), "ms": ms})
            if log:
                log(f"[autotune] {cand} -> {ms:.1f} ms")
            if ms < best_ms:
                best, best_ms = cand, ms
    return {"params": best, "sample": len(smp), "kernel_ms": best_ms, "history": history}


def save(path: str, result: dict) -> None:
    with open(path, "w") as f:
        json.dump(result["params"], f, indent=1, sort_keys=True)


def load(path: str) -> dict:
    with open(path) as f:
        return {k: int(v) for k, v in json.load(f).items()}


if __name__ == "__main__":
    import argparse

    from . import ORDER_LEARNED, generate
    from . import learn

    ap = argparse.ArgumentParser(description="Tune the launch configuration for a synthetic batch")
    ap.add_argument("--n", type=int, default=200000)
    ap.add_argument("--emin", type=int, default=10)
    ap.add_argument("--emax", type=int, default=18)
    ap.add_argument("--sample", type=int, default=20000)
    ap.add_argument("--out", default="go2_tuned.json")
    a = ap.parse_args()
    base = default_config()
    if os.path.exists(learn.WEIGHTS_PATH):
        learn.install()
        base.ordering = ORDER_LEARNED
    b = generate(3, 3, a.n, a.emin, a.emax)
    r = autotune(b, a.emax, base, a.sample, log=lambda m: print(m, file=sys.stderr, flush=True))
    save(a.out, r)
    print(json.dumps({"params": r["params"], "kernel_ms": r["kernel_ms"], "sample": r["sample"]}))
