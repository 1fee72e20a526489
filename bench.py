"""Benchmark: exact-solved positions/sec at 16 empties on one MI355X,
learned-order batched solver vs the plain batched kernel, every score checked.

One step = one exact solve of the whole batch (default 100,000 random-playout
positions at 16 empties, seed 20261203, config id 4), boards resident in HBM
(a go2 task uploaded once), timed from launch to completion.  The timed
solver is the headline configuration: learned move ordering, a shared
lock-free transposition table of 2^31 buckets probed at nodes with >= 8
empties, 40 boards per wavefront.

Beside the timed steps, in the same run:
  * A/B on the first 8,192 positions of the same batch, three legs timed with
    HIP events on the solve stream:
      plain     static corner / X-square ordering, no table   (the plain kernel's search)
      learned   learned ordering, no table, same launch shape (differs from plain in ordering only)
      headline  the timed configuration                       (ordering, table, boards per wavefront)
    reported as ratio_learned_vs_plain and ratio_headline_vs_plain, with the
    knobs each pair differs in;
  * every score of the last timed step (all 100,000) and of every A/B leg
    against the oracle's scores committed in tests/golden/c4_16e.npz
    (orc_solve_ref, the restatement of the reference search), after checking
    the boards' SHA-256 against the fixture's;
  * the CPU baseline: the oracle's reference-method search on a bounded
    sample of the same positions on the host cores (16 threads, plus one).

Prints ONE JSON line.  `--gpus N` with N > 1 exits: this solver runs on one
GPU.  Exit status 1 if any score differs.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gpuothello2_amd as go2  # noqa: E402

# Integer work per search node of the reference formulation (one mobility
# ~131 u64 ops + one flip ~82 + ~10 bookkeeping = 223 u64 ops = 446 int32
# lane-ops): the algorithmic unit of the VALU roofline.
ALGO_INT_OPS_PER_NODE = 446
# MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz peak engine clock.
PEAK_INT32_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
FIXTURE = os.path.join(ROOT, "tests", "golden", "c4_16e.npz")
# HBM bytes per headline launch from two rocprofv3 --pmc runs (FETCH_SIZE,
# WRITE_SIZE, copy-calibrated); the file names the command and kernel build.
TRAFFIC = os.path.join(ROOT, "profiles", "solve_kernel_16e_traffic.json")


def log(msg: str) -> None:
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def boards_sha256(b: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(b).view("<u8").tobytes()).hexdigest()


def run_leg(boards, depth, cfg, label):
    """One solve of `boards` on a fresh task: (scores, stats)."""
    t = go2.Task(len(boards), depth, cfg)
    try:
        t.upload(boards)
        t.launch()
        t.wait()
        s, st = t.download()
    finally:
        t.close()
    log(f"[bench] {label}: {len(boards)} positions, kernel {st['kernel_ms']:.1f} ms, "
        f"{st['nodes'] / len(boards):.0f} nodes/position")
    return s, st


def leg_summary(n, st, knobs):
    return {"value": n / (st["kernel_ms"] / 1e3), "kernel_ms": st["kernel_ms"],
            "nodes_per_position": st["nodes"] / n, "knobs": knobs}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100000, help="positions per step")
    ap.add_argument("--empties", type=int, default=16)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--tt-log2", type=int, default=31,
                    help="headline solver: shared table of 2^N 16-byte buckets (0 = none; 31 = 32 GiB)")
    ap.add_argument("--tt-min", type=int, default=8, help="headline solver: probe/store nodes with >= N empties")
    ap.add_argument("--ppw", type=int, default=40, help="headline solver: boards a wavefront owns at once")
    ap.add_argument("--ab-n", type=int, default=8192, help="positions of the A/B legs (0 = skip)")
    ap.add_argument("--ab-limit-s", type=int, default=240, help="time limit of each A/B leg")
    ap.add_argument("--cpu-sample", type=int, default=2048,
                    help="positions for the CPU baseline leg (about 15 s on 16 threads at 16 empties)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cores)")
    ap.add_argument("--cpu-single-sample", type=int, default=128,
                    help="positions for the one-thread CPU run (0 = skip)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device (HBM counter calibration for rocprofv3)")
    args = ap.parse_args()
    if args.copy_calibration:
        from gpuothello2_amd.hip import DeviceBuffer, hip
        src, dst = DeviceBuffer(args.copy_calibration), DeviceBuffer(args.copy_calibration)
        for _ in range(3):
            assert hip().hipMemcpy(dst.ptr, src.ptr, args.copy_calibration, 3) == 0
        print(json.dumps({"copy_bytes": args.copy_calibration, "copies": 3}))
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this solver runs on one GPU; multi-GPU runs are not part of this build "
              "(positions shard with no exchange, see the design notes).", file=sys.stderr)
        return 2

    from gpuothello2_amd import learn
    if not os.path.exists(learn.WEIGHTS_PATH):
        raise SystemExit(f"no trained evaluator at {learn.WEIGHTS_PATH}")
    learned_meta = learn.install()
    E, n = args.empties, args.n
    head_knobs = {"ordering": "learned", "tt_log2_buckets": args.tt_log2, "tt_min_empties": args.tt_min,
                  "positions_per_wave": args.ppw}
    cfg = go2.default_config(ordering=go2.ORDER_LEARNED, tt_log2_buckets=args.tt_log2,
                             tt_min_empties=args.tt_min, positions_per_wave=args.ppw)

    log(f"[bench] generating {n} positions at {E} empties on the GPU")
    boards = go2.generate(args.seed, 4, n, E)
    # Oracle scores of this batch (or of its prefix), when the fixture covers it.
    # Position i depends only on (seed, config id, i), so a shorter batch is
    # a prefix of the fixture's and a longer one starts with it.
    want = None
    if os.path.exists(FIXTURE):
        fx = np.load(FIXTURE)
        fn = int(fx["n"])
        if (int(fx["seed"]), int(fx["config_id"]), int(fx["empties_min"])) == (args.seed, 4, E):
            full = boards[:fn] if n >= fn else go2.generate(args.seed, 4, fn, E)
            if boards_sha256(full) == str(fx["boards_sha256"]):
                want = fx["score"][:min(n, fn)]
    task = go2.Task(n, E, cfg)
    # Warmup: the same task on a small prefix (code load, clocks, the task's
    # table and stack allocated); the timed steps then solve the whole batch.
    # Table entries carry a launch generation, so nothing a warmup stored is
    # visible to a timed launch.
    wn = min(n, 4096)
    task.upload(boards[:wn])
    for i in range(args.warmup):
        task.launch()
        task.wait()
        log(f"[bench] warmup {i + 1}/{args.warmup} done ({wn} positions)")
    task.upload(boards)
    kernel_ms, nodes, step_scores, step_stats = [], [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        task.launch()
        task.wait()
        scores, st = task.download()
        kernel_ms.append(st["kernel_ms"])
        nodes.append(st["nodes"])
        step_scores.append(scores)
        step_stats.append(st)
        log(f"[bench] step {i + 1}/{args.steps}: kernel {st['kernel_ms']:.1f} ms, nodes {st['nodes']}")
    t1 = time.perf_counter()
    # download() (n bytes of scores + counters) sits inside the timed loop.
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n / (ms_per_step / 1e3)
    k_ms = float(np.mean(kernel_ms))
    nodes_per_launch = float(np.mean(nodes))
    achieved = nodes_per_launch * ALGO_INT_OPS_PER_NODE / (k_ms / 1e3) / 1e12
    task.close()
    steps_agree = all(np.array_equal(s, step_scores[0]) for s in step_scores)
    match = {"steps_agree": steps_agree}
    if want is not None:
        k = len(want)
        bad = int(np.sum(step_scores[-1][:k] != want))
        match["headline_vs_oracle_fixture"] = {"checked": k, "mismatches": bad}
        log(f"[bench] headline scores vs oracle fixture: {k - bad}/{k} equal")

    # ---- A/B on the first ab_n positions ---------------------------------------------
    ab = None
    an = min(args.ab_n, n)
    if an > 0:
        limit = args.ab_limit_s * 1000
        b = boards[:an]
        legs = {
            "plain": (go2.default_config(ordering=go2.ORDER_STATIC, time_limit_ms=limit),
                      {"ordering": "static", "tt_log2_buckets": 0, "positions_per_wave": "auto"}),
            "learned": (go2.default_config(ordering=go2.ORDER_LEARNED, time_limit_ms=limit),
                        {"ordering": "learned", "tt_log2_buckets": 0, "positions_per_wave": "auto"}),
            "headline": (go2.Config.from_buffer_copy(cfg), head_knobs),
        }
        ab = {"positions": an, "sample": f"first {an} of the same positions",
              "same_launch_shape": "persistent resident grid, helper lanes and cross-wavefront help in every leg"}
        got = {}
        for name, (c, knobs) in legs.items():
            c.time_limit_ms = limit
            try:
                leg_scores, leg_st = run_leg(b, E, c, f"A/B {name}")
                ab[name] = leg_summary(an, leg_st, knobs)
                got[name] = leg_scores
            except go2.Go2Error as e:
                log(f"[bench] A/B {name} did not finish: {e}")
                ab[name] = {"value_upper_bound": an / args.ab_limit_s, "knobs": knobs, "error": str(e)}
        if "value" in ab.get("plain", {}):
            for name in ("learned", "headline"):
                if "value" in ab.get(name, {}):
                    ab[f"ratio_{name}_vs_plain"] = ab[name]["value"] / ab["plain"]["value"]
        # "ordering" = learned: evaluator move keys near the root and an
        # aspiration window around the evaluator's estimate of each board
        ab["knobs_differing"] = {"learned_vs_plain": ["ordering"],
                                 "headline_vs_plain": ["ordering", "tt_log2_buckets", "tt_min_empties",
                                                       "positions_per_wave"]}
        ref = want[:an] if want is not None and len(want) >= an else step_scores[-1][:an]
        match["ab_legs"] = {name: int(np.sum(s != ref)) for name, s in got.items()}
        match["ab_reference"] = "oracle fixture" if want is not None and len(want) >= an else "headline step"

    # ---- CPU baseline: the reference method on the host cores ----------------------------
    cpu = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as ol

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        k = min(args.cpu_sample, n)
        b2 = np.stack([boards["player"][:k], boards["opponent"][:k]], 1)
        cs, cnodes, secs = ol.solve_batch(b2, E, threads=threads)
        cpu_bad = int(np.sum(cs != step_scores[-1][:k]))
        k1 = min(args.cpu_single_sample, k)
        single = None
        if k1 > 0:
            c1, _, secs1 = ol.solve_batch(b2[:k1], E, threads=1)
            cpu_bad += int(np.sum(c1 != step_scores[-1][:k1]))
            single = {"value": k1 / secs1, "positions": k1, "seconds": secs1}
        match["cpu_sample_vs_gpu"] = {"checked": k, "mismatches": cpu_bad}
        model = None
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
        except OSError:
            pass
        cpu = {"value": k / secs, "unit": "positions/s", "cores": threads, "kind": "port",
               "cpu_model": model, "host_cores": os.cpu_count(), "single_thread": single,
               "sample": f"first {k} of the same {E}-empty positions, reference method "
                         f"(fastest-first in the top DEPTH-9 plies, corner/LSB order below, no table), "
                         f"{threads} threads, {secs:.1f} s"}
        log(f"[bench] cpu baseline {k / secs:.1f} pos/s on {threads} threads; mismatches {cpu_bad}")

    traffic, traffic_src = None, None
    if os.path.exists(TRAFFIC) and n == 100000 and E == 16:
        with open(TRAFFIC) as f:
            tj = json.load(f)
        if tj.get("config") == head_knobs:
            traffic = tj["bytes_per_launch"]
            traffic_src = tj["source"]
    mismatches = [match[k_]["mismatches"] for k_ in ("headline_vs_oracle_fixture", "cpu_sample_vs_gpu")
                  if k_ in match] + list(match.get("ab_legs", {}).values())
    ok = steps_agree and sum(mismatches) == 0
    match["all_equal"] = bool(ok)
    line = {
        "metric": "exact-solved positions/sec @16 empties, 1 GPU (learned-order vs plain), 100% match",
        "value": value,
        "unit": "positions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"exact solve of {n} random-playout positions at {E} empties",
            "positions": n, "empties": E, **head_knobs,
            "nodes_per_position": nodes_per_launch / n,
            "table": {k_: step_stats[-1][k_] for k_ in ("tt_probes", "tt_hits", "tt_cuts", "tt_stores")},
            "ab": ab,
            "score_match": match,
            "evaluator": {k_: learned_meta[k_] for k_ in ("positions", "empties", "split", "valid_mse",
                                                          "valid_mse_constant_predictor")},
        },
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_INT32_TOPS,
                     "unit": "Tops/s (int32 lane-ops)", "frac": achieved / PEAK_INT32_TOPS,
                     "traffic": traffic, "traffic_source": traffic_src, "kernel_ms": k_ms,
                     "ops_per_node": ALGO_INT_OPS_PER_NODE},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
