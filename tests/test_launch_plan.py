"""Launch plan of a solve (gpuothello2_amd/csrc/plan.hpp), checked on the CPU
through a g++ build of the same header (tests/plan_check.cpp).

The plan decides the stack levels, the grid and the cross-wavefront export
levels; the kernel shifts a 32-bit mask by (export level - 3) and indexes
frames and mailboxes by level, so every plan must keep those in range."""
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def plan_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("plan") / "plan_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "gpuothello2_amd", "csrc"),
                    os.path.join(HERE, "plan_check.cpp"), "-o", out], check=True)
    return out


def plan(binary, n, depth, cus=256, per_cu=2, **fields):
    args = [binary, str(n), str(depth), str(cus), str(per_cu)] + [f"{k}={v}" for k, v in fields.items()]
    r = subprocess.run(args, capture_output=True, text=True, check=True)
    return {k: int(v) for k, v in (ln.split() for ln in r.stdout.splitlines())}


def test_headline_shape(plan_bin):
    # two 256-lane workgroups per CU (the kernel's launch bound), each with
    # nine stack levels (4..12) in its half of the CU's 160 KB of LDS
    p = plan(plan_bin, 100000, 16)
    assert (p["depth"], p["maxe"], p["levels"], p["lds_levels"]) == (16, 16, 13, 9)
    assert p["lds_bytes"] == 9 * 256 * 32 + 2048 + 4 * 64
    assert (p["grid"], p["waves"], p["total_lanes"]) == (512, 2048, 512 * 256)
    assert p["K"] == -(-100000 // 2048)
    assert (p["help"], p["ext_min"], p["ext_lo"], p["split_min"]) == (1, 12, 9, 8)
    p = plan(plan_bin, 100000, 16, ordering=2, positions_per_wave=24)
    assert p["lds_bytes"] == 9 * 256 * 32 + 2048 + 2048 + 4 * 64 and p["K"] == 24
    assert 2 * p["lds_bytes"] <= 160 * 1024  # two workgroups per CU
    assert 2 * (p["lds_bytes"] + 256 * 32) > 160 * 1024  # a tenth level would not fit
    # smaller workgroups: more of them per CU, the same LDS per lane
    for bs, levels in ((128, 8), (64, 7)):
        q = plan(plan_bin, 100000, 16, block_size=bs, ordering=2)
        assert q["lds_levels"] == levels
        assert (512 // bs) * q["lds_bytes"] <= 160 * 1024


def test_depth_beyond_kernel_limit_sizes_for_32(plan_bin):
    # DEPTH may be anything a game allows; the stack covers at most 32 empties.
    for depth in (33, 40, 60):
        p = plan(plan_bin, 1000, depth)
        assert (p["depth"], p["maxe"], p["levels"]) == (32, 32, 29)
    p = plan(plan_bin, 1000, 5)
    assert (p["depth"], p["maxe"], p["levels"]) == (5, 5, 2)
    p = plan(plan_bin, 1000, 1)
    assert (p["depth"], p["maxe"], p["levels"]) == (1, 3, 1)


def test_export_level_above_stack_is_capped(plan_bin):
    # An export level beyond the deepest level is capped to maxe + 1 (nothing
    # to export there), so (level - 3) stays below 32 and mailbox indices
    # (level - ext_lo, level <= maxe) stay in range.
    p = plan(plan_bin, 5000, 14, export_min_empties=40)
    assert p["help"] == 1 and p["ext_min"] == 15 and p["ext_lo"] == 9
    p = plan(plan_bin, 5000, 14, export_min_empties=61)
    assert p["help"] == 0  # > 60 means never


def test_plan_invariants_random_sweep(plan_bin):
    rng = random.Random(7)
    for _ in range(300):
        depth = rng.randint(1, 60)
        f = dict(export_min_empties=rng.choice([0, rng.randint(1, 70)]),
                 export_low_empties=rng.choice([0, rng.randint(1, 70)]),
                 split_min_empties=rng.choice([0, rng.randint(1, 70)]),
                 lds_levels=rng.choice([0, rng.randint(1, 30)]),
                 positions_per_wave=rng.randint(0, 64))
        n = rng.choice([1, 63, 1000, 100000, 2000000])
        p = plan(plan_bin, n, depth, **f)
        assert 3 <= p["maxe"] <= 32 and p["levels"] == max(1, p["maxe"] - 3)
        assert 1 <= p["lds_levels"] <= p["levels"]
        assert p["g_levels"] == p["levels"] - p["lds_levels"]
        assert 1 <= p["K"] <= 64 and p["grid"] >= 1
        if p["help"]:
            assert p["split_min"] <= p["ext_lo"] <= p["ext_min"] <= p["maxe"] + 1, (depth, f, p)
            assert p["ext_lo"] <= p["maxe"] and p["ext_min"] - 3 < 32
        else:
            assert p["grid"] <= max(1, -(-n // 4))  # no more wavefronts than boards


def test_think_mode_plan(plan_bin):
    p = plan(plan_bin, 1000, 60, think_depth=6)
    assert (p["think"], p["depth"], p["maxe"], p["split_min"], p["help"]) == (1, 60, 8, 64, 0)


NOALLOC = 128 << 20  # GO2_NOALLOC_SCRATCH_BYTES


def test_noalloc_scratch_fit(plan_bin):
    # The no-allocation entry point's scratch holds the full grid up to 24
    # empties; deeper stacks shrink the grid, and every DEPTH fits somehow.
    p = plan(plan_bin, 100000, 16, capacity=NOALLOC)
    assert p["fits"] == 1 and p["grid"] == 512 and p["help"] == 1
    assert plan(plan_bin, 100000, 24, capacity=NOALLOC)["grid"] == 512
    assert p["scratch_bytes"] <= NOALLOC
    full = plan(plan_bin, 100000, 16)
    assert full["scratch_bytes"] == p["scratch_bytes"]
    for depth in range(1, 61):
        q = plan(plan_bin, 100000, depth, capacity=NOALLOC)
        assert q["fits"] == 1 and q["scratch_bytes"] <= NOALLOC, depth
        assert q["grid"] >= 256, depth  # at least one workgroup per CU
    q = plan(plan_bin, 100000, 32, capacity=NOALLOC)
    assert q["grid"] < 512  # shrunk
    q = plan(plan_bin, 100000, 32, capacity=1 << 20)
    assert q["fits"] == 0


def test_scratch_bytes_layout(plan_bin):
    p = plan(plan_bin, 100000, 16)
    cells = p["total_lanes"] * p["g_levels"]
    r = lambda b: (b + 255) // 256 * 256  # noqa: E731
    assert p["frame_bytes"] == 3 * r(cells * 8) + 2 * r(cells * 4)
    mb = r((p["maxe"] + 1 - p["ext_lo"]) * p["total_lanes"] * 4)
    assert p["help_bytes"] == 4096 * 64 + 7 * mb
    assert p["scratch_bytes"] == 256 + p["help_bytes"] + p["frame_bytes"]
