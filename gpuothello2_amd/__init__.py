"""gpuothello2_amd — batched exact Othello endgame solver for MI355X (gfx950).

Python view of the C ABI in ``include/go2.h`` (the HIP library
``gpuothello2_amd/libgo2.so``), for tests, benchmarks and scripting.  Boards
are ``(player, opponent)`` uint64 pairs, square = row*8 + col (bit 0 = A1),
``player`` to move; scores are exact disc differences in [-64, 64].

The product path is the HIP library only: if ``libgo2.so`` is missing this
module raises on first use instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgo2.so")
SOLVER_BIN = os.path.join(_HERE, "bin", "solver")

GO2_OK = 0
GO2_E_ARG = -1
GO2_E_BOARD = -2
GO2_E_DEVICE = -3
GO2_E_NOMEM = -4
GO2_E_STATE = -5
GO2_E_PARSE = -6
GO2_E_TIMEOUT = -7
GO2_MAX_EMPTIES = 32

MOVE_PASS = 64  # best-move codes besides squares 0..63 (bit 0 = A1)
MOVE_NONE = 65

ORDER_STATIC = 0
ORDER_MOBILITY = 1
ORDER_LEARNED = 2

# Every symbol include/go2.h declares.
EXPORTED_SYMBOLS = (
    "go2_version", "go2_last_error", "go2_set_device", "go2_config_default",
    "go2_board_from_text", "go2_board_to_text", "go2_solve", "go2_solve_device",
    "go2_solve_moves", "go2_solve_device_moves", "go2_task_download_moves", "go2_think",
    "go2_task_create", "go2_task_upload", "go2_task_launch", "go2_task_query",
    "go2_task_wait", "go2_task_download", "go2_task_destroy", "go2_generate",
    "go2_generate_device", "go2_pattern_info", "go2_pattern_features_device",
    "go2_pattern_eval_device", "go2_sgd_gradient_device", "go2_sgd_apply_device",
    "go2_quantize_weights_device", "go2_set_evaluator", "go2_solve_device_noalloc",
    "go2_task_upload_async", "go2_host_alloc", "go2_host_free",
)
NOALLOC_SCRATCH_BYTES = 128 << 20

BOARD_DTYPE = np.dtype([("player", "<u8"), ("opponent", "<u8")])


class Go2Error(RuntimeError):
    """A libgo2 call returned a negative status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libgo2 error {code}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [
        ("ordering", ctypes.c_int32),
        ("mobility_min_empties", ctypes.c_int32),
        ("learned_min_empties", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("positions_per_wave", ctypes.c_int32),
        ("split_min_empties", ctypes.c_int32),
        ("lds_levels", ctypes.c_int32),
        ("time_limit_ms", ctypes.c_int32),
        ("split_cut_after", ctypes.c_int32),
        ("learned_key_eval", ctypes.c_int32),
        ("learned_key_mobility", ctypes.c_int32),
        ("max_pops_per_trip", ctypes.c_int32),
        ("tt_log2_buckets", ctypes.c_int32),
        ("tt_min_empties", ctypes.c_int32),
        ("graph_launch", ctypes.c_int32),
        ("export_min_empties", ctypes.c_int32),
        ("export_low_empties", ctypes.c_int32),
        ("think_depth", ctypes.c_int32),
        ("leaf_batch_min", ctypes.c_int32),
        ("score_batch_min", ctypes.c_int32),
        ("aspiration_width", ctypes.c_int32),
        ("pvs_min_empties", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 1),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_uint64),
        ("nodes", ctypes.c_uint64),
        ("score_steps", ctypes.c_uint64),
        ("iterations", ctypes.c_uint64),
        ("kernel_ms", ctypes.c_double),
        ("steals", ctypes.c_uint64),
        ("lane_slots", ctypes.c_uint64),
        ("grid_waves", ctypes.c_uint64),
        ("tt_probes", ctypes.c_uint64),
        ("tt_hits", ctypes.c_uint64),
        ("tt_cuts", ctypes.c_uint64),
        ("tt_stores", ctypes.c_uint64),
        ("exports", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class Board(ctypes.Structure):
    _fields_ = [("player", ctypes.c_uint64), ("opponent", ctypes.c_uint64)]


_lib = None


def lib() -> ctypes.CDLL:
    """Load libgo2.so (built in-tree by ``make -C gpuothello2_amd``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; build it with `make -C gpuothello2_amd` first")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, u64, i32, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32
    L.go2_version.restype = i32
    L.go2_last_error.restype = ctypes.c_char_p
    L.go2_set_device.argtypes = [i32]
    L.go2_config_default.argtypes = [ctypes.POINTER(Config)]
    L.go2_config_default.restype = None
    L.go2_board_from_text.argtypes = [ctypes.c_char_p, ctypes.POINTER(Board)]
    L.go2_board_to_text.argtypes = [ctypes.POINTER(Board), ctypes.c_char_p]
    L.go2_board_to_text.restype = None
    L.go2_solve.argtypes = [vp, sz, i32, vp, ctypes.POINTER(Config), ctypes.POINTER(Stats)]
    L.go2_solve_device.argtypes = [vp, sz, i32, vp, ctypes.POINTER(Config), ctypes.POINTER(Stats), vp]
    L.go2_solve_moves.argtypes = [vp, sz, i32, vp, vp, ctypes.POINTER(Config), ctypes.POINTER(Stats)]
    L.go2_solve_device_moves.argtypes = [vp, sz, i32, vp, vp, ctypes.POINTER(Config), ctypes.POINTER(Stats), vp]
    L.go2_task_download_moves.argtypes = [vp, vp]
    L.go2_solve_device_noalloc.argtypes = [vp, sz, i32, vp, vp, ctypes.POINTER(Config), ctypes.POINTER(Stats), vp]
    L.go2_think.argtypes = [vp, sz, i32, vp, vp, ctypes.POINTER(Config), ctypes.POINTER(Stats)]
    L.go2_task_create.argtypes = [sz, i32, ctypes.POINTER(Config)]
    L.go2_task_create.restype = vp
    L.go2_task_upload.argtypes = [vp, vp, sz]
    L.go2_task_launch.argtypes = [vp]
    L.go2_task_upload_async.argtypes = [vp, vp, sz]
    L.go2_host_alloc.argtypes = [sz, ctypes.POINTER(ctypes.c_void_p)]
    L.go2_host_free.argtypes = [vp]
    L.go2_host_free.restype = None
    L.go2_task_query.argtypes = [vp]
    L.go2_task_wait.argtypes = [vp]
    L.go2_task_download.argtypes = [vp, vp, ctypes.POINTER(Stats)]
    L.go2_task_destroy.argtypes = [vp]
    L.go2_task_destroy.restype = None
    L.go2_generate.argtypes = [u64, u32, sz, i32, i32, vp]
    L.go2_generate_device.argtypes = [u64, u32, sz, i32, i32, vp, vp]
    ip = ctypes.POINTER(ctypes.c_int)
    L.go2_pattern_info.argtypes = [ip, ip, ip]
    L.go2_pattern_features_device.argtypes = [vp, sz, vp, vp, vp]
    L.go2_pattern_eval_device.argtypes = [vp, vp, vp, sz, vp, vp]
    L.go2_sgd_gradient_device.argtypes = [vp, vp, vp, vp, vp, sz, vp, vp, vp, vp]
    L.go2_sgd_apply_device.argtypes = [vp, vp, vp, vp, ctypes.c_float, vp]
    L.go2_quantize_weights_device.argtypes = [vp, vp, ctypes.c_float, vp, vp, vp]
    L.go2_set_evaluator.argtypes = [vp, vp, ctypes.c_float]
    for name in EXPORTED_SYMBOLS:
        fn = getattr(L, name)
        if name not in ("go2_last_error", "go2_config_default", "go2_board_to_text",
                        "go2_task_create", "go2_task_destroy", "go2_host_free"):
            fn.restype = i32
    _lib = L
    return L


def last_error() -> str:
    return lib().go2_last_error().decode()


def _check(rc: int) -> int:
    if rc < 0:
        raise Go2Error(rc, last_error())
    return rc


def default_config(**overrides) -> Config:
    c = Config()
    lib().go2_config_default(ctypes.byref(c))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def as_boards(player, opponent=None) -> np.ndarray:
    """Pack player/opponent uint64 arrays (or an (n,2) array) into go2_board records."""
    if opponent is None:
        a = np.asarray(player, dtype=np.uint64).reshape(-1, 2)
        player, opponent = a[:, 0], a[:, 1]
    p = np.asarray(player, dtype=np.uint64).ravel()
    o = np.asarray(opponent, dtype=np.uint64).ravel()
    out = np.empty(p.shape[0], dtype=BOARD_DTYPE)
    out["player"] = p
    out["opponent"] = o
    return out


def solve(boards: np.ndarray, max_depth: int, config: Config | None = None):
    """Exact scores of host boards; returns (int8 scores, stats dict)."""
    b = np.ascontiguousarray(boards, dtype=BOARD_DTYPE)
    n = b.shape[0]
    scores = np.zeros(max(n, 1), dtype=np.int8)
    st = Stats()
    _check(lib().go2_solve(b.ctypes.data if n else None, n, max_depth, scores.ctypes.data,
                           ctypes.byref(config) if config is not None else None, ctypes.byref(st)))
    return scores[:n], st.as_dict()


def solve_moves(boards: np.ndarray, max_depth: int, config: Config | None = None):
    """Exact scores and best moves of host boards; returns (int8 scores,
    uint8 moves, stats dict).  A move is a square 0..63 whose child has the
    board's score, MOVE_PASS or MOVE_NONE."""
    b = np.ascontiguousarray(boards, dtype=BOARD_DTYPE)
    n = b.shape[0]
    scores = np.zeros(max(n, 1), dtype=np.int8)
    moves = np.zeros(max(n, 1), dtype=np.uint8)
    st = Stats()
    _check(lib().go2_solve_moves(b.ctypes.data if n else None, n, max_depth, scores.ctypes.data,
                                 moves.ctypes.data, ctypes.byref(config) if config is not None else None,
                                 ctypes.byref(st)))
    return scores[:n], moves[:n], st.as_dict()


def think(boards: np.ndarray, depth: int, config: Config | None = None):
    """The reference's think mode: depth-limited search, the position after
    `depth` moves scored by the installed evaluator (set_evaluator); boards
    may have up to 60 empties.  Returns (int8 scores, uint8 moves, stats)."""
    b = np.ascontiguousarray(boards, dtype=BOARD_DTYPE)
    n = b.shape[0]
    scores = np.zeros(max(n, 1), dtype=np.int8)
    moves = np.zeros(max(n, 1), dtype=np.uint8)
    st = Stats()
    _check(lib().go2_think(b.ctypes.data if n else None, n, depth, scores.ctypes.data, moves.ctypes.data,
                           ctypes.byref(config) if config is not None else None, ctypes.byref(st)))
    return scores[:n], moves[:n], st.as_dict()


def move_text(m: int) -> str:
    """The reference's move notation (main.cu:31-39): column letter, row digit; "ps" = pass."""
    if m == MOVE_PASS:
        return "ps"
    if m == MOVE_NONE:
        return "--"
    return chr(ord("A") + m % 8) + chr(ord("1") + m // 8)


def solve_device(d_boards_ptr: int, n: int, max_depth: int, d_scores_ptr: int,
                 config: Config | None = None, stream: int | None = None) -> dict:
    """Solve boards already in device memory (raw pointers); returns stats."""
    st = Stats()
    _check(lib().go2_solve_device(d_boards_ptr, n, max_depth, d_scores_ptr,
                                  ctypes.byref(config) if config is not None else None,
                                  ctypes.byref(st), stream))
    return st.as_dict()


def solve_device_noalloc(d_boards_ptr: int, n: int, max_depth: int, d_scores_ptr: int,
                         d_moves_ptr: int | None = None, config: Config | None = None,
                         stream: int | None = None) -> dict:
    """go2_solve_device_noalloc: boards already in device memory, no device
    allocation by the library (its scratch is a static device array shared
    by every call on the device); returns stats."""
    st = Stats()
    _check(lib().go2_solve_device_noalloc(d_boards_ptr, n, max_depth, d_scores_ptr, d_moves_ptr,
                                          ctypes.byref(config) if config is not None else None,
                                          ctypes.byref(st), stream))
    return st.as_dict()


def generate(seed: int, config_id: int, n: int, empties_min: int, empties_max: int | None = None) -> np.ndarray:
    """Random-playout positions generated on the GPU (see go2_generate)."""
    if empties_max is None:
        empties_max = empties_min
    out = np.zeros(max(n, 1), dtype=BOARD_DTYPE)
    _check(lib().go2_generate(seed, config_id, n, empties_min, empties_max, out.ctypes.data))
    return out[:n]


def board_from_text(text: str) -> tuple[int, int]:
    b = Board()
    _check(lib().go2_board_from_text(text.encode("latin-1"), ctypes.byref(b)))
    return b.player, b.opponent


def board_to_text(player: int, opponent: int) -> str:
    b = Board(player, opponent)
    buf = ctypes.create_string_buffer(17)
    lib().go2_board_to_text(ctypes.byref(b), buf)
    return buf.value.decode("latin-1")


@dataclass
class Task:
    """Mirror of the reference's BatchedTask (solver.cuh:8-25): create, upload,
    launch asynchronously, poll, download."""

    n: int
    max_depth: int
    config: Config | None = None

    def __post_init__(self):
        h = lib().go2_task_create(self.n, self.max_depth,
                                  ctypes.byref(self.config) if self.config is not None else None)
        if not h:
            raise Go2Error(GO2_E_ARG, last_error())
        self._h = h
        self.count = 0

    def upload(self, boards: np.ndarray) -> None:
        b = np.ascontiguousarray(boards, dtype=BOARD_DTYPE)
        _check(lib().go2_task_upload(self._h, b.ctypes.data if b.shape[0] else None, b.shape[0]))
        self.count = b.shape[0]

    def launch(self) -> None:
        _check(lib().go2_task_launch(self._h))

    def is_ready(self) -> bool:
        return _check(lib().go2_task_query(self._h)) == 1

    def wait(self) -> None:
        _check(lib().go2_task_wait(self._h))

    def download(self):
        scores = np.zeros(max(self.count, 1), dtype=np.int8)
        st = Stats()
        _check(lib().go2_task_download(self._h, scores.ctypes.data, ctypes.byref(st)))
        return scores[: self.count], st.as_dict()

    def download_moves(self) -> np.ndarray:
        moves = np.zeros(max(self.count, 1), dtype=np.uint8)
        _check(lib().go2_task_download_moves(self._h, moves.ctypes.data))
        return moves[: self.count]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().go2_task_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pattern_info() -> tuple[int, int, int]:
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().go2_pattern_info(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def set_evaluator(w16: np.ndarray, bias16: np.ndarray, scale: float) -> None:
    """Install the int16 evaluator used by ORDER_LEARNED solves."""
    _, nw, nb = pattern_info()
    w = np.ascontiguousarray(w16, dtype=np.int16)
    b = np.ascontiguousarray(bias16, dtype=np.int16)
    assert w.shape == (nw,) and b.shape == (nb,)
    _check(lib().go2_set_evaluator(w.ctypes.data, b.ctypes.data, float(scale)))
