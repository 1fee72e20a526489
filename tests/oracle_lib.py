"""ctypes view of the CPU oracle (oracle/liboracle.so) and, when built, of the
reference's own host functions (oracle/_ref/libref.so).

Test infrastructure only: the oracle is the checker for the HIP solver.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libref.so")

u64, i32, u32, sz = ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32, ctypes.c_size_t
P64 = ctypes.POINTER(ctypes.c_uint64)

_orc = None
_ref = None


def oracle() -> ctypes.CDLL:
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_PATH):
            raise ImportError(f"{ORACLE_PATH} missing; run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_PATH)
        L.orc_final_score.argtypes = [u64, u64]
        for f in ("orc_flip", "orc_flip_walk"):
            getattr(L, f).argtypes = [u64, u64, i32]
            getattr(L, f).restype = u64
        for f in ("orc_mobility", "orc_mobility_walk"):
            getattr(L, f).argtypes = [u64, u64]
            getattr(L, f).restype = u64
        L.orc_next_bit.argtypes = [u64]
        L.orc_next_bit.restype = u64
        L.orc_encode.argtypes = [u64, u64, ctypes.c_char_p]
        L.orc_encode.restype = None
        L.orc_decode.argtypes = [ctypes.c_char_p, P64, P64]
        L.orc_solve_ref.argtypes = [u64, u64, i32, P64]
        L.orc_minimax.argtypes = [u64, u64, P64]
        L.orc_solve_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, sz, i32, i32, ctypes.c_void_p, P64]
        L.orc_best_moves.argtypes = [u64, u64, i32, P64, ctypes.POINTER(ctypes.c_int)]
        L.orc_best_moves_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, sz, i32, i32, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p]
        L.orc_best_moves_batch.restype = None
        L.orc_think_leaf.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, u64, u64]
        L.orc_think.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, u64, u64, i32, P64,
                                ctypes.POINTER(ctypes.c_int)]
        L.orc_think_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                      ctypes.c_void_p, sz, i32, i32, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]
        L.orc_think_batch.restype = None
        L.orc_think_value.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, u64, u64, i32]
        L.orc_solve_batch.restype = ctypes.c_double
        L.orc_position_key.argtypes = [u64, u32, u64, u32]
        L.orc_position_key.restype = u64
        L.orc_empties_for.argtypes = [u64, u32, u64, i32, i32]
        L.orc_generate.argtypes = [u64, u32, u64, i32, P64, P64]
        L.orc_pattern_layout.argtypes = [ctypes.c_void_p] * 4
        L.orc_pattern_indices.argtypes = [u64, u64, ctypes.c_void_p]
        L.orc_pattern_indices.restype = None
        L.orc_ref_pattern_index.argtypes = [u64, u64, u64]
        L.orc_ref_pattern_index.restype = u32
        for f in ("orc_pext",):
            getattr(L, f).argtypes = [u64, u64]
            getattr(L, f).restype = u64
        for f in ("orc_flip_vertical", "orc_flip_diag", "orc_rot90"):
            getattr(L, f).argtypes = [u64]
            getattr(L, f).restype = u64
        L.orc_eval_int.argtypes = [ctypes.c_void_p, ctypes.c_void_p, u64, u64]
        L.orc_eval_float.argtypes = [ctypes.c_void_p, ctypes.c_void_p, u64, u64]
        L.orc_eval_float.restype = ctypes.c_float
        L.orc_sgd_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, sz, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_sgd_step.restype = ctypes.c_double
        _orc = L
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def reference() -> ctypes.CDLL:
    """The reference's own host codec and PEXT (built only where its sources are)."""
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_PATH)
        L.ref_pext.argtypes = [u64, u64]
        L.ref_pext.restype = u64
        L.ref_to_board.argtypes = [ctypes.c_char_p, P64, P64]
        L.ref_to_board.restype = None
        L.ref_from_board.argtypes = [u64, u64, ctypes.c_char_p]
        L.ref_from_board.restype = None
        _ref = L
    return _ref


# ---- convenience wrappers -------------------------------------------------------

def generate(seed: int, config: int, n: int, empties: int, start: int = 0) -> np.ndarray:
    """Oracle playout positions: (n, 2) uint64 array of (player, opponent)."""
    L = oracle()
    out = np.zeros((n, 2), dtype=np.uint64)
    p, o = u64(), u64()
    for i in range(n):
        rc = L.orc_generate(seed, config, start + i, empties, ctypes.byref(p), ctypes.byref(o))
        assert rc >= 1
        out[i] = (p.value, o.value)
    return out


def empties_for(seed: int, config: int, index: int, emin: int, emax: int) -> int:
    return oracle().orc_empties_for(seed, config, index, emin, emax)


def solve_batch(boards: np.ndarray, depth: int, threads: int = 8):
    """Reference-method scores of (n,2) boards, multithreaded; returns (scores, nodes, seconds)."""
    b = np.ascontiguousarray(boards, dtype=np.uint64).reshape(-1, 2)
    P = np.ascontiguousarray(b[:, 0])
    O = np.ascontiguousarray(b[:, 1])
    scores = np.zeros(b.shape[0], dtype=np.int8)
    nodes = u64(0)
    secs = oracle().orc_solve_batch(P.ctypes.data, O.ctypes.data, b.shape[0], depth, threads,
                                    scores.ctypes.data, ctypes.byref(nodes))
    return scores, nodes.value, secs


MOVES_PLAY, MOVES_PASS, MOVES_GAME_OVER = 0, 1, 2


def best_moves(p: int, o: int, depth: int = 60):
    """(score, mask of optimal squares, kind) of one board."""
    m = u64(0)
    k = ctypes.c_int(0)
    s = oracle().orc_best_moves(p, o, depth, ctypes.byref(m), ctypes.byref(k))
    return s, m.value, k.value


def best_moves_batch(boards: np.ndarray, depth: int, threads: int = 8):
    """(scores, optimal-square masks, kinds) of (n,2) boards, multithreaded."""
    b = np.ascontiguousarray(boards, dtype=np.uint64).reshape(-1, 2)
    P = np.ascontiguousarray(b[:, 0])
    O = np.ascontiguousarray(b[:, 1])
    n = b.shape[0]
    scores = np.zeros(n, dtype=np.int8)
    masks = np.zeros(n, dtype=np.uint64)
    kinds = np.zeros(n, dtype=np.uint8)
    oracle().orc_best_moves_batch(P.ctypes.data, O.ctypes.data, n, depth, threads, scores.ctypes.data,
                                  masks.ctypes.data, kinds.ctypes.data)
    return scores, masks, kinds


def think(w16: np.ndarray, b16: np.ndarray, scale: float, p: int, o: int, depth: int):
    """(value, mask of optimal squares, kind) of the depth-limited search."""
    m = u64(0)
    k = ctypes.c_int(0)
    v = oracle().orc_think(w16.ctypes.data, b16.ctypes.data, scale, p, o, depth, ctypes.byref(m), ctypes.byref(k))
    return v, m.value, k.value


def think_value(w16: np.ndarray, b16: np.ndarray, scale: float, p: int, o: int, depth: int) -> int:
    """Depth-limited value by alpha-beta (equal to think()'s value)."""
    return oracle().orc_think_value(w16.ctypes.data, b16.ctypes.data, scale, p, o, depth)


def think_leaf(w16: np.ndarray, b16: np.ndarray, scale: float, p: int, o: int) -> int:
    return oracle().orc_think_leaf(w16.ctypes.data, b16.ctypes.data, scale, p, o)


def think_batch(w16: np.ndarray, b16: np.ndarray, scale: float, boards: np.ndarray, depth: int, threads: int = 8):
    """(values, optimal-square masks, kinds) of (n,2) boards."""
    b = np.ascontiguousarray(boards, dtype=np.uint64).reshape(-1, 2)
    P = np.ascontiguousarray(b[:, 0])
    O = np.ascontiguousarray(b[:, 1])
    n = b.shape[0]
    scores = np.zeros(n, dtype=np.int8)
    masks = np.zeros(n, dtype=np.uint64)
    kinds = np.zeros(n, dtype=np.uint8)
    w = np.ascontiguousarray(w16, dtype=np.int16)
    bb = np.ascontiguousarray(b16, dtype=np.int16)
    oracle().orc_think_batch(w.ctypes.data, bb.ctypes.data, scale, P.ctypes.data, O.ctypes.data, n, depth, threads,
                             scores.ctypes.data, masks.ctypes.data, kinds.ctypes.data)
    return scores, masks, kinds


def minimax(p: int, o: int) -> int:
    n = u64(0)
    return oracle().orc_minimax(p, o, ctypes.byref(n))


def encode(p: int, o: int) -> str:
    buf = ctypes.create_string_buffer(17)
    oracle().orc_encode(p, o, buf)
    return buf.value.decode("latin-1")


def decode(s: str):
    p, o = u64(), u64()
    rc = oracle().orc_decode(s.encode("latin-1"), ctypes.byref(p), ctypes.byref(o))
    return (p.value, o.value) if rc == 0 else None


def empties(p: int, o: int) -> int:
    return 64 - bin(p | o).count("1")
