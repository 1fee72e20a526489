"""Learned move ordering: data, GPU training and the int16 evaluator file.

The evaluator is a linear pattern model (46 ternary instances in 11 classes
plus a bias per number of empties) predicting the exact disc difference for
the side to move.  Training data are random-playout positions labelled by
this project's exact GPU solver; training runs on the GPU through the HIP
kernels of libgo2 (features, SGD forward/gradient, count-normalised update,
int16 quantisation).  Solves with ORDER_LEARNED use the installed int16
tables to order moves at shallow nodes; scores never depend on them.
"""
from __future__ import annotations

import ctypes
import json
import os
import sys
import time

import numpy as np

from . import (BOARD_DTYPE, ORDER_MOBILITY, Task, _check, default_config, generate, lib,
               pattern_info, set_evaluator)
from .hip import DeviceBuffer

WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "pattern_int16.npz")
SCALE = 64.0  # int16 units per disc


def log(msg: str) -> None:
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def label(boards: np.ndarray, empties: int, chunk: int = 250000) -> np.ndarray:
    """Exact scores of `boards` from the GPU solver (mobility ordering)."""
    out = np.zeros(len(boards), dtype=np.int8)
    cfg = default_config(ordering=ORDER_MOBILITY)
    t = Task(min(chunk, len(boards)), empties, cfg)
    for i in range(0, len(boards), chunk):
        b = boards[i:i + chunk]
        t.upload(b)
        t.launch()
        t.wait()
        s, st = t.download()
        out[i:i + len(b)] = s
        log(f"[learn] labelled {i + len(b)}/{len(boards)} ({st['kernel_ms']:.0f} ms)")
    t.close()
    return out


def make_dataset(n: int, empties: int, seed: int, config_id: int = 40):
    boards = generate(seed, config_id, n, empties)
    y = label(boards, empties)
    return boards, y


class Trainer:
    """FP32 master weights on the GPU, trained by libgo2's SGD kernels."""

    def __init__(self):
        self.ni, self.nw, self.nb = pattern_info()
        self.w = DeviceBuffer(self.nw * 4)
        self.bias = DeviceBuffer(self.nb * 4)
        self.grad = DeviceBuffer((self.nw + self.nb) * 4)
        self.cnt = DeviceBuffer((self.nw + self.nb) * 4)
        self.sse = DeviceBuffer(8)
        zero_w = np.zeros(self.nw, dtype=np.float32)
        self.w.upload(zero_w)
        self.bias.upload(np.zeros(self.nb, dtype=np.float32))
        self.grad.upload(np.zeros(self.nw + self.nb, dtype=np.float32))
        self.cnt.upload(np.zeros(self.nw + self.nb, dtype=np.uint32))

    def load_data(self, boards: np.ndarray, y: np.ndarray):
        """Upload boards, compute their features on the GPU; returns device handles."""
        n = len(boards)
        b = DeviceBuffer(n * 16)
        b.upload(np.ascontiguousarray(boards, dtype=BOARD_DTYPE))
        feat = DeviceBuffer(n * self.ni * 4)
        emp = DeviceBuffer(n)
        _check(lib().go2_pattern_features_device(b.addr, n, feat.addr, emp.addr, None))
        yd = DeviceBuffer(n * 4)
        yd.upload(np.ascontiguousarray(y, dtype=np.float32))
        return {"n": n, "boards": b, "feat": feat, "emp": emp, "y": yd}

    def _sse_reset(self):
        self.sse.upload(np.zeros(1, dtype=np.float64))

    def _sse_read(self) -> float:
        v = np.zeros(1, dtype=np.float64)
        self.sse.download(v)
        return float(v[0])

    def step(self, data, start: int, count: int, lr: float) -> None:
        L = lib()
        _check(L.go2_sgd_gradient_device(self.w.addr, self.bias.addr, data["feat"].addr + start * self.ni * 4,
                                         data["emp"].addr + start, data["y"].addr + start * 4, count,
                                         self.grad.addr, self.cnt.addr, self.sse.addr, None))
        _check(L.go2_sgd_apply_device(self.w.addr, self.bias.addr, self.grad.addr, self.cnt.addr, lr, None))

    def mse(self, data, batch: int = 65536) -> float:
        """Mean squared error on `data` (gradients computed, then discarded)."""
        L = lib()
        self._sse_reset()
        for s in range(0, data["n"], batch):
            c = min(batch, data["n"] - s)
            _check(L.go2_sgd_gradient_device(self.w.addr, self.bias.addr, data["feat"].addr + s * self.ni * 4,
                                             data["emp"].addr + s, data["y"].addr + s * 4, c,
                                             self.grad.addr, self.cnt.addr, self.sse.addr, None))
            # discard: lr = 0 clears grad/count without moving the weights
            _check(L.go2_sgd_apply_device(self.w.addr, self.bias.addr, self.grad.addr, self.cnt.addr, 0.0, None))
        return self._sse_read() / data["n"]

    def train(self, tr, va, epochs: int, batch: int, lr: float, decay: float = 0.75):
        # Every prediction sums 47 weights and each seen weight moves by
        # lr * (mean 2 * residual), so lr near 1/47 takes a full step.
        hist = []
        for ep in range(epochs):
            t0 = time.time()
            for s in range(0, tr["n"], batch):
                self.step(tr, s, min(batch, tr["n"] - s), lr)
            m_tr, m_va = self.mse(tr), self.mse(va)
            hist.append({"epoch": ep + 1, "lr": lr, "train_mse": m_tr, "valid_mse": m_va,
                         "seconds": time.time() - t0})
            log(f"[learn] epoch {ep + 1}: train mse {m_tr:.2f}, valid mse {m_va:.2f}, lr {lr:.3f}")
            lr *= decay
        return hist

    def quantize(self, scale: float = SCALE):
        w16 = DeviceBuffer(self.nw * 2)
        b16 = DeviceBuffer(self.nb * 2)
        _check(lib().go2_quantize_weights_device(self.w.addr, self.bias.addr, scale, w16.addr, b16.addr, None))
        hw = np.zeros(self.nw, dtype=np.int16)
        hb = np.zeros(self.nb, dtype=np.int16)
        w16.download(hw)
        b16.download(hb)
        return hw, hb

    def weights(self):
        hw = np.zeros(self.nw, dtype=np.float32)
        hb = np.zeros(self.nb, dtype=np.float32)
        self.w.download(hw)
        self.bias.download(hb)
        return hw, hb


def save(path: str, w16: np.ndarray, b16: np.ndarray, scale: float, meta: dict) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, w16=w16, b16=b16, scale=np.float32(scale), meta=json.dumps(meta))


def load(path: str = WEIGHTS_PATH):
    d = np.load(path, allow_pickle=False)
    return d["w16"], d["b16"], float(d["scale"]), json.loads(str(d["meta"]))


def export_raw(out_path: str, path: str = WEIGHTS_PATH) -> None:
    """Raw little-endian int16 weights then biases (the CLI's GO2_WEIGHTS file)."""
    w16, b16, _, _ = load(path)
    with open(out_path, "wb") as f:
        f.write(np.ascontiguousarray(w16, dtype="<i2").tobytes())
        f.write(np.ascontiguousarray(b16, dtype="<i2").tobytes())


def install(path: str = WEIGHTS_PATH) -> dict:
    """Install the saved int16 evaluator for ORDER_LEARNED solves; returns its metadata."""
    w16, b16, scale, meta = load(path)
    set_evaluator(w16, b16, scale)
    return meta


def train_main(n: int = 2000000, empties: int = 14, seed: int = 1401, epochs: int = 8,
               batch: int = 4096, lr: float = 0.01, out: str = WEIGHTS_PATH) -> dict:
    t0 = time.time()
    boards, y = make_dataset(n, empties, seed)
    t_label = time.time() - t0
    k = int(n * 0.8)  # 80/20 split
    tr_b, tr_y, va_b, va_y = boards[:k], y[:k], boards[k:], y[k:]
    tr = Trainer()
    dtr, dva = tr.load_data(tr_b, tr_y), tr.load_data(va_b, va_y)
    base = float(np.mean((va_y.astype(np.float64) - tr_y.mean()) ** 2))
    hist = tr.train(dtr, dva, epochs, batch, lr)
    w16, b16 = tr.quantize(SCALE)
    meta = {"positions": n, "empties": empties, "seed": seed, "split": "80/20", "epochs": epochs,
            "batch": batch, "lr": lr, "scale": SCALE, "label_seconds": t_label,
            "train_seconds": time.time() - t0 - t_label, "valid_mse_constant_predictor": base,
            "train_mse": hist[-1]["train_mse"], "valid_mse": hist[-1]["valid_mse"], "history": hist}
    save(out, w16, b16, SCALE, meta)
    log(f"[learn] wrote {out}")
    return meta


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="Train the int16 move-ordering evaluator on the GPU")
    ap.add_argument("--n", type=int, default=2000000)
    ap.add_argument("--empties", type=int, default=14)
    ap.add_argument("--seed", type=int, default=1401)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=0.01)
    ap.add_argument("--out", default=WEIGHTS_PATH)
    a = ap.parse_args()
    m = train_main(a.n, a.empties, a.seed, a.epochs, a.batch, a.lr, a.out)
    print(json.dumps({k: v for k, v in m.items() if k != "history"}))
