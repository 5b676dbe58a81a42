"""ctypes face of the learned ABC closure (lps_closure_* in include/lps_amd.h).

A `Closure` is the network (1600 -> 256 -> 256 -> 256 -> 4) with its Adam state and a sample store
on one GPU. It has no counterpart in the reference, whose ABC is the least-squares G_u / G_v update
(src/main.cpp:883-974) that the closure corrects. There is no CPU fallback: without the HIP
library or a GPU, creating a closure raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from .solver import LPS_OK, LpsError, _lps_problem, _ptr, load_library

N_IN = 1600
N_HID = 256
N_OUT = 4

# flat parameter layout (weights [out][in])
SHAPES = (("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_HID, N_HID)), ("b3", (N_HID,)), ("W4", (N_OUT, N_HID)), ("b4", (N_OUT,)))


def split_params(p: np.ndarray):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = p[o:o + n].reshape(shp)
        o += n
    return out


def join_params(d) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float32).ravel() for name, _ in SHAPES])


class Closure:
    def __init__(self, device: int = 0, seed: int = 0):
        L = load_library()
        h = C.c_void_p()
        rc = L.lps_closure_create(device, seed, C.byref(h))
        if rc != LPS_OK:
            raise LpsError(rc, "lps_closure_create failed (no HIP device?)")
        self._L, self._h = L, h
        self.n_params = int(L.lps_closure_param_count())

    def close(self):
        if getattr(self, "_h", None):
            self._L.lps_closure_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc < 0:
            raise LpsError(rc, self._L.lps_closure_error(self._h).decode())
        return rc

    # parameters
    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        self._check(self._L.lps_closure_get_params(self._h, _ptr(p)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self._check(self._L.lps_closure_set_params(self._h, _ptr(p)))

    def set_scales(self, sigma_u: float, sigma_v: float):
        self._check(self._L.lps_closure_set_scales(self._h, float(sigma_u), float(sigma_v)))

    def get_scales(self):
        s = np.zeros(2, dtype=np.float32)
        self._check(self._L.lps_closure_get_scales(self._h, _ptr(s)))
        return float(s[0]), float(s[1])

    # compute
    def forward(self, x_bits: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x_bits, dtype=np.uint16)
        n = x.shape[0]
        out = np.zeros((n, N_OUT), dtype=np.float32)
        self._check(self._L.lps_closure_forward(self._h, _ptr(x), n, _ptr(out)))
        return out

    def forward_noalloc(self, x_bits: np.ndarray) -> np.ndarray:
        """forward() through lps_closure_forward_noalloc: no device allocation, at most 2048 rows."""
        x = np.ascontiguousarray(x_bits, dtype=np.uint16)
        n = x.shape[0]
        out = np.zeros((n, N_OUT), dtype=np.float32)
        self._check(self._L.lps_closure_forward_noalloc(self._h, _ptr(x), n, _ptr(out)))
        return out

    def train_batch(self, x_bits, target, weight, lr: float, want_grads: bool = False):
        x = np.ascontiguousarray(x_bits, dtype=np.uint16)
        t = np.ascontiguousarray(target, dtype=np.float32)
        w = np.ascontiguousarray(weight, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32) if want_grads else None
        self._check(self._L.lps_closure_train_batch(self._h, _ptr(x), _ptr(t), _ptr(w),
                                                    x.shape[0], float(lr), _ptr(loss), _ptr(g)))
        return float(loss[0]), g

    # samples
    def reserve(self, capacity: int):
        self._check(self._L.lps_closure_dataset_reserve(self._h, int(capacity)))
        self._cap = max(getattr(self, "_cap", 0), int(capacity))

    def capacity(self) -> int:
        """Samples the store holds (appends beyond it are counted but dropped)."""
        return getattr(self, "_cap", 0)

    def count(self) -> int:
        return self._check(self._L.lps_closure_dataset_count(self._h))

    def clear(self):
        self._check(self._L.lps_closure_dataset_clear(self._h))

    def append(self, x_bits, aux):
        x = np.ascontiguousarray(x_bits, dtype=np.uint16)
        a = np.ascontiguousarray(aux, dtype=np.float32)
        self._check(self._L.lps_closure_dataset_append(self._h, _ptr(x), _ptr(a), x.shape[0]))

    def read(self, first: int, n: int):
        x = np.zeros((n, N_IN), dtype=np.uint16)
        a = np.zeros((n, 8), dtype=np.float32)
        self._check(self._L.lps_closure_dataset_read(self._h, first, n, _ptr(x), _ptr(a)))
        return x, a

    def train(self, nsteps: int, batch: int, lr0: float, lr1: float, first: int, count: int,
              seed: int = 0, input_noise: float = 0.0) -> np.ndarray:
        curve = np.zeros(max(nsteps, 1), dtype=np.float32)
        self._check(self._L.lps_closure_train(self._h, nsteps, batch, lr0, lr1, first, count,
                                              seed, float(input_noise), _ptr(curve)))
        return curve[:nsteps]

    def evaluate(self, first: int, count: int):
        """(sum w |f - t/sigma|^2, sum w |t/sigma|^2, sum w) over samples [first, first+count)."""
        out = np.zeros(3, dtype=np.float64)
        self._check(self._L.lps_closure_evaluate(self._h, first, count, _ptr(out)))
        return float(out[0]), float(out[1]), float(out[2])

    def collect_closed(self, truth, trunc, nsteps: int, first_step: int = 8, sample_every: int = 1,
                       keep: float = 1.0, amp_min: float = 0.0, y_max: float = 1e30, seed: int = 0):
        """Closed-loop samples: `trunc` (a Solver with this closure attached) and `truth` (a Solver
        on a larger domain without ABC) advance nsteps together; samples append to the store."""
        rc = self._L.lps_closure_collect_closed(self._h, truth._h, trunc._h, nsteps, first_step,
                                                sample_every, float(keep), float(amp_min),
                                                float(y_max), seed)
        if rc < 0:  # the message sits on the truncated context
            raise LpsError(rc, self._L.lps_last_error(trunc._h).decode())
        return self.count()

    def collect(self, truth, trunc_prob, nsteps: int, first_step: int = 8, sample_every: int = 1,
                keep: float = 1.0, amp_min: float = 0.0, y_max: float = 1e30, seed: int = 0):
        """Teacher-forced samples: run `truth` (a Solver on a larger domain) for nsteps steps and
        append samples for the ABC rows of trunc_prob (a preprocess.Problem)."""
        keep_arrays = []
        p = _lps_problem(trunc_prob, keep_arrays)
        rc = self._L.lps_closure_collect(self._h, truth._h, C.byref(p), nsteps, first_step,
                                         sample_every, float(keep), float(amp_min), float(y_max),
                                         seed)
        if rc < 0:  # the message sits on the truth context
            raise LpsError(rc, self._L.lps_last_error(truth._h).decode())
        return self.count()


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> BF16 bit patterns, round to nearest even."""
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = ((b >> 16) & 1) + 0x7FFF
    return ((b + r) >> 16).astype(np.uint16)


def bf16_to_f32(bits: np.ndarray) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)
