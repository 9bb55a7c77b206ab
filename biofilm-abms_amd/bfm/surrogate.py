"""JKR surrogate: GPU training, evaluation and weight files (over libbfm.so).

The network and its training kernels live in csrc/bfm_surrogate.hip; this
module drives them through the C ABI and stores trained weights as .npz
(FP32 master parameters, BF16 copies are made on upload).
"""
import ctypes as C
import os

import numpy as np

from .core import BFM_OK, BFMError, Context, load_library

NPARAMS = 64 * 4 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                               "data", "jkr_surrogate.npz")
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


def _bind(lib):
    if getattr(lib, "_sur_bound", False):
        return lib
    vp = C.c_void_p
    lib.bfm_surrogate_train.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64,
                                        _fp, _fp]
    lib.bfm_surrogate_eval.argtypes = [vp, _dp, _dp, C.c_int, C.c_int, _fp, _dp]
    lib.bfm_surrogate_grad.argtypes = [vp, _fp, _fp, _fp, C.c_int, _fp, _fp, _fp]
    lib.bfm_surrogate_generate.argtypes = [vp, C.c_int, C.c_uint64, _fp, _fp,
                                           C.POINTER(C.c_int)]
    lib.bfm_surrogate_set_polish.argtypes = [vp, C.c_int]
    for f in ("bfm_surrogate_train", "bfm_surrogate_eval", "bfm_surrogate_grad",
              "bfm_surrogate_generate", "bfm_surrogate_set_polish"):
        getattr(lib, f).restype = C.c_int
    lib._sur_bound = True
    return lib


def split(params):
    p = np.asarray(params, dtype=np.float32)
    o = [0]
    for n in (256, 64, 4096, 64, 4096, 64, 64, 1):
        o.append(o[-1] + n)
    return [p[o[k]:o[k + 1]] for k in range(8)]


def to_bf16_bits(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def install(ctx: Context, params, envelope=None):
    """bfm_set_surrogate with FP32 master parameters."""
    w1, b1, w2, b2, w3, b3, w4, b4 = split(params)
    bits = [np.ascontiguousarray(to_bf16_bits(w)) for w in (w1, w2, w3, w4)]
    bs = [np.ascontiguousarray(b, dtype=np.float32) for b in (b1, b2, b3, b4)]
    env = None if envelope is None else np.ascontiguousarray(envelope, dtype=np.float64)
    rc = ctx.lib.bfm_set_surrogate(ctx.h, bits[0].ctypes.data, bs[0].ctypes.data, bits[1].ctypes.data,
                                   bs[1].ctypes.data, bits[2].ctypes.data, bs[2].ctypes.data,
                                   bits[3].ctypes.data, bs[3].ctypes.data,
                                   None if env is None else env.ctypes.data_as(_dp))
    ctx._check(rc)


def train(ctx: Context, n_samples=1_000_000, epochs=30, batch=4096, lr=3e-3, seed=42):
    lib = _bind(load_library())
    params = np.zeros(NPARAMS, dtype=np.float32)
    losses = np.zeros(epochs, dtype=np.float32)
    ctx._check(lib.bfm_surrogate_train(ctx.h, n_samples, epochs, batch, lr, seed,
                                       params.ctypes.data_as(_fp), losses.ctypes.data_as(_fp)))
    return params, losses


def evaluate(ctx: Context, delta, gamma, polish=1):
    lib = _bind(load_library())
    d = np.ascontiguousarray(delta, dtype=np.float64)
    g = np.ascontiguousarray(gamma, dtype=np.float64)
    t = np.zeros(len(d), dtype=np.float32)
    r = np.zeros(len(d))
    ctx._check(lib.bfm_surrogate_eval(ctx.h, d.ctypes.data_as(_dp), g.ctypes.data_as(_dp), len(d),
                                      polish, t.ctypes.data_as(_fp), r.ctypes.data_as(_dp)))
    return t, r


def grad(ctx: Context, params, X, T):
    lib = _bind(load_library())
    p = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    B = len(T)
    loss = C.c_float()
    y = np.zeros(B, dtype=np.float32)
    g = np.zeros(NPARAMS, dtype=np.float32)
    ctx._check(lib.bfm_surrogate_grad(ctx.h, p.ctypes.data_as(_fp), X.ctypes.data_as(_fp),
                                      T.ctypes.data_as(_fp), B, C.byref(loss),
                                      y.ctypes.data_as(_fp), g.ctypes.data_as(_fp)))
    return loss.value, y, g


def generate(ctx: Context, n, seed=42):
    lib = _bind(load_library())
    X = np.zeros((n, 4), dtype=np.float32)
    T = np.zeros(n, dtype=np.float32)
    ok = np.zeros(n, dtype=np.int32)
    ctx._check(lib.bfm_surrogate_generate(ctx.h, n, seed, X.ctypes.data_as(_fp),
                                          T.ctypes.data_as(_fp),
                                          ok.ctypes.data_as(C.POINTER(C.c_int))))
    return X, T, ok


def set_polish(ctx: Context, steps):
    lib = _bind(load_library())
    ctx._check(lib.bfm_surrogate_set_polish(ctx.h, int(steps)))


def save(path, params, losses=None, **meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=np.asarray(params, dtype=np.float32),
             losses=np.asarray([] if losses is None else losses, dtype=np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def load(path=DEFAULT_WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)
