"""CPU reference of the JKR surrogate network (TEST INFRASTRUCTURE ONLY).

float64 numpy restatement of the network the GPU kernels run: inputs and
hidden activations rounded to BF16 (round to nearest even), weights rounded
to BF16 for the forward, SiLU, last layer a dot product, loss = mean squared
error. The gradient is taken with respect to the FP32 master parameters
through the BF16 rounding (straight-through), as the GPU backward does.
Parameter layout: W1[64][4] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4.
"""
import numpy as np

SIZES = [("W1", (64, 4)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)),
         ("b3", (64,)), ("W4", (64,)), ("b4", (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(x):
    """Round to float32, then to bfloat16 (nearest even); returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    out, k = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(params[k:k + n], dtype=np.float64).reshape(shape)
        k += n
    return out


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def forward(params, X):
    P = unpack(params)
    x = bf16(X)
    acts = {"x": x}
    h = x
    for i, (W, b) in enumerate((("W1", "b1"), ("W2", "b2"), ("W3", "b3")), start=1):
        Wq = bf16(P[W])
        z = h @ Wq.T + P[b]
        h = bf16(silu(z))
        acts[f"z{i}"], acts[f"h{i}"] = z, h
    y = h @ bf16(P["W4"]) + P["b4"][0]
    return y, acts


def loss_and_grad(params, X, T):
    P = unpack(params)
    y, a = forward(params, X)
    B = X.shape[0]
    e = y - T
    loss = float(np.mean(e * e))
    dy = 2 * e / B
    g = {"W4": dy @ a["h3"], "b4": np.array([dy.sum()])}
    dh = np.outer(dy, bf16(P["W4"]))
    for i, (W, b) in ((3, ("W3", "b3")), (2, ("W2", "b2")), (1, ("W1", "b1"))):
        dz = dh * dsilu(a[f"z{i}"])
        hin = a[f"h{i - 1}"] if i > 1 else a["x"]
        g[W] = dz.T @ hin
        g[b] = dz.sum(0)
        dh = dz @ bf16(P[W])
    flat = np.concatenate([g[name].ravel() for name, _ in SIZES])
    return loss, y, flat
