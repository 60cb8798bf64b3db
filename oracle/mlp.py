"""FP64 restatement of the order-model MLP (test infrastructure).

No counterpart in the reference: the reference uses one expansion order for
every M2L pair (3d/kernel.h:233-260).  This is the checker for the device
kernels in minimal_amd/csrc/order.hip: same layout, same BF16 rounding of the
weights, FP64 arithmetic.  Pinned by central finite differences in
tests/test_order_model.py.

Layout of the 1579 parameters: W1[32][4], b1[32], W2[32][32], b2[32],
W3[11][32], b3[11]; ReLU after layers 1 and 2.  Loss = sum over samples and
non-NaN outputs of (y - t)^2 / n_samples.
"""
import numpy as np

N_IN, N_H, N_OUT = 4, 32, 11
SIZES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_OUT, N_H)),
         ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def bf16_round(w):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def unpack(w):
    w = np.asarray(w, np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = w[o:o + n].reshape(shape)
        o += n
    return out


def forward(w, X, round_bf16=True):
    """Outputs [n, 11] and the hidden activations."""
    wv = bf16_round(w).astype(np.float64) if round_bf16 else np.asarray(w, np.float64)
    p = unpack(wv)
    X = np.asarray(X, np.float64)
    a1 = X @ p["W1"].T + p["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ p["W3"].T + p["b3"]
    return y, (X, a1, h1, a2, h2)


def loss_and_grad(w, X, T, round_bf16=True):
    """Loss and its gradient with respect to the (rounded) weights."""
    wv = bf16_round(w).astype(np.float64) if round_bf16 else np.asarray(w, np.float64)
    p = unpack(wv)
    y, (X, a1, h1, a2, h2) = forward(wv, X, round_bf16=False)
    T = np.asarray(T, np.float64)
    ok = ~np.isnan(T)
    n = len(X)
    e = np.where(ok, y - np.where(ok, T, 0.0), 0.0)
    loss = float(np.sum(e * e) / n)
    dy = 2.0 * e / n
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dh2 = (dy @ p["W3"]) * (a2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ p["W2"]) * (a1 > 0)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
