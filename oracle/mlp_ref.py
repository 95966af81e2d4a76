"""CPU restatement (float64 numpy) of the mixed solver's step-controller MLP.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the library's forward
and backward kernels (jkg_ctl_grad) and its in-graph controller.  The network
has no counterpart in the reference.  Layout of the parameter blob (5378
floats): W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]; the forward
pass uses the weights rounded to BF16 (nearest even) and FP32 biases.
"""
import numpy as np

IN, H, OUT = 16, 64, 2
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)), ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shp in SIZES:
        k = int(np.prod(shp))
        out[name] = p[o:o + k].reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def init(seed=0, scale=1.0):
    """He-style initialisation (a starting point for training)."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.standard_normal((H, IN)) * np.sqrt(2.0 / IN) * scale, "b1": np.zeros(H),
         "W2": rng.standard_normal((H, H)) * np.sqrt(2.0 / H) * scale, "b2": np.zeros(H),
         "W3": rng.standard_normal((OUT, H)) * np.sqrt(1.0 / H) * scale, "b3": np.full(OUT, 0.5)}
    return pack(d).astype(np.float32)


def forward(p, X):
    d = unpack(p)
    W1, W2, W3 = bf16(d["W1"]), bf16(d["W2"]), bf16(d["W3"])
    z1 = X @ W1.T + d["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + d["b2"]
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3.T + d["b3"]
    return y, (W1, W2, W3, z1, h1, z2, h2)


def loss_grad(p, X, T, W=None):
    """Loss sum_s w_s |y_s - t_s|^2 / (2 sum w) and its gradient with respect to the
    blob (gradients of the BF16-rounded weights, as the library applies them)."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    w = np.ones(len(X)) if W is None else np.asarray(W, dtype=np.float64)
    ws = w.sum()
    y, (W1, W2, W3, z1, h1, z2, h2) = forward(p, X)
    r = y - T
    loss = 0.5 * np.sum(w[:, None] * r * r) / ws
    dy = w[:, None] * r / ws
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    d2 = (dy @ W3) * (z2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    return loss, pack(g), y


def action(p, f):
    """The controller's action for one feature vector: (N_2 fraction, log10 tau)."""
    y, _ = forward(p, np.asarray(f, dtype=np.float64)[None, :])
    y = np.clip(y[0], 0.0, 1.0)
    return 0.75 + 0.25 * y[0], -16.0 + 10.0 * y[1]
