"""The CPU restatement of the step controller's MLP (oracle/mlp_ref.py) pinned
by what fixes it independently: finite differences of its own loss for the
gradient, and torch's float32 -> bfloat16 conversion for the weight rounding."""
import numpy as np
import pytest


def test_bf16_rounding_is_torchs():
    torch = pytest.importorskip("torch")
    from oracle import mlp_ref
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.uniform(-30, 30, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.3895314e38, 1e-40], np.float32)])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    got = mlp_ref.bf16(x)
    assert np.array_equal(got, ref)


def test_gradient_matches_finite_differences():
    from oracle import mlp_ref
    rng = np.random.default_rng(1)
    p = mlp_ref.init(3).astype(np.float64)
    # round the weights once, so that the finite differences see the same (BF16) weights as the gradient
    d = mlp_ref.unpack(p)
    for k in ("W1", "W2", "W3"):
        d[k] = mlp_ref.bf16(d[k])
    p = mlp_ref.pack(d)
    X = rng.standard_normal((40, 16))
    T = rng.random((40, 2))
    W = rng.random(40) + 0.5
    loss, g, _ = mlp_ref.loss_grad(p, X, T, W)

    def f(q):  # the same loss without re-rounding the perturbed weights
        dd = mlp_ref.unpack(q)
        z1 = X @ dd["W1"].T + dd["b1"]
        h1 = np.maximum(z1, 0)
        z2 = h1 @ dd["W2"].T + dd["b2"]
        h2 = np.maximum(z2, 0)
        y = h2 @ dd["W3"].T + dd["b3"]
        return 0.5 * np.sum(W[:, None] * (y - T) ** 2) / W.sum()

    assert abs(f(p) - loss) <= 1e-12 * max(1.0, abs(loss))
    idx = rng.choice(mlp_ref.NPARAM, 300, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(mlp_ref.NPARAM)
        e[i] = h
        fd = (f(p + e) - f(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_action_mapping():
    from oracle import mlp_ref
    p = np.zeros(mlp_ref.NPARAM, np.float32)
    d = mlp_ref.unpack(p)
    d["b3"][:] = [2.0, -1.0]  # outputs clamp to [0, 1]
    n2f, lt = mlp_ref.action(mlp_ref.pack(d), np.zeros(16))
    assert n2f == 1.0 and lt == -16.0
