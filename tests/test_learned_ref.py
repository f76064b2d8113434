"""Pins the CPU reference of the learned form's network (tests/mlp_ref.py)
against finite differences of its own loss, in float64: the gradients the
GPU kernels are checked against are the derivatives of the loss."""
import numpy as np

from tests import mlp_ref


def test_reference_gradients_are_derivatives():
    rng = np.random.default_rng(1)
    NC = 6
    dims = [NC + 2, 16, 16, 16, NC + 1]
    W = [rng.normal(0, 0.5, (dims[l], dims[l + 1])) for l in range(4)]
    b = [rng.normal(0, 0.1, dims[l + 1]) for l in range(4)]
    B = 7
    X = rng.uniform(0, 1, (B, NC + 2))
    lnK = rng.normal(0, 1, (B, NC))
    label = (rng.uniform(size=B) > 0.5).astype(float)
    mask = label.copy()
    loss, grads = mlp_ref.loss_and_grads(W, b, X, lnK, label, mask, bf16=False)
    params = W + b
    h = 1e-6
    for p, g in zip(params, grads):
        for idx in list(np.ndindex(p.shape))[:12]:
            old = p[idx]
            p[idx] = old + h
            lp = mlp_ref.loss_and_grads(W, b, X, lnK, label, mask, bf16=False)[0].mean()
            p[idx] = old - h
            lm = mlp_ref.loss_and_grads(W, b, X, lnK, label, mask, bf16=False)[0].mean()
            p[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, g[idx])


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 sits halfway between two BF16 values: ties go to the even one
    x = np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0], dtype=np.float32)
    r = mlp_ref.to_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 + 2 * 2.0 ** -7 and r[2] == -2.5 and r[3] == 0.0
