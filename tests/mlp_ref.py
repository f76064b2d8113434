"""CPU reference of the learned form's network (test infrastructure): the
same layers, loss and gradients in numpy.  With bf16=True the operands of
every product are rounded to BF16 as the GPU's MFMA kernel rounds them;
with bf16=False it is plain float64 (used to pin the backward pass against
finite differences)."""
import numpy as np
from scipy.special import erf


def to_bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1 + erf(x / np.sqrt(2)))


def gelu_grad(x):
    return 0.5 * (1 + erf(x / np.sqrt(2))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def mm(a, b, bf16):
    if bf16:
        return to_bf16(a) @ to_bf16(b)
    return np.asarray(a, dtype=np.float64) @ np.asarray(b, dtype=np.float64)


def forward(W, b, X, bf16=True):
    H = np.asarray(X, dtype=np.float64)
    Zs, Hs = [], [H]
    for l in range(4):
        Z = mm(H, W[l], bf16) + b[l]
        Zs.append(Z)
        if l < 3:
            H = gelu(Z)
            Hs.append(H)
        else:
            H = Z
    return H, Zs, Hs


def loss_and_grads(W, b, X, lnK, label, mask, w_stab=1.0, bf16=True):
    B = X.shape[0]
    NC = lnK.shape[1]
    Y, Zs, Hs = forward(W, b, X, bf16)
    e = Y[:, :NC] - lnK
    z = Y[:, NC]
    loss = (mask[:, None] * e * e / NC).sum(axis=1) + w_stab * (np.maximum(z, 0) - z * label +
                                                                np.log1p(np.exp(-np.abs(z))))
    dY = np.zeros_like(Y)
    dY[:, :NC] = mask[:, None] * 2 * e / NC / B
    dY[:, NC] = w_stab * (1 / (1 + np.exp(-z)) - label) / B
    gW, gb = [None] * 4, [None] * 4
    dZ = dY
    for l in range(3, -1, -1):
        gW[l] = mm(Hs[l].T, dZ, bf16)
        gb[l] = dZ.sum(axis=0)
        if l > 0:
            dH = mm(dZ, np.asarray(W[l]).T, bf16)
            dZ = dH * gelu_grad(Zs[l - 1])
    return loss, gW + gb
