"""Learned form on the GPU: the BF16 MFMA product against the CPU (operands
rounded to BF16 the same way), the network's gradients against the CPU
reference (tests/mlp_ref.py), training that lowers the loss, and the guided
flash: same converged results as the plain flash, fewer iterations."""
import numpy as np
import pytest

from opencaeporo_amd import fluids, learned, lib
from tests import gpu_util, mlp_ref

pytestmark = pytest.mark.gpu


def torch():
    import torch as t
    return t


def test_mfma_gemm_matches_cpu():
    gpu_util.say("[learned] MFMA GEMM layout and split-K")
    t = torch()
    rng = np.random.default_rng(0)
    net = learned.MLP(6)
    for (M, N, K, splitK) in [(37, 64, 8, 1), (64, 7, 300, 4), (5, 19, 33, 1)]:
        A = rng.normal(size=(M, K)).astype(np.float32)
        B = rng.normal(size=(K, N)).astype(np.float32)
        ref = mlp_ref.to_bf16(A) @ mlp_ref.to_bf16(B)
        Ad, Bd = t.tensor(A, device="cuda:0"), t.tensor(B, device="cuda:0")
        Cd = t.zeros((M, N), dtype=t.float32, device="cuda:0")
        part = t.zeros(splitK * M * N, dtype=t.float32, device="cuda:0") if splitK > 1 else None
        net._gemm(M, N, K, Ad, K, 1, Bd, N, 1, Cd, N, 1, 0, splitK, part)
        got = Cd.cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
        # transposed operand through strides: C^T = B^T A^T
        Ct = t.zeros((N, M), dtype=t.float32, device="cuda:0")
        net._gemm(N, M, K, Bd, 1, N, Ad, 1, K, Ct, M, 1)
        np.testing.assert_allclose(Ct.cpu().numpy(), ref.T, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


def test_gradients_match_cpu_reference():
    gpu_util.say("[learned] gradients vs CPU reference")
    t = torch()
    rng = np.random.default_rng(3)
    NC = 6
    net = learned.MLP(NC, seed=5)
    B = 300
    X = rng.uniform(0, 1, (B, NC + 2)).astype(np.float32)
    lnK = rng.normal(0, 1, (B, NC)).astype(np.float32)
    label = (rng.uniform(size=B) > 0.5).astype(np.float32)
    dev = lambda a: t.tensor(a, device="cuda:0")
    loss, grads = net.loss_and_grads(dev(X), dev(lnK), dev(label), dev(label))
    W = [w.cpu().numpy().astype(np.float64) for w in net.W]
    b = [x.cpu().numpy().astype(np.float64) for x in net.b]
    l_ref, g_ref = mlp_ref.loss_and_grads(W, b, X, lnK, label, label)
    np.testing.assert_allclose(loss.cpu().numpy(), l_ref, rtol=2e-4, atol=1e-5)
    for g, r in zip(grads, g_ref):
        g = g.cpu().numpy()
        err = np.abs(g - r).max() / max(np.abs(r).max(), 1e-12)
        assert err < 2e-3, err


def test_training_and_guided_flash():
    gpu_util.say("[learned] train, then guided vs plain cold flash")
    t = torch()
    eos = lib.eos_from_deck(fluids.bench6_deck())
    net, tr = learned.train(eos, n_train=60_000, steps=600, batch=4096, seed=1)
    hist = tr["loss_history"]
    gpu_util.say("  loss", hist[0], "->", hist[-1])
    assert hist[-1][1] < 0.5 * hist[0][1]
    ev = learned.evaluate(eos, net, nb=20_000, seed=77)
    gpu_util.say("  ", ev)
    for kind in ("cold", "warm"):
        a = ev[kind]["agreement"]
        assert a["phase_count_same"] >= 0.999, kind
        assert a["p99_abs_S_diff"] < 1e-10, kind
    it_p, it_l = ev["cold"]["plain"]["iters"], ev["cold"]["learned"]["iters"]
    # the guessed K proves instability in fewer successive substitutions
    assert it_l["ssm_sta"] < it_p["ssm_sta"]


def test_fused_inference_matches_layer_path_and_cpu():
    """ocpg_mlp_forward (one kernel, weights as BF16 fragments in LDS,
    activations in registers) against the layer-by-layer GEMM path and the
    CPU reference with the same BF16 roundings: both differ from it only by
    FP32 accumulation order (and the rare BF16 rounding it flips)."""
    gpu_util.say("[learned] fused inference vs layer path")
    t = torch()
    rng = np.random.default_rng(9)
    for NC, B in ((6, 1000), (10, 333), (6, 17)):
        net = learned.MLP(NC, seed=NC)
        for l in range(4):
            net.b[l].copy_(t.tensor(rng.normal(0, 0.1, net.b[l].shape[0]).astype(np.float32), device="cuda:0"))
        X = rng.uniform(0, 1, (B, NC + 2)).astype(np.float32)
        Xd = t.tensor(X, device="cuda:0")
        fused = net.infer(Xd).cpu().numpy()
        layered = net.forward(Xd).cpu().numpy()
        W = [w.cpu().numpy() for w in net.W]
        b = [x.cpu().numpy() for x in net.b]
        ref = mlp_ref.forward(W, b, X)[0]
        scale = np.abs(ref).max()
        for got in (fused, layered):
            err = np.abs(got - ref) / scale
            assert err.max() < 2e-2 and np.quantile(err, 0.99) < 2e-3, (NC, B, err.max())
        assert np.abs(fused - layered).max() / scale < 2e-2


def test_committed_network_on_the_full_field():
    """The committed network (trained on 10M GPU flash results) on the
    1.122M-cell 6-component field, cold: converged results unchanged but for
    a handful of near-boundary cells where a trial from the guess finds an
    instability Wilson's four trials miss, fewer successive substitutions in
    the stability test."""
    gpu_util.say("[learned] committed network on the 1.122M-cell field")
    eos = lib.eos_from_deck(fluids.bench6_deck())
    net = learned.MLP.load(learned.weights_path(6))
    ev = learned.evaluate(eos, net, nb=1_122_000, warm=False)
    c = ev["cold"]
    gpu_util.say("  ", c["agreement"], "gain", c["gain"], "iters", c["plain"]["iters"], c["learned"]["iters"])
    a = c["agreement"]
    assert a["phase_count_differs"] <= 30
    assert a["p99_abs_S_diff"] < 1e-10 and a["max_abs_S_diff"] < 1e-8
    assert c["learned"]["iters"]["ssm_sta"] < 0.8 * c["plain"]["iters"]["ssm_sta"]
    assert c["learned"]["nonconverged"] <= c["plain"]["nonconverged"]
