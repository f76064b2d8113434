"""Learned form: a small MLP that proposes the flash's initial K-values and a
stability prior, trained on the GPU on flash results the simulator itself
produces.

Network: features (P/5000, T/600, z_1..z_NC) -> 64 -> 64 -> 64 (GELU) ->
(ln K_1..ln K_NC, stability logit).  Every product runs on the library's BF16
MFMA kernel with FP32 accumulation (`ocpg_mlp_gemm`); bias/GELU, the loss
and Adam are the library's kernels too.  PyTorch only holds device memory and
provides the stream.

Use: `ocpg_flash_fim_guided` takes the network's output.  The FP64
stability test and split still decide every cell; the guess only changes
where the iterations start (extra stability trials from the guessed K, the
split from the guessed K instead of Wilson's), so converged results agree
with the plain flash within the flash's own tolerances and the network only
cuts iterations.

    python -m opencaeporo_amd.learned      # train + evaluate on the bench field
"""
import ctypes as C
import json
import sys
import time

import numpy as np

from . import abi, fluids, lib, workloads

HIDDEN = 64


def _torch():
    import torch
    return torch


class MLP:
    """FP32 master weights on the device; W_l is (in x out) row-major."""

    def __init__(self, NC, seed=0, device="cuda:0"):
        torch = _torch()
        self.NC = NC
        self.nin = NC + 2
        self.nout = NC + 1
        dims = [self.nin, HIDDEN, HIDDEN, HIDDEN, self.nout]
        self.dims = dims
        g = np.random.default_rng(seed)
        self.W, self.b = [], []
        for l in range(4):
            w = g.normal(0.0, np.sqrt(2.0 / dims[l]), (dims[l], dims[l + 1])).astype(np.float32)
            self.W.append(torch.tensor(w, device=device))
            self.b.append(torch.zeros(dims[l + 1], dtype=torch.float32, device=device))
        self.params = self.W + self.b
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.step = 0
        self.L = lib.load()

    # ---- helpers over the C-ABI ----
    def _stream(self):
        return C.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def _gemm(self, M, N, K, A, sAm, sAk, B, sBk, sBn, Cm, sCm, sCn, accumulate=0, splitK=1, partial=None):
        rc = self.L.ocpg_mlp_gemm(self._stream(), M, N, K, A.data_ptr(), sAm, sAk, B.data_ptr(), sBk, sBn,
                                  Cm.data_ptr(), sCm, sCn, accumulate, splitK,
                                  partial.data_ptr() if partial is not None else None)
        lib.check(rc, "ocpg_mlp_gemm")

    def forward(self, X, keep=False):
        """X: (B, nin) float32 on the device.  Returns Y (B, nout); with keep,
        also the pre-activations Z_l and activations H_l for backward."""
        torch = _torch()
        B = X.shape[0]
        H = X
        Zs, Hs = [], [X]
        for l in range(4):
            n_in, n_out = self.dims[l], self.dims[l + 1]
            Z = torch.empty((B, n_out), dtype=torch.float32, device=X.device)
            self._gemm(B, n_out, n_in, H, n_in, 1, self.W[l], n_out, 1, Z, n_out, 1)
            Hn = torch.empty_like(Z) if l < 3 else None
            lib.check(self.L.ocpg_mlp_bias_act(self._stream(), B, n_out, Z.data_ptr(), self.b[l].data_ptr(),
                                               Hn.data_ptr() if Hn is not None else None), "ocpg_mlp_bias_act")
            Zs.append(Z)
            if l < 3:
                H = Hn
                Hs.append(Hn)
            else:
                H = Z
        return (H, Zs, Hs) if keep else H

    def loss_and_grads(self, X, lnK, label, mask, w_stab=1.0):
        torch = _torch()
        B = X.shape[0]
        Y, Zs, Hs = self.forward(X, keep=True)
        dY = torch.empty_like(Y)
        loss = torch.empty(B, dtype=torch.float32, device=X.device)
        lib.check(self.L.ocpg_mlp_loss(self._stream(), B, self.NC, self.nout, Y.data_ptr(), lnK.data_ptr(),
                                       label.data_ptr(), mask.data_ptr(), float(w_stab), dY.data_ptr(),
                                       loss.data_ptr()), "ocpg_mlp_loss")
        gW = [None] * 4
        gb = [None] * 4
        dZ = dY
        splitK = max(1, B // 512)
        for l in range(3, -1, -1):
            n_in, n_out = self.dims[l], self.dims[l + 1]
            Hin = Hs[l]
            gW[l] = torch.empty((n_in, n_out), dtype=torch.float32, device=X.device)
            partial = torch.empty(splitK * n_in * n_out, dtype=torch.float32, device=X.device) if splitK > 1 \
                else None
            # dW = H_in^T dZ
            self._gemm(n_in, n_out, B, Hin, 1, n_in, dZ, n_out, 1, gW[l], n_out, 1, 0, splitK, partial)
            gb[l] = torch.empty(n_out, dtype=torch.float32, device=X.device)
            lib.check(self.L.ocpg_mlp_colsum(self._stream(), B, n_out, dZ.data_ptr(), gb[l].data_ptr()),
                      "ocpg_mlp_colsum")
            if l > 0:
                dH = torch.empty((B, n_in), dtype=torch.float32, device=X.device)
                # dH = dZ W^T
                self._gemm(B, n_in, n_out, dZ, n_out, 1, self.W[l], 1, n_out, dH, n_in, 1)
                lib.check(self.L.ocpg_mlp_act_backward(self._stream(), B * n_in, Zs[l - 1].data_ptr(),
                                                       dH.data_ptr()), "ocpg_mlp_act_backward")
                dZ = dH
        return loss, gW + gb

    def adam(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        for p, g, m, v in zip(self.params, grads, self.m, self.v):
            lib.check(self.L.ocpg_mlp_adam(self._stream(), p.numel(), p.data_ptr(), g.data_ptr(), m.data_ptr(),
                                           v.data_ptr(), lr, b1, b2, eps, self.step), "ocpg_mlp_adam")


def features(P, T, z):
    """(B, NC + 2) float32: P / 5000, T / 600, z."""
    z = np.asarray(z, dtype=np.float64)
    X = np.empty((z.shape[0], z.shape[1] + 2), dtype=np.float32)
    X[:, 0] = np.asarray(P) / 5000.0
    X[:, 1] = np.asarray(T) / 600.0
    X[:, 2:] = z
    return X


def flash_labels(eos, P, T, z, seed):
    """Cold FlashFIM of feeds (P, T, z) on the GPU: stability label (two
    hydrocarbon phases) and, for two-phase cells, ln K_i = ln(x_gas,i /
    x_oil,i)."""
    nb, NC = z.shape
    nc = NC + 1
    b = abi.Bulk(nb, NC, fields=["P", "T", "Ni", "S", "phaseNum", "xij", "skipFlag"])
    b.P[:] = P
    b.T[:] = T
    b.Ni[:] = workloads.moles_from_z(z, seed).ravel()
    b.S.reshape(nb, 3)[:, 2] = 0.2
    ctx = lib.Context(eos)
    ctx.set_cells(nb)
    ctx.upload(b)
    ctx.flash_fim(stats=False)
    out = abi.Bulk(nb, NC, fields=["phaseNum", "xij"])
    ctx.download(out)
    ctx.close()
    two = out.phaseNum == 3
    xij = out.xij.reshape(nb, 3, nc)
    lnK = np.zeros((nb, NC), dtype=np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        lk = np.log(xij[:, 1, :NC] / xij[:, 0, :NC])
    lnK[two] = lk[two]
    return lnK, two.astype(np.float32)


def make_dataset(eos, n, seed, p_range=(1000.0, 5000.0)):
    NC = eos.nc
    P, z, _ = workloads.cell_inputs(n, NC, seed=seed, p_range=p_range)
    T = np.full(n, workloads.T_F)
    lnK, label = flash_labels(eos, P, T, z, seed + 1)
    return features(P, T, z), lnK, label


def init_device():
    """Bring up PyTorch's device before the library's first HIP call (the
    two share the process's HIP runtime; PyTorch must see the GPU first)."""
    torch = _torch()
    torch.cuda.init()
    return torch


def train(eos, n_train=1_000_000, steps=3000, batch=8192, lr=2e-3, seed=0, log=lambda *a: None):
    torch = init_device()
    X, lnK, label = make_dataset(eos, n_train, seed)
    dev = "cuda:0"
    Xd = torch.tensor(X, device=dev)
    Kd = torch.tensor(lnK, device=dev)
    Ld = torch.tensor(label, device=dev)
    Md = Ld.clone()
    net = MLP(eos.nc, seed=seed)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    hist = []
    t0 = time.time()
    for s in range(steps):
        idx = torch.randint(0, n_train, (batch,), device=dev, generator=g)
        lr_s = lr * (0.5 * (1 + np.cos(np.pi * s / steps)))
        loss, grads = net.loss_and_grads(Xd[idx].contiguous(), Kd[idx].contiguous(), Ld[idx].contiguous(),
                                         Md[idx].contiguous())
        net.adam(grads, lr=lr_s)
        if s % 250 == 0 or s == steps - 1:
            v = float(loss.mean())
            hist.append((s, v))
            log(f"[learned] step {s} loss {v:.4f}")
    torch.cuda.synchronize()
    return net, dict(train_seconds=time.time() - t0, loss_history=hist, n_train=n_train, steps=steps,
                     batch=batch, two_phase_fraction=float(label.mean()))


def evaluate(eos, net, nb=1_122_000, seed=2024, log=lambda *a: None):
    """Cold flash of the bench field with and without the network: iteration
    totals, flash time and agreement of the results."""
    torch = init_device()
    NC = eos.nc
    P, z, Sw = workloads.cell_inputs(nb, NC, seed=seed + 1)
    T = np.full(nb, workloads.T_F)
    b = abi.Bulk(nb, NC, fields=["P", "T", "Ni", "S", "phaseNum", "skipFlag"])
    b.P[:] = P
    b.T[:] = T
    b.Ni[:] = workloads.moles_from_z(z, seed + 7).ravel()
    b.S.reshape(nb, 3)[:, 2] = Sw
    X = torch.tensor(features(P, T, z), device="cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    guess = net.forward(X)
    torch.cuda.synchronize()
    infer_ms = (time.perf_counter() - t0) * 1e3
    res = {}
    outs = {}
    for mode in ("plain", "learned"):
        ctx = lib.Context(eos)
        ctx.set_cells(nb)
        ctx.upload(b)
        ctx.save_flash_inputs()
        # one untimed pass, then a timed one from the same inputs
        for rep in range(2):
            ctx.restore_flash_inputs()
            ctx.synchronize()
            ctx.timing(rep == 1)
            t0 = time.perf_counter()
            if mode == "plain":
                st = ctx.flash_fim(stats=True)
            else:
                st = ctx.flash_fim_guided(guess, NC + 1)
            ctx.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
        stages = {k: ctx.timer(k)[0] for k in ("flash_stability", "flash_split", "flash_finish")}
        ctx.timing(False)
        o = abi.Bulk(nb, NC, fields=["phaseNum", "S", "xij"])
        ctx.download(o)
        ctx.close()
        outs[mode] = o
        res[mode] = dict(flash_ms=ms, stage_ms=stages, iters=dict(zip(["ssm_sta", "nr_sta", "ssm_sp", "nr_sp", "rr"],
                                                     [int(v) for v in st.iters])),
                         two_phase=int(st.two_phase))
        log(f"[learned] {mode}: {ms:.2f} ms, stages {stages}, iterations {res[mode]['iters']}")
    same = outs["plain"].phaseNum == outs["learned"].phaseNum
    S_p = outs["plain"].S.reshape(nb, 3)[same]
    S_l = outs["learned"].S.reshape(nb, 3)[same]
    x_p = outs["plain"].xij.reshape(nb, -1)[same]
    x_l = outs["learned"].xij.reshape(nb, -1)[same]
    res["agreement"] = dict(phase_count_same=float(same.mean()),
                            max_abs_S_diff=float(np.abs(S_p - S_l).max()),
                            max_abs_x_diff=float(np.abs(x_p - x_l).max()),
                            p99_abs_S_diff=float(np.quantile(np.abs(S_p - S_l).max(axis=1), 0.99)))
    res["inference_ms"] = infer_ms
    return res


def main():
    init_device()
    deck = fluids.bench6_deck()
    eos = lib.eos_from_deck(deck)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    net, tr = train(eos, log=log)
    ev = evaluate(eos, net, log=log)
    print(json.dumps(dict(training=tr, evaluation=ev)), flush=True)


if __name__ == "__main__":
    main()
