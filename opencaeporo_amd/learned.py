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

    def infer(self, X, Y=None):
        """Y = the network on X (B, nin) in one fused kernel (ocpg_mlp_forward)."""
        torch = _torch()
        B = X.shape[0]
        if Y is None:
            Y = torch.empty((B, self.nout), dtype=torch.float32, device=X.device)
        args = []
        for l in range(4):
            args += [self.W[l].data_ptr(), self.b[l].data_ptr()]
        lib.check(self.L.ocpg_mlp_forward(self._stream(), B, self.nin, self.nout, X.data_ptr(), X.shape[1], *args,
                                          Y.data_ptr(), Y.shape[1]), "ocpg_mlp_forward")
        return Y

    def save(self, path):
        np.savez(path, **{f"W{l}": self.W[l].cpu().numpy() for l in range(4)},
                 **{f"b{l}": self.b[l].cpu().numpy() for l in range(4)}, NC=self.NC)

    @classmethod
    def load(cls, path, device="cuda:0"):
        torch = _torch()
        d = np.load(path, allow_pickle=False)
        net = cls(int(d["NC"]), device=device)
        for l in range(4):
            net.W[l].copy_(torch.tensor(d[f"W{l}"], device=device))
            net.b[l].copy_(torch.tensor(d[f"b{l}"], device=device))
        return net

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


def workload_dataset(eos, n, seed, chunk=2_000_000, log=lambda *a: None):
    """About n flash results of the configurations' own feeds, made by the
    GPU's cold flash in chunks: 80% the standalone field's (P uniform in
    [1000, 5000] psia, z biased towards two phases as the 1.122M-cell field
    draws it), 20% the 147-cell FIM case's (z ~ Dirichlet(1) and the
    pressures its wells span, 1000..8000 psia).  Seeds are disjoint from the
    evaluation field's."""
    NC = eos.nc
    rng = np.random.default_rng(seed)
    parts = []
    made = 0
    k = 0
    while made < n:
        m = min(chunk, n - made)
        sd = int(rng.integers(1 << 30)) + 100_000
        if k % 5 == 4:
            P = np.random.default_rng(sd).uniform(1000.0, 8000.0, m)
            z = np.random.default_rng(sd + 1).dirichlet(np.ones(NC), m)
        else:
            P, z, _ = workloads.cell_inputs(m, NC, seed=sd)
        T = np.full(m, workloads.T_F)
        lnK, label = flash_labels(eos, P, T, z, sd + 2)
        parts.append((features(P, T, z), lnK, label))
        made += m
        k += 1
        log(f"[learned] dataset {made}/{n} ({float(label.mean()):.2f} two-phase)")
    return tuple(np.concatenate([p[q] for p in parts]) for q in range(3))


def init_device():
    """Bring up PyTorch's device before the library's first HIP call (the
    two share the process's HIP runtime; PyTorch must see the GPU first)."""
    torch = _torch()
    torch.cuda.init()
    return torch


def train(eos, n_train=1_000_000, steps=3000, batch=8192, lr=2e-3, seed=0, log=lambda *a: None, data=None):
    """data: (X, lnK, label) host arrays; made by make_dataset when None."""
    torch = init_device()
    X, lnK, label = data if data is not None else make_dataset(eos, n_train, seed)
    n_train = X.shape[0]
    dev = "cuda:0"
    Xd = torch.tensor(X, device=dev)
    Kd = torch.tensor(lnK, device=dev)
    Ld = torch.tensor(label, device=dev)
    Md = Ld.clone()
    net = MLP(eos.nc, seed=seed)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    hist = []
    torch.cuda.synchronize()
    t0 = time.time()
    for s in range(steps):
        idx = torch.randint(0, n_train, (batch,), device=dev, generator=g)
        lr_s = lr * (0.5 * (1 + np.cos(np.pi * s / steps)))
        loss, grads = net.loss_and_grads(Xd[idx].contiguous(), Kd[idx].contiguous(), Ld[idx].contiguous(),
                                         Md[idx].contiguous())
        net.adam(grads, lr=lr_s)
        if s % 500 == 0 or s == steps - 1:
            v = float(loss.mean())
            hist.append((s, v))
            log(f"[learned] step {s} loss {v:.4f}")
    torch.cuda.synchronize()
    return net, dict(train_seconds=time.time() - t0, loss_history=hist, n_train=n_train, steps=steps,
                     batch=batch, two_phase_fraction=float(label.mean()))


def _field(eos, nb, seed):
    """The standalone field's cold state as the benchmark builds it: P, z and
    Sw drawn as there, the moles from InitFlashFIM (pore volume of a
    100 x 100 x 10 ft cell, porosity 0.3), no previous phase information."""
    NC = eos.nc
    P, z, Sw = workloads.cell_inputs(nb, NC, seed=seed + 1)
    b = abi.Bulk(nb, NC, fields=["P", "T", "S", "v", "depth", "poroInit"])
    b.P[:] = P
    b.T[:] = workloads.T_F
    b.S.reshape(nb, 3)[:, 2] = Sw
    b.v[:] = 100.0 * 100.0 * 10.0
    b.depth[:] = 8325.0
    b.poroInit[:] = 0.3
    ctx = lib.Context(eos, fluids.make_satfunc(), fluids.make_rock())
    ctx.set_cells(nb)
    ctx.upload(b)
    ctx.calrock()
    ctx.init_flash_fim(z)
    out = abi.Bulk(nb, NC, fields=["P", "T", "Ni", "S", "phaseNum", "skipFlag"])
    ctx.download(out)
    ctx.close()
    out.phaseNum[:] = 0
    out.skipFlag[:] = 0
    return out


def _weights(net):
    return [w.cpu().numpy() for w in net.W], [x.cpu().numpy() for x in net.b]


def _run(eos, state, weights, NC, reps=3):
    """Flash of `state` from the same inputs, plain (weights None) or with
    the learned guess (ocpg_flash_fim_learned): the median of reps timed
    passes after one untimed; stats, inference time and outputs of the
    last."""
    nb = state.nb
    ctx = lib.Context(eos)
    ctx.set_cells(nb)
    ctx.upload(state)
    ctx.save_flash_inputs()
    if weights is not None:
        ctx.set_learned(*weights)
    times = []
    for rep in range(reps + 1):
        ctx.restore_flash_inputs()
        ctx.synchronize()
        ctx.timing(rep == reps)
        t0 = time.perf_counter()
        st = ctx.flash_fim(stats=True) if weights is None else ctx.flash_fim_learned(stats=True)
        ctx.synchronize()
        if rep > 0:
            times.append((time.perf_counter() - t0) * 1e3)
    stages = {k: ctx.timer(k)[0] for k in ("flash_stability", "flash_split", "flash_finish", "mlp")}
    ctx.timing(False)
    o = abi.Bulk(nb, NC, fields=["phaseNum", "S", "xij", "skipFlag", "skipMinEig", "skipP", "skipT", "skipZi"])
    ctx.download(o)
    ctx.close()
    res = dict(flash_ms=float(np.median(times)), stage_ms=stages,
               iters=dict(zip(["ssm_sta", "nr_sta", "ssm_sp", "nr_sp", "rr"], [int(v) for v in st.iters])),
               two_phase=int(st.two_phase), nonconverged=int(st.nonconverged))
    return res, o


def _agreement(a, b, nb):
    same = a.phaseNum == b.phaseNum
    S_a, S_b = a.S.reshape(nb, 3)[same], b.S.reshape(nb, 3)[same]
    x_a, x_b = a.xij.reshape(nb, -1)[same], b.xij.reshape(nb, -1)[same]
    dS = np.abs(S_a - S_b).max(axis=1) if same.any() else np.zeros(1)
    return dict(phase_count_same=float(same.mean()), phase_count_differs=int((~same).sum()),
                max_abs_S_diff=float(dS.max()), max_abs_x_diff=float(np.abs(x_a - x_b).max()) if same.any() else 0.0,
                p99_abs_S_diff=float(np.quantile(dS, 0.99)))


def evaluate(eos, net, nb=1_122_000, seed=2024, log=lambda *a: None, warm=True):
    """The flash of the standalone field with and without the network
    (ocpg_flash_fim_learned: the fused inference from the cells' own state,
    then the guided flash, timed together): cold (no previous phase
    information) and, with warm, the warm Newton pass (the plain cold pass's
    phase state, pressures moved by U[-5, 5] psia).  Flash time, stage times
    (with the inference), iteration totals and the agreement of the
    converged results."""
    init_device()
    NC = eos.nc
    b = _field(eos, nb, seed)
    wts = _weights(net)
    out = {}
    res, outs = {}, {}
    for mode in ("plain", "learned"):
        res[mode], outs[mode] = _run(eos, b, None if mode == "plain" else wts, NC)
        log(f"[learned] cold {mode}: {res[mode]['flash_ms']:.2f} ms, stages {res[mode]['stage_ms']}, "
            f"iterations {res[mode]['iters']}")
    res["agreement"] = _agreement(outs["plain"], outs["learned"], nb)
    res["gain"] = res["plain"]["flash_ms"] / res["learned"]["flash_ms"]
    out["cold"] = res
    out["inference_ms"] = res["learned"]["stage_ms"]["mlp"]
    if warm:
        # the warm Newton pass: the plain cold pass's phase state, P moved
        kept = ["S", "phaseNum", "xij", "skipFlag", "skipMinEig", "skipP", "skipT", "skipZi"]
        w2 = abi.Bulk(nb, NC, fields=["P", "T", "Ni"] + kept)
        for f in ("P", "T", "Ni"):
            w2.arrays[f][:] = b.arrays[f]
        for f in kept:
            w2.arrays[f][:] = outs["plain"].arrays[f]
        w2.P[:] += np.random.default_rng(seed + 3).uniform(-5.0, 5.0, nb)
        rw, ow = {}, {}
        for mode in ("plain", "learned"):
            rw[mode], ow[mode] = _run(eos, w2, None if mode == "plain" else wts, NC)
            log(f"[learned] warm {mode}: {rw[mode]['flash_ms']:.2f} ms, stages {rw[mode]['stage_ms']}, "
                f"iterations {rw[mode]['iters']}")
        rw["agreement"] = _agreement(ow["plain"], ow["learned"], nb)
        rw["gain"] = rw["plain"]["flash_ms"] / rw["learned"]["flash_ms"]
        out["warm"] = rw
    return out


WEIGHTS = {6: "weights/mlp6.npz", 10: "weights/mlp10.npz"}


def weights_path(NC):
    import os
    return os.path.join(os.path.dirname(__file__), WEIGHTS[NC])


def main():
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=10_000_000)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--out", default=None, help="save the trained weights (.npz)")
    ap.add_argument("--load", default=None, help="evaluate saved weights instead of training")
    ap.add_argument("--nc", type=int, default=6, choices=(6, 10))
    a = ap.parse_args()
    init_device()
    deck = fluids.bench6_deck() if a.nc == 6 else fluids.ten_deck()
    eos = lib.eos_from_deck(deck)
    log = lambda *x: print(*x, file=sys.stderr, flush=True)
    out = {}
    if a.load:
        net = MLP.load(a.load)
    else:
        t0 = time.time()
        data = workload_dataset(eos, a.samples, seed=12345, log=log)
        gen_s = time.time() - t0
        net, tr = train(eos, steps=a.steps, batch=a.batch, lr=2e-3, seed=0, log=log, data=data)
        tr["dataset_seconds"] = gen_s
        out["training"] = tr
        if a.out:
            net.save(a.out)
    out["evaluation"] = evaluate(eos, net, log=log)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
