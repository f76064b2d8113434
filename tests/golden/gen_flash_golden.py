"""Writes the flash golden vectors from the REFERENCE's own MixtureComp.

Runs only where oracle/_ref/libocpref.so was compiled from the reference
sources (``make -C oracle ref``); the committed .npz files are what the tests
use, on machines without the reference.  Inputs are seeded; each file records the seeds, the
deck, the EoS description the reference constructor derived, and for each
pass (cold: lastNP = 0 and empty skip memory; warm: the cold result with
P + U[-5, 5] psia) every array PassFlashValue hands to the Bulk.

    python tests/golden/gen_flash_golden.py [--fimn]

(--fimn writes only the IsoT_FIMn files.)
"""
import copy
import ctypes as C
import json
import os
import platform
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from opencaeporo_amd import abi, fluids, workloads  # noqa: E402
from oracle.oracle_py import Reference  # noqa: E402

OUT_FIELDS = ["phaseNum", "phaseExist", "S", "xij", "rho", "xi", "mu", "rhoP", "xiP", "muP",
              "rhox", "xix", "mux", "Nt", "vf", "vfP", "vfi", "pSderExist", "pVnumCom",
              "bRowSizedSdP", "dSec_dPri", "skipFlag", "skipMinEig", "skipP", "skipT",
              "skipZi", "ftype", "iters"]
IN_FIELDS = ["P", "T", "Ni", "S", "phaseNum", "xij", "skipFlag", "skipMinEig", "skipP",
             "skipT", "skipZi"]


def deck_arrays(deck):
    NC = deck.nc
    return dict(
        deck_nc=np.int32(NC), deck_Tc=np.array(deck.Tc[:NC]), deck_Pc=np.array(deck.Pc[:NC]),
        deck_MW=np.array(deck.MW[:NC]), deck_Acf=np.array(deck.Acf[:NC]),
        deck_Zc=np.array(deck.Zc[:NC]),
        deck_BIC_lower=np.array(deck.BIC_lower[:NC * (NC - 1) // 2]),
        deck_LBCcoef=np.array(deck.LBCcoef[:]),
        deck_pvtw=np.array([list(deck.pvtw[r]) for r in range(deck.pvtw_rows)]),
        deck_std_rhow=np.float64(deck.std_rhow),
        deck_controls=np.array([deck.ssmsta_maxit, deck.ssmsta_tol, deck.ssmsta_eyt,
                                deck.nrsta_maxit, deck.nrsta_tol, deck.ssmsp_maxit,
                                deck.ssmsp_tol, deck.nrsp_maxit, deck.nrsp_tol, deck.rr_maxit,
                                deck.rr_tol]),
    )


def eos_arrays(eos):
    NC = eos.nc
    bic = np.array(eos.BIC[:]).reshape(abi.MAXNC, abi.MAXNC)[:NC, :NC]
    return dict(eos_Vc=np.array(eos.Vc[:NC]), eos_OmegaA=np.array(eos.OmegaA[:NC]),
                eos_OmegaB=np.array(eos.OmegaB[:NC]), eos_Vshift=np.array(eos.Vshift[:NC]),
                eos_Vcvis=np.array(eos.Vcvis[:NC]), eos_BIC=bic,
                eos_LBCcoef=np.array(eos.LBCcoef[:]))


def used_dsec(b):
    L = abi.max_len_dsdp(b.nc)
    used = int(b.bRowSizedSdP.max()) * (b.nc + 1)
    return b.dSec_dPri.reshape(b.nb, L)[:, :used].copy()


def snapshot(b, fields, prefix):
    out = {}
    for f in fields:
        a = used_dsec(b) if f == "dSec_dPri" else getattr(b, f).copy()
        out[prefix + f] = a
    return out


FIMN_FIELDS = ["nj", "res_n", "resPc"]


def gen_flash(name, deck, nb, seeds, dp_seed, limits=None, method="fim", T=None, bias=None, p_range=None):
    """limits: (SSMSTA, NRSTA, SSMSP, NRSP) iteration caps replacing the
    deck's (the non-convergence paths); such files also record the per-cell
    non-convergence flag.  method "fimn": the IsoT_FIMn flash (FlashFIMn),
    whose files also record nj, res_n and resPc.  T (F), bias (the feed the
    random compositions scatter around) and p_range (psia) default to the
    workloads' values."""
    out_fields = OUT_FIELDS + (FIMN_FIELDS if method == "fimn" else [])
    flash = "flash_fimn" if method == "fimn" else "flash_fim"
    if limits is not None:
        deck = copy.copy(deck)
        deck.ssmsta_maxit, deck.nrsta_maxit, deck.ssmsp_maxit, deck.nrsp_maxit = limits
        out_fields = out_fields + ["nonconv"]
    NC = deck.nc
    ref = Reference(deck, nb)
    kw = dict(bias=bias) if bias is not None else {}
    if p_range is not None:
        kw["p_range"] = p_range
    P, z, Sw = workloads.cell_inputs(nb, NC, seed=seeds[0], **kw)
    Ni = workloads.moles_from_z(z, seed=seeds[1])
    b = abi.Bulk(nb, NC)
    b.P[:] = P
    b.T[:] = workloads.T_F if T is None else T
    b.Ni[:] = Ni.ravel()
    data = dict(seeds=np.array(list(seeds) + [dp_seed]))
    data.update(deck_arrays(deck))
    data.update(eos_arrays(ref.eos()))
    data.update(snapshot(b, IN_FIELDS, "cold_in_"))
    getattr(ref, flash)(b)
    data.update(snapshot(b, out_fields, "cold_out_"))
    rng = np.random.default_rng(dp_seed)
    b.P[:] = P + rng.uniform(-5.0, 5.0, nb)
    data.update(snapshot(b, IN_FIELDS, "warm_in_"))
    getattr(ref, flash)(b)
    data.update(snapshot(b, out_fields, "warm_out_"))
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **data)
    two = np.mean(b.phaseNum == 3)
    print(f"{name}: {nb} cells NC={NC}, two-phase {two:.2f}, {os.path.getsize(path)} bytes")


def gen_init(name, deck, nb, seeds, method="fim", T=None, bias=None, p_range=None):
    NC = deck.nc
    ref = Reference(deck, nb)
    kw = dict(bias=bias) if bias is not None else {}
    if p_range is not None:
        kw["p_range"] = p_range
    P, z, Sw = workloads.cell_inputs(nb, NC, seed=seeds[0], **kw)
    rng = np.random.default_rng(seeds[1])
    b = abi.Bulk(nb, NC)
    b.P[:] = P
    b.T[:] = workloads.T_F if T is None else T
    b.S.reshape(nb, 3)[:, 2] = Sw
    b.rockVp[:] = rng.uniform(1.0e4, 5.0e4, nb)
    data = dict(seeds=np.array(seeds), init_z=z.copy(), init_rockVp=b.rockVp.copy())
    data.update(deck_arrays(deck))
    data.update(snapshot(b, ["P", "T", "S"], "init_in_"))
    if method == "fimn":
        ref.init_flash_fimn(b, z)
        data.update(snapshot(b, OUT_FIELDS + FIMN_FIELDS + ["Ni"], "init_out_"))
    else:
        ref.init_flash_fim(b, z)
        data.update(snapshot(b, OUT_FIELDS + ["Ni"], "init_out_"))
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **data)
    print(f"{name}: {nb} cells NC={NC}, {os.path.getsize(path)} bytes")


def mkl_version():
    try:
        meta = "/opt/conda/conda-meta"
        return sorted(f for f in os.listdir(meta) if f.startswith("mkl-2"))[0][:-5]
    except OSError:
        return "unknown"


def gen_fimn():
    """The IsoT_FIMn files (FlashFIMn / InitFlashFIMn)."""
    gen_flash("flashn_spe5.npz", fluids.spe5_deck(), 256, (111, 112), 113, method="fimn")
    gen_flash("flashn_ten.npz", fluids.ten_deck(), 128, (311, 312), 313, method="fimn")
    gen_init("initflashn_spe5.npz", fluids.spe5_deck(), 128, (411, 412), method="fimn")
    # capped iterations: unconverged splits, whose correction res_n is large
    gen_flash("flashn_limits_bench6.npz", fluids.bench6_deck(), 256, (511, 512), 513, limits=(10, 2, 8, 2),
              method="fimn")


def gen_decks():
    """The fluids of the reference's other decks: 9 components (SPE3, at its
    200 F) and 7 (the CO2 + SPE5 fluid of spe5large, at 160 F), compositions
    scattered around each deck's feed."""
    s3, s7 = fluids.SPE3, fluids.SPE5L
    gen_flash("flash_spe3.npz", fluids.spe3_deck(), 192, (701, 702), 703, T=s3["T_F"], bias=s3["feed"])
    gen_flash("flash_spe5large.npz", fluids.spe5large_deck(), 192, (711, 712), 713, T=s7["T_F"],
              bias=s7["feed"], p_range=(500.0, 4000.0))
    gen_flash("flashn_spe3.npz", fluids.spe3_deck(), 128, (721, 722), 723, method="fimn", T=s3["T_F"],
              bias=s3["feed"])
    gen_flash("flashn_spe5large.npz", fluids.spe5large_deck(), 128, (731, 732), 733, method="fimn",
              T=s7["T_F"], bias=s7["feed"], p_range=(500.0, 4000.0))
    gen_init("initflash_spe5large.npz", fluids.spe5large_deck(), 128, (741, 742), T=s7["T_F"], bias=s7["feed"],
             p_range=(500.0, 4000.0))


def main():
    if not Reference.available():
        sys.exit("oracle/_ref/libocpref.so is not built (make -C oracle ref)")
    if "--fimn" in sys.argv:
        gen_fimn()
        return
    if "--decks" in sys.argv:
        gen_decks()
        return
    gen_decks()
    gen_fimn()
    gen_flash("flash_spe5.npz", fluids.spe5_deck(), 384, (101, 102), 103)
    gen_flash("flash_bench6.npz", fluids.bench6_deck(), 256, (201, 202), 203)
    gen_flash("flash_ten.npz", fluids.ten_deck(), 192, (301, 302), 303)
    gen_init("initflash_spe5.npz", fluids.spe5_deck(), 128, (401, 402))
    # iteration caps: some cells stop unconverged, and with every cap at 0
    # every cell does (MixtureComp's ePEC = 1e8 paths)
    gen_flash("flash_limits_bench6.npz", fluids.bench6_deck(), 256, (501, 502), 503, limits=(10, 2, 8, 2))
    gen_flash("flash_maxit0_bench6.npz", fluids.bench6_deck(), 128, (601, 602), 603, limits=(0, 0, 0, 0))
    meta = dict(reference="OpenCAEPoro 0.5.0 MixtureComp (src/MixtureComp.cpp) compiled by "
                          "oracle/Makefile 'ref'",
                lapack=mkl_version() + " (sequential layer)",
                compiler="g++ -O3 -std=c++14", host=platform.processor() or platform.machine(),
                ten_component_bic_seed=20241203)
    with open(os.path.join(HERE, "flash_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
