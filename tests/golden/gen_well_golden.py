"""Writes golden vectors of the per-step well preparation's fluid pieces from
the REFERENCE's own MixtureComp: XiPhase / RhoPhase (the injected fluid's
densities, src/MixtureComp.cpp:711-768) and Flash (the perforation stream of
Well::CalProddG01, src/MixtureComp.cpp:167-193).

Runs only where oracle/_ref/libocpref.so was compiled from the reference
sources (``make -C oracle ref``); the committed .npz is what the tests use.

    python tests/golden/gen_well_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from opencaeporo_amd import fluids  # noqa: E402
from oracle.oracle_py import Reference  # noqa: E402

SEED = 20241204


def cases(NC, rng, n):
    nc = NC + 1
    # injected fluids: a solvent (C1/C3/C6 as examples/spe5/spe5.data's
    # WELLSTRE), pure C1, random light mixtures, and water
    feeds = []
    solvent = np.zeros(nc)
    solvent[:3] = (0.77, 0.20, 0.03)
    c1 = np.zeros(nc)
    c1[0] = 1.0
    feeds += [(solvent, 1), (c1, 1)]
    for _ in range(6):
        z = np.zeros(nc)
        z[:NC] = rng.exponential(1.0, NC) * np.geomspace(1.0, 0.02, NC)
        z /= z.sum()
        feeds.append((z, 1))
    water = np.zeros(nc)
    water[-1] = 1.0
    feeds.append((water, 2))
    P = np.concatenate([[14.7, 500.0, 3000.0, 10000.0], rng.uniform(200.0, 9000.0, n - 4)])
    T = np.concatenate([[60.0, 160.0, 160.0, 200.0], rng.uniform(60.0, 250.0, n - 4)])
    return feeds, P, T


def main():
    out = {}
    rng = np.random.default_rng(SEED)
    for tag, deck in (("spe5", fluids.spe5_deck()), ("bench6", fluids.bench6_deck()), ("ten", fluids.ten_deck())):
        NC = deck.nc
        nc = NC + 1
        ref = Reference(deck, 1)
        feeds, P, T = cases(NC, rng, 24)
        Z, PH, PP, TT, XI, RHO = [], [], [], [], [], []
        for z, ph in feeds:
            for p, t in zip(P, T):
                xi, rho = ref.phase_density(p, t, z, ph)
                Z.append(z)
                PH.append(ph)
                PP.append(p)
                TT.append(t)
                XI.append(xi)
                RHO.append(rho)
        out[f"{tag}_dens_z"] = np.array(Z)
        out[f"{tag}_dens_phase"] = np.array(PH, dtype=np.int32)
        out[f"{tag}_dens_P"] = np.array(PP)
        out[f"{tag}_dens_T"] = np.array(TT)
        out[f"{tag}_dens_xi"] = np.array(XI)
        out[f"{tag}_dens_rho"] = np.array(RHO)
        # perforation streams: feeds of every kind (single-phase gas and oil,
        # two-phase), with water, over the reservoir's pressure range
        m = 160
        Ni = np.empty((m, nc))
        bias = np.ones(NC)
        bias[0] = 10.0
        Ni[:, :NC] = rng.exponential(1.0, (m, NC)) * bias[None, :] * rng.uniform(1.0, 500.0, (m, 1))
        Ni[:, NC] = rng.uniform(0.0, 300.0, m)
        Ni[::7, NC] = 0.0
        Pf = rng.uniform(500.0, 6000.0, m)
        Tf = rng.uniform(100.0, 220.0, m)
        EX, VJ, XIJ, RHOJ, NT = [], [], [], [], []
        for k in range(m):
            ex, vj, xi, rho, Nt = ref.flash_phases(Pf[k], Tf[k], Ni[k])
            EX.append(ex)
            VJ.append(vj)
            XIJ.append(xi)
            RHOJ.append(rho)
            NT.append(Nt)
        out[f"{tag}_flash_Ni"] = Ni
        out[f"{tag}_flash_P"] = Pf
        out[f"{tag}_flash_T"] = Tf
        out[f"{tag}_flash_ex"] = np.array(EX)
        out[f"{tag}_flash_vj"] = np.array(VJ)
        out[f"{tag}_flash_xi"] = np.array(XIJ)
        out[f"{tag}_flash_rho"] = np.array(RHOJ)
        out[f"{tag}_flash_Nt"] = np.array(NT)
    out["seed"] = np.int64(SEED)
    path = os.path.join(HERE, "wells_density.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
