"""Builds libocpgpu.so (HIP, gfx950) in-tree.

    python -m opencaeporo_amd.build [--incremental] [--nc 6,10]

Each flash instantiation (one per supported hydrocarbon component count) is
its own translation unit so the instantiations compile in parallel.  The
library is written next to this file; the tests, the smoke check and the
benchmark load it from there.
"""
import argparse
import concurrent.futures as cf
import hashlib
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libocpgpu.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

DEFAULT_NC = (6, 7, 9, 10)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# No FP contraction: the kernels keep the reference's rounding order.
# Packed FP32 instructions are disabled (wrong results have been seen from
# some of their forms on this hardware).
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", CSRC, "-I", INCLUDE,
         "-Wno-pass-failed"]


def _headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))] + \
        [os.path.join(INCLUDE, "ocpgpu.h")]


# Optimisation level per component count of the flash translation units
# (the default -O3 for all).  Earlier versions of the 10-component kernels
# gave wrong phase splits on the GPU at -O3 and -O2 (other iteration paths
# in most two-phase cells) while -O1, a CPU build of the same source and the
# oracle agreed; the current source passes every 10-component parity test at
# -O3 (a few cells with another iteration count, none with another phase
# count), and those tests guard it.  OCPG_NC10_OPT=-O1 rebuilds it the old
# way.
NC_OPT = {10: os.environ["OCPG_NC10_OPT"]} if os.environ.get("OCPG_NC10_OPT") else {}


def _jobs(ncs):
    nc_list = " ".join(f"X({n})" for n in ncs)
    jobs = []
    for n in ncs:
        opt = [NC_OPT[n]] if n in NC_OPT else []
        jobs.append(("flash_nc.hip", f"flash_nc{n}.o", [f"-DOCPG_NC={n}"] + opt))
        jobs.append(("flashn_nc.hip", f"flashn_nc{n}.o", [f"-DOCPG_NC={n}"] + opt))
        jobs.append(("well_nc.hip", f"well_nc{n}.o", [f"-DOCPG_NC={n}"] + opt))
    jobs.append(("flash_dispatch.cpp", "flash_dispatch.o", [f"-DOCPG_NC_LIST(X)={nc_list}"]))
    jobs.append(("fim_kernels.hip", "fim_kernels.o", []))
    jobs.append(("ocpgpu_capi.cpp", "ocpgpu_capi.o", []))
    jobs.append(("fim_capi.cpp", "fim_capi.o", []))
    jobs.append(("mlp_kernels.hip", "mlp_kernels.o", []))
    return jobs


def _deps(depfile):
    """The files a make-style dependency file lists (None if it is missing)."""
    if not os.path.exists(depfile):
        return None
    text = open(depfile).read().replace("\\\n", " ")
    files = []
    for line in text.splitlines():
        if ":" in line:
            files += line.split(":", 1)[1].split()
    return sorted(set(f for f in files if f.startswith(os.path.dirname(HERE))))


def _digest(cmd, files):
    """Hash of the command and of the texts of the source and the project
    headers it includes, read before compiling (an edit made while the
    compiler runs forces the next build to compile again)."""
    h = hashlib.sha256(" ".join(cmd).encode())
    for f in files:
        h.update(f.encode())
        try:
            with open(f, "rb") as fh:
                h.update(fh.read())
        except OSError:
            h.update(b"missing")
    return h.hexdigest()


def _compile(src, obj, extra, incremental, stamp):
    srcp = os.path.join(CSRC, src)
    objp = os.path.join(BUILD, obj)
    depfile = objp + ".d"
    cmd = [HIPCC] + FLAGS + extra + ["-c", srcp, "-o", objp]
    deps = _deps(depfile)
    if incremental and deps is not None and os.path.exists(objp) and os.path.exists(stamp):
        if open(stamp).read() == _digest(cmd, deps):
            return obj, 0.0, ""
    # the texts this compile reads, hashed before it starts
    pre = {f: open(f, "rb").read() for f in [srcp] + _headers()}
    t0 = time.time()
    r = subprocess.run(cmd + ["-MD", "-MF", depfile], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    deps = _deps(depfile) or [srcp]
    h = hashlib.sha256(" ".join(cmd).encode())
    for f in deps:
        h.update(f.encode())
        h.update(pre[f] if f in pre else open(f, "rb").read())
    with open(stamp, "w") as f:
        f.write(h.hexdigest())
    return obj, time.time() - t0, r.stderr


def build(ncs=DEFAULT_NC, incremental=False, verbose=True, workers=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = _jobs(ncs)
    workers = workers or min(len(jobs), os.cpu_count() or 4, 16)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        futs = [ex.submit(_compile, s, o, e, incremental, os.path.join(BUILD, o + ".flags"))
                for s, o, e in jobs]
        for f in cf.as_completed(futs):
            obj, dt, _ = f.result()
            objs.append(obj)
            if verbose and dt > 0:
                print(f"  built {obj} in {dt:.1f}s", flush=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libocpgpu.so", "-Wl,--no-undefined",
           "-o", tmp] + \
        [os.path.join(BUILD, o) for _, o, _ in jobs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"  linked {LIB}", flush=True)
    _build_shim_driver(verbose)
    return LIB


SHIM_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "shim_driver.cpp")
SHIM_LIB = os.path.join(os.path.dirname(HERE), "tests", "cpp", "libshim_driver.so")


def _build_shim_driver(verbose):
    """The C++ test driver of the IsoT_FIM-shaped host shim (include/ocpgpu_fim.hpp),
    linked against libocpgpu.so."""
    if not os.path.exists(SHIM_SRC):
        return
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", INCLUDE, SHIM_SRC, "-L", HERE, "-locpgpu",
           "-Wl,-rpath,$ORIGIN/../../opencaeporo_amd", "-o", SHIM_LIB]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"shim driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"  linked {SHIM_LIB}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--nc", default=",".join(str(n) for n in DEFAULT_NC))
    a = ap.parse_args()
    build(tuple(int(x) for x in a.nc.split(",")), incremental=a.incremental)


if __name__ == "__main__":
    sys.exit(main())
