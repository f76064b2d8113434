"""Loader and thin Python mirror of the C-ABI (include/ocpgpu.h).

Loading fails loudly when libocpgpu.so is not built: there is no CPU
fallback in the product path.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("OCPG_LIBRARY") or os.path.join(HERE, "libocpgpu.so")

_lib = None


class OcpgError(RuntimeError):
    pass


def load():
    """The HIP library, with argument types set."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OcpgError(f"{LIB_PATH} is not built; run `python -m opencaeporo_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "ocpg_version": ([], C.c_char_p),
        "ocpg_sizeof": ([C.c_char_p], C.c_int64),
        "ocpg_eos_from_deck": ([P, P], C.c_int),
        "ocpg_create": ([P, P, P, P, C.c_int], C.c_int),
        "ocpg_destroy": ([P], C.c_int),
        "ocpg_set_cells": ([P, C.c_int64], C.c_int),
        "ocpg_set_method": ([P, C.c_int], C.c_int),
        "ocpg_get_method": ([P], C.c_int),
        "ocpg_critical_saturations": ([P, P], C.c_int),
        "ocpg_upload_bulk": ([P, P], C.c_int),
        "ocpg_download_bulk": ([P, P], C.c_int),
        "ocpg_flash_fim": ([P, P], C.c_int),
        "ocpg_init_flash_fim": ([P, P], C.c_int),
        "ocpg_update_last_step": ([P], C.c_int),
        "ocpg_reset_to_last_step": ([P], C.c_int),
        "ocpg_set_conn": ([P, P], C.c_int),
        "ocpg_bsr_nnzb": ([P], C.c_int64),
        "ocpg_bsr_rows": ([P], C.c_int64),
        "ocpg_capture_iteration": ([P, C.c_double], C.c_int),
        "ocpg_flash_fim_guided": ([P, P, C.c_int, P], C.c_int),
        "ocpg_set_learned": ([P, C.c_int, C.c_int, P, P, P, P, P, P, P, P], C.c_int),
        "ocpg_flash_fim_learned": ([P, P], C.c_int),
        "ocpg_mlp_gemm": ([P, C.c_int, C.c_int, C.c_int, P, C.c_int64, C.c_int64, P, C.c_int64, C.c_int64,
                           P, C.c_int64, C.c_int64, C.c_int, C.c_int, P], C.c_int),
        "ocpg_mlp_gemm_noalloc": ([P, C.c_int, C.c_int, C.c_int, P, C.c_int64, C.c_int64, P, C.c_int64,
                                   C.c_int64, P, C.c_int64, C.c_int64, C.c_int, C.c_int], C.c_int),
        "ocpg_mlp_bias_act": ([P, C.c_int, C.c_int, P, P, P], C.c_int),
        "ocpg_mlp_forward": ([P, C.c_int, C.c_int, C.c_int, P, C.c_int, P, P, P, P, P, P, P, P, P, C.c_int],
                             C.c_int),
        "ocpg_mlp_act_backward": ([P, C.c_int64, P, P], C.c_int),
        "ocpg_mlp_colsum": ([P, C.c_int, C.c_int, P, P], C.c_int),
        "ocpg_mlp_loss": ([P, C.c_int, C.c_int, C.c_int, P, P, P, P, C.c_float, P, P], C.c_int),
        "ocpg_mlp_adam": ([P, C.c_int64, P, P, P, P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int],
                          C.c_int),
        "ocpg_launch_iteration": ([P], C.c_int),
        "ocpg_iteration_done": ([P], C.c_int),
        "ocpg_set_wells": ([P, P], C.c_int),
        "ocpg_set_well_state": ([P, P], C.c_int),
        "ocpg_set_well_setup": ([P, P], C.c_int),
        "ocpg_prepare_wells": ([P], C.c_int),
        "ocpg_check_wells": ([P, P], C.c_int),
        "ocpg_get_well_state": ([P, P, P, P, P, P], C.c_int),
        "ocpg_phase_density": ([P, C.c_double, C.c_double, P, C.c_int, P, P], C.c_int),
        "ocpg_check_bulk": ([P, P], C.c_int),
        "ocpg_max_change": ([P, P], C.c_int),
        "ocpg_get_well_flux": ([P, P, P, P], C.c_int),
        "ocpg_calkrpc": ([P], C.c_int),
        "ocpg_calrock": ([P], C.c_int),
        "ocpg_calkrpc_rock": ([P], C.c_int),
        "ocpg_calres": ([P, C.c_double, P], C.c_int),
        "ocpg_assemble": ([P, C.c_double], C.c_int),
        "ocpg_update_assemble": ([P, C.c_double, P, P], C.c_int),
        "ocpg_get_solution": ([P, P, C.c_int, C.c_double, P], C.c_int),
        "ocpg_nr_change": ([P, P], C.c_int),
        "ocpg_get_well_bhp": ([P, P], C.c_int),
        "ocpg_get_bsr": ([P, P, P, P, P], C.c_int),
        "ocpg_get_bsr_async": ([P, P, P, P, P], C.c_int),
        "ocpg_host_alloc": ([C.c_int64], C.c_void_p),
        "ocpg_host_free": ([P], C.c_int),
        "ocpg_get_upwind": ([P, P, P, P], C.c_int),
        "ocpg_pool_init": ([P, C.c_int64, C.c_int32], C.c_int),
        "ocpg_pool_place": ([P, C.c_int64, P, P], C.c_int),
        "ocpg_pool_activate": ([P, C.c_int64, C.c_int], C.c_int),
        "ocpg_pool_system": ([P, C.c_int64, P, P, P, P, P], C.c_int),
        "ocpg_pool_download": ([P, C.c_int64, P, C.c_int], C.c_int),
        "ocpg_pool_remove": ([P, C.c_int64], C.c_int),
        "ocpg_pool_image_bytes": ([P, C.c_int64], C.c_int64),
        "ocpg_pool_export": ([P, C.c_int64, P, C.c_int], C.c_int),
        "ocpg_pool_import": ([P, C.c_int64, P], C.c_int),
        "ocpg_pool_sync_copies": ([P], C.c_int),
        "ocpg_save_flash_inputs": ([P], C.c_int),
        "ocpg_restore_flash_inputs": ([P], C.c_int),
        "ocpg_timing_enable": ([P, C.c_int], C.c_int),
        "ocpg_timing_get": ([P, C.c_char_p, P, P], C.c_int),
        "ocpg_synchronize": ([P], C.c_int),
        "ocpg_stream": ([P], C.c_void_p),
        "ocpg_device_bytes": ([P], C.c_int64),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc, what):
    if rc != abi.OK:
        raise OcpgError(f"{what} failed with code {rc}")


def eos_from_deck(deck):
    e = abi.Eos()
    check(load().ocpg_eos_from_deck(C.byref(deck), C.byref(e)), "ocpg_eos_from_deck")
    return e


class Context:
    """One ocpg_ctx: EoS + saturation tables + rock on one device."""

    def __init__(self, eos, sat=None, rock=None, device=0):
        self.lib = load()
        self.eos = eos
        self.NC = eos.nc
        self.nc = self.NC + 1
        self.ncol = self.nc + 1
        h = C.c_void_p()
        check(self.lib.ocpg_create(C.byref(h), C.byref(eos), C.byref(sat) if sat else None,
                                   C.byref(rock) if rock else None, device), "ocpg_create")
        self.h = h
        self.nb = 0
        self.grid = None

    def close(self):
        if self.h:
            self.lib.ocpg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_method(self, method):
        """abi.METHOD_FIM or abi.METHOD_FIMN (before set_cells)."""
        check(self.lib.ocpg_set_method(self.h, method), "ocpg_set_method")

    def critical_saturations(self):
        out = np.zeros(3)
        check(self.lib.ocpg_critical_saturations(self.h, out.ctypes.data), "ocpg_critical_saturations")
        return out

    def set_cells(self, nb):
        check(self.lib.ocpg_set_cells(self.h, nb), "ocpg_set_cells")
        self.nb = nb

    def upload(self, bulk):
        s = bulk.cstruct()
        check(self.lib.ocpg_upload_bulk(self.h, C.byref(s)), "ocpg_upload_bulk")

    def download(self, bulk):
        s = bulk.cstruct()
        check(self.lib.ocpg_download_bulk(self.h, C.byref(s)), "ocpg_download_bulk")

    def flash_fim(self, stats=True):
        st = abi.FlashStats() if stats else None
        check(self.lib.ocpg_flash_fim(self.h, C.byref(st) if st else None), "ocpg_flash_fim")
        return st

    def init_flash_fim(self, zi):
        zi = np.ascontiguousarray(zi, dtype=np.float64)
        check(self.lib.ocpg_init_flash_fim(self.h, zi.ctypes.data), "ocpg_init_flash_fim")

    def update_last_step(self):
        check(self.lib.ocpg_update_last_step(self.h), "ocpg_update_last_step")

    def reset_to_last_step(self):
        check(self.lib.ocpg_reset_to_last_step(self.h), "ocpg_reset_to_last_step")

    def set_conn(self, grid):
        conn = abi.conn_struct(grid.bId, grid.eId, grid.area)
        check(self.lib.ocpg_set_conn(self.h, C.byref(conn)), "ocpg_set_conn")
        self.grid = grid

    def nnzb(self):
        return int(self.lib.ocpg_bsr_nnzb(self.h))

    def calkrpc(self):
        check(self.lib.ocpg_calkrpc(self.h), "ocpg_calkrpc")

    def calrock(self):
        check(self.lib.ocpg_calrock(self.h), "ocpg_calrock")

    def calkrpc_rock(self):
        check(self.lib.ocpg_calkrpc_rock(self.h), "ocpg_calkrpc_rock")

    def calres(self, dt, norms=True):
        nr = abi.ResNorm() if norms else None
        check(self.lib.ocpg_calres(self.h, dt, C.byref(nr) if nr else None), "ocpg_calres")
        return nr

    def assemble(self, dt):
        check(self.lib.ocpg_assemble(self.h, dt), "ocpg_assemble")

    def update_assemble(self, dt, norms=False, stats=False):
        """flash_fim + calkrpc_rock + calres + assemble in one call, the
        flash's slowest cells overlapped with the rest (ocpg_update_assemble).
        Returns (norms or None, stats or None)."""
        nr = abi.ResNorm() if norms else None
        st = abi.FlashStats() if stats else None
        check(self.lib.ocpg_update_assemble(self.h, dt, C.byref(nr) if nr else None,
                                            C.byref(st) if st else None), "ocpg_update_assemble")
        return nr, st

    def get_solution(self, u, dSmaxlim, device_ptr=None):
        """IsoT_FIM::GetSolution on the device from the solution u (host
        array of rows * ncol, or device_ptr: an address in device memory).
        Returns (NRdPmax, NRdNmax)."""
        out = abi.NRUpdate()
        if device_ptr is not None:
            check(self.lib.ocpg_get_solution(self.h, C.c_void_p(device_ptr), 1, dSmaxlim, C.byref(out)),
                  "ocpg_get_solution")
        else:
            u = np.ascontiguousarray(u, dtype=np.float64)
            check(self.lib.ocpg_get_solution(self.h, u.ctypes.data, 0, dSmaxlim, C.byref(out)),
                  "ocpg_get_solution")
        return out.NRdPmax, out.NRdNmax

    def nr_change(self):
        """After the flash that followed get_solution: NRdSmax and maxNRdSSP
        with their cells."""
        out = abi.NRChange()
        check(self.lib.ocpg_nr_change(self.h, C.byref(out)), "ocpg_nr_change")
        return dict(NRdSmax=out.NRdSmax, NRdSmax_cell=out.NRdSmax_cell, maxNRdSSP=out.maxNRdSSP,
                    maxNRdSSP_cell=out.maxNRdSSP_cell)

    def get_well_bhp(self):
        out = np.zeros(self.wells.nw)
        check(self.lib.ocpg_get_well_bhp(self.h, out.ctypes.data), "ocpg_get_well_bhp")
        return out

    def rows(self):
        return int(self.lib.ocpg_bsr_rows(self.h))

    def flash_fim_guided(self, guess, ld):
        """FlashFIM with a learned guess: `guess` is a device tensor of nb rows
        of ld float32 (ln K_1..ln K_NC, stability logit)."""
        st = abi.FlashStats()
        check(self.lib.ocpg_flash_fim_guided(self.h, C.c_void_p(guess.data_ptr()), ld, C.byref(st)),
              "ocpg_flash_fim_guided")
        return st

    def set_learned(self, W=None, b=None):
        """The learned form's network (W_l (in x out), b_l float32 arrays, 4
        layers); None removes it."""
        if W is None:
            check(self.lib.ocpg_set_learned(self.h, 0, 0, *([None] * 8)), "ocpg_set_learned")
            self._learned = None
            return
        arrs = []
        for l in range(4):
            arrs += [np.ascontiguousarray(W[l], dtype=np.float32), np.ascontiguousarray(b[l], dtype=np.float32)]
        self._learned = arrs
        check(self.lib.ocpg_set_learned(self.h, arrs[0].shape[0], arrs[6].shape[1],
                                        *[a.ctypes.data for a in arrs]), "ocpg_set_learned")

    def flash_fim_learned(self, stats=True):
        st = abi.FlashStats() if stats else None
        check(self.lib.ocpg_flash_fim_learned(self.h, C.byref(st) if st else None), "ocpg_flash_fim_learned")
        return st

    def capture_iteration(self, dt):
        """Capture one Newton iteration (restore, flash, kr/Pc + rock, CalRes,
        AssembleMat) into a HIP graph."""
        check(self.lib.ocpg_capture_iteration(self.h, dt), "ocpg_capture_iteration")

    def launch_iteration(self):
        check(self.lib.ocpg_launch_iteration(self.h), "ocpg_launch_iteration")

    def iteration_done(self):
        r = self.lib.ocpg_iteration_done(self.h)
        if r < 0:
            raise OcpgError(f"ocpg_iteration_done failed with code {r}")
        return r == 1

    def set_wells(self, wells):
        """abi.Wells -> ocpg_set_wells (rebuilds the BSR pattern)."""
        ws = wells.cstruct()
        check(self.lib.ocpg_set_wells(self.h, C.byref(ws)), "ocpg_set_wells")
        self.wells = wells

    def set_well_state(self, wells):
        ws = wells.cstruct()
        check(self.lib.ocpg_set_well_state(self.h, C.byref(ws)), "ocpg_set_well_state")

    def set_well_setup(self, setup):
        """abi.WellSetup -> ocpg_set_well_setup (after set_wells)."""
        su = setup.cstruct()
        check(self.lib.ocpg_set_well_setup(self.h, C.byref(su)), "ocpg_set_well_setup")

    def prepare_wells(self):
        """AllWells::PrepareWell on the device."""
        check(self.lib.ocpg_prepare_wells(self.h), "ocpg_prepare_wells")

    def check_wells(self):
        """AllWells::CheckP: abi.WELL_SUCCESS / WELL_NEGATIVE_PRESSURE / WELL_SWITCH."""
        st = C.c_int32()
        check(self.lib.ocpg_check_wells(self.h, C.byref(st)), "ocpg_check_wells")
        return st.value

    def get_well_state(self):
        """(mode, bhp, perfDG, perfXi, perfMult) held on the device."""
        w = self.wells
        out = (np.zeros(w.nw, dtype=np.int32), np.zeros(w.nw), np.zeros(max(w.nperf, 1)),
               np.zeros(max(w.nperf, 1)), np.zeros(max(w.nperf, 1)))
        check(self.lib.ocpg_get_well_state(self.h, *[a.ctypes.data for a in out]), "ocpg_get_well_state")
        return tuple(a[:w.nperf] if i >= 2 else a for i, a in enumerate(out))

    def phase_density(self, P, T, zi, phase):
        """MixtureComp::XiPhase / RhoPhase: (xi, rho)."""
        zi = np.ascontiguousarray(zi, dtype=np.float64)
        xi, rho = C.c_double(), C.c_double()
        check(self.lib.ocpg_phase_density(self.h, P, T, zi.ctypes.data, phase, C.byref(xi), C.byref(rho)),
              "ocpg_phase_density")
        return xi.value, rho.value

    def check_bulk(self):
        """Bulk::CheckNi then CheckP: 0 or abi.BULK_NEGATIVE_*."""
        st = C.c_int32()
        check(self.lib.ocpg_check_bulk(self.h, C.byref(st)), "ocpg_check_bulk")
        return st.value

    def max_change(self):
        """Bulk::CalMaxChange + AllWells::CalMaxBHPChange against the last step."""
        out = abi.MaxChange()
        check(self.lib.ocpg_max_change(self.h, C.byref(out)), "ocpg_max_change")
        return dict(dPmax=out.dPmax, dNmax=out.dNmax, dSmax=out.dSmax, eVmax=out.eVmax, dBHPmax=out.dBHPmax)

    def get_well_flux(self):
        w = self.wells
        nc = w.nc
        out = (np.zeros(w.nperf * nc), np.zeros(w.nw * nc), np.zeros(w.nw))
        check(self.lib.ocpg_get_well_flux(self.h, *[a.ctypes.data for a in out]), "ocpg_get_well_flux")
        return out

    def get_bsr(self):
        nnzb = self.nnzb()
        nrow = self.rows()
        IA = np.zeros(nrow + 1, dtype=np.int32)
        JA = np.zeros(nnzb, dtype=np.int32)
        val = np.zeros(nnzb * self.ncol * self.ncol)
        rhs = np.zeros(nrow * self.ncol)
        check(self.lib.ocpg_get_bsr(self.h, IA.ctypes.data, JA.ctypes.data, val.ctypes.data,
                                    rhs.ctypes.data), "ocpg_get_bsr")
        return IA, JA, val, rhs

    def pinned_bsr(self):
        """Page-locked host arrays (IA, JA, val, rhs) for get_bsr_async; freed
        with free_pinned."""
        rows, nnzb, ncol = self.rows(), self.nnzb(), self.ncol
        specs = [(rows + 1, np.int32), (nnzb, np.int32), (nnzb * ncol * ncol, np.float64),
                 (rows * ncol, np.float64)]
        out = []
        for n, dt in specs:
            nbytes = max(int(n) * np.dtype(dt).itemsize, 1)
            p = self.lib.ocpg_host_alloc(nbytes)
            if not p:
                raise OcpgError("ocpg_host_alloc failed")
            buf = (C.c_char * nbytes).from_address(p)
            out.append((p, np.frombuffer(buf, dtype=dt, count=int(n))))
        return out

    def free_pinned(self, arrays):
        for p, _ in arrays:
            self.lib.ocpg_host_free(C.c_void_p(p))

    def get_bsr_async(self, arrays):
        """Queue the BSR hand-off into pinned arrays (pinned_bsr); wait with
        synchronize()."""
        check(self.lib.ocpg_get_bsr_async(self.h, *[C.c_void_p(p) for p, _ in arrays]), "ocpg_get_bsr_async")

    def get_upwind(self):
        n = self.grid.nconn * abi.NP
        up = np.zeros(n, dtype=np.int32)
        urho = np.zeros(n)
        uvel = np.zeros(n)
        check(self.lib.ocpg_get_upwind(self.h, up.ctypes.data, urho.ctypes.data, uvel.ctypes.data),
              "ocpg_get_upwind")
        return up, urho, uvel

    # ---- serving pool ----
    def pool_init(self, capacity, max_neighbours=6):
        check(self.lib.ocpg_pool_init(self.h, int(capacity), int(max_neighbours)), "ocpg_pool_init")
        self.nb = int(capacity)
        self.grid = None

    def pool_place(self, base, bulk, grid):
        s = bulk.cstruct()
        conn = abi.conn_struct(grid.bId, grid.eId, grid.area)
        check(self.lib.ocpg_pool_place(self.h, int(base), C.byref(s), C.byref(conn)), "ocpg_pool_place")

    def pool_activate(self, base, on=True):
        check(self.lib.ocpg_pool_activate(self.h, int(base), 1 if on else 0), "ocpg_pool_activate")

    def pool_system(self, base, nb):
        """(IA, JA, val, rhs) of the realisation at base, in its own indexing."""
        nnzb = C.c_int64()
        check(self.lib.ocpg_pool_system(self.h, int(base), C.byref(nnzb), None, None, None, None),
              "ocpg_pool_system")
        k = nnzb.value
        IA = np.zeros(nb + 1, dtype=np.int32)
        JA = np.zeros(k, dtype=np.int32)
        val = np.zeros(k * self.ncol * self.ncol)
        rhs = np.zeros(nb * self.ncol)
        check(self.lib.ocpg_pool_system(self.h, int(base), None, IA.ctypes.data, JA.ctypes.data,
                                        val.ctypes.data, rhs.ctypes.data), "ocpg_pool_system")
        return IA, JA, val, rhs

    def pool_download(self, base, bulk, saved=True):
        s = bulk.cstruct()
        check(self.lib.ocpg_pool_download(self.h, int(base), C.byref(s), 1 if saved else 0), "ocpg_pool_download")

    def pool_remove(self, base):
        check(self.lib.ocpg_pool_remove(self.h, int(base)), "ocpg_pool_remove")

    def pool_image_bytes(self, base):
        n = int(self.lib.ocpg_pool_image_bytes(self.h, int(base)))
        if n < 0:
            raise OcpgError(f"ocpg_pool_image_bytes failed with code {n}")
        return n

    def pool_export(self, base, image_ptr, all=True):
        check(self.lib.ocpg_pool_export(self.h, int(base), image_ptr, 1 if all else 0), "ocpg_pool_export")

    def pool_import(self, base, image_ptr):
        check(self.lib.ocpg_pool_import(self.h, int(base), image_ptr), "ocpg_pool_import")

    def pool_sync_copies(self):
        check(self.lib.ocpg_pool_sync_copies(self.h), "ocpg_pool_sync_copies")

    def save_flash_inputs(self):
        check(self.lib.ocpg_save_flash_inputs(self.h), "ocpg_save_flash_inputs")

    def restore_flash_inputs(self):
        check(self.lib.ocpg_restore_flash_inputs(self.h), "ocpg_restore_flash_inputs")

    def timing(self, on=True):
        check(self.lib.ocpg_timing_enable(self.h, 1 if on else 0), "ocpg_timing_enable")

    def timer(self, name):
        ms = C.c_double()
        cnt = C.c_int64()
        check(self.lib.ocpg_timing_get(self.h, name.encode(), C.byref(ms), C.byref(cnt)),
              "ocpg_timing_get")
        return ms.value, cnt.value

    def synchronize(self):
        check(self.lib.ocpg_synchronize(self.h), "ocpg_synchronize")

    def device_bytes(self):
        return int(self.lib.ocpg_device_bytes(self.h))
