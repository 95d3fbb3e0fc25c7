"""ctypes binding of libfest_hip.so, the C ABI declared in include/fest_hip.h.

A `GpuBlock` mirrors the reference's `update` module surface for one block:
``set_grid``/``upload_geometry`` (setup_geometry), ``upload_state``,
``get_next_solution`` (src/update.f90:129-226), ``residual``
(get_total_conservative_Residue, src/update.f90:495-547), ``time_step``
(compute_time_step, src/time.f90:291-334) and ``resnorm`` (find_resnorm's
absolute norms, src/resnorm.f90:171-225).  Errors raise `FestError`
carrying the C ABI code, the counterpart of the reference's Fatal_error.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .cases import FLUX, LIMITER, RECON, TURB, TIME_ACC, TRANSITION, Case

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libfest_hip.so")

FEST_E = {-1: "FEST_E_NONPHYSICAL", -2: "FEST_E_NAN", -3: "FEST_E_CONFIG", -4: "FEST_E_HIP", -5: "FEST_E_ARG"}

DL = {"qp": 1 << 0, "temp": 1 << 1, "residue": 1 << 2, "F": 1 << 3, "G": 1 << 4, "H": 1 << 5, "dt": 1 << 6,
      "mu": 1 << 7, "mut": 1 << 8, "grads": 1 << 9, "f1": 1 << 10, "dist": 1 << 11, "cells": 1 << 12,
      "faces": 1 << 13}


class FestError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{FEST_E.get(code, code)}: {msg}")
        self.code = code


class Extent(ctypes.Structure):
    _fields_ = [("imx", ctypes.c_int32), ("jmx", ctypes.c_int32), ("kmx", ctypes.c_int32), ("n_var", ctypes.c_int32)]


class Scheme(ctypes.Structure):
    _fields_ = [("flux", ctypes.c_int32), ("recon", ctypes.c_int32), ("turb", ctypes.c_int32),
                ("time_local", ctypes.c_int32), ("time_accuracy", ctypes.c_int32), ("accur", ctypes.c_int32),
                ("lim", ctypes.c_int32 * 3), ("tlim", ctypes.c_int32 * 3), ("global_time_step", ctypes.c_double),
                ("transition", ctypes.c_int32), ("pbs", ctypes.c_int32 * 3), ("muscl_limiter", ctypes.c_int32)]


class Flow(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("density_inf", "x_speed_inf", "y_speed_inf", "z_speed_inf", "pressure_inf", "tk_inf", "tw_inf",
                 "vel_mag", "MInf", "gm", "R_gas", "mu_ref", "T_ref", "Sutherland_temp", "Pr", "tPr")] + \
               [("mu_sutherland", ctypes.c_int32), ("tv_inf", ctypes.c_double), ("tu_inf", ctypes.c_double),
                ("tkl_inf", ctypes.c_double), ("tgm_inf", ctypes.c_double)]


class BC(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int32 * 6)] + [(n, ctypes.c_double * 6) for n in
                ("fixed_density", "fixed_x_speed", "fixed_y_speed", "fixed_z_speed", "fixed_pressure",
                 "fixed_tk", "fixed_tw", "fixed_wall_temperature", "fixed_Tpressure", "fixed_tv", "fixed_tkl",
                 "fixed_tgm")]


_lib = None


def lib():
    """Loads libfest_hip.so; there is no fallback path."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built: run __graft_entry__.build() or make -C fest-3d_amd")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
        L.fest_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Extent), ctypes.POINTER(Scheme),
                                  ctypes.POINTER(Flow), ctypes.POINTER(BC), ip]
        L.fest_destroy.argtypes = [vp]
        L.fest_last_error.argtypes = [vp]
        L.fest_last_error.restype = ctypes.c_char_p
        L.fest_upload_geometry.argtypes = [vp, dp, dp, dp, dp, dp]
        L.fest_set_grid.argtypes = [vp, dp, ip]
        L.fest_upload_state.argtypes = [vp, dp, dp]
        L.fest_get_next_solution.argtypes = [vp, ip, ctypes.c_double]
        L.fest_residual.argtypes = [vp, ip]
        L.fest_vjp.argtypes = [vp, dp, dp, dp]
        L.fest_time_step.argtypes = [vp, ctypes.c_double]
        L.fest_download.argtypes = [vp, ctypes.c_uint32] + [dp] * 14
        L.fest_resnorm.argtypes = [vp, dp]
        L.fest_device_ptr.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t)]
        L.fest_sync.argtypes = [vp]
        L.fest_group_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), ip, ctypes.POINTER(FestLink), ip]
        L.fest_group_destroy.argtypes = [vp]
        L.fest_group_last_error.argtypes = [vp]
        L.fest_group_last_error.restype = ctypes.c_char_p
        L.fest_group_exchange.argtypes = [vp]
        L.fest_group_vjp.argtypes = [vp, ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.POINTER(dp)]
        L.fest_group_wall_dist.argtypes = [vp]
        L.fest_group_resnorm.argtypes = [vp, dp]
        L.fest_group_get_next_solution.argtypes = [vp, ip, ctypes.c_double]
        L.fest_timing.argtypes = [vp, ctypes.c_char_p]
        L.fest_kernel_time.argtypes = [vp, ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_long)]
        L.fest_set_option.argtypes = [vp, ctypes.c_char_p, ctypes.c_long]
        L.fest_group_set_option.argtypes = [vp, ctypes.c_char_p, ctypes.c_long]
        L.fest_group_autotune.argtypes = [vp, ip, dp, ip, ctypes.POINTER(ip), ctypes.POINTER(ip)]
        L.fest_group_autotune.restype = ctypes.c_int
        for f in ("fest_set_option", "fest_group_set_option", "fest_create", "fest_upload_geometry",
                  "fest_set_grid", "fest_upload_state",
                  "fest_get_next_solution", "fest_residual", "fest_vjp", "fest_time_step", "fest_download", "fest_resnorm",
                  "fest_device_ptr", "fest_sync", "fest_timing", "fest_kernel_time", "fest_group_create",
                  "fest_group_exchange", "fest_group_vjp", "fest_group_wall_dist", "fest_group_resnorm",
                  "fest_group_get_next_solution"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


class FestLink(ctypes.Structure):
    """fest_link (include/fest_hip.h)."""
    _fields_ = [(n, ctypes.c_int32) for n in
                ("block", "face", "from_block", "from_face", "lo1", "hi1", "lo2", "hi2", "dir_switch")]


def c_structs(case: Case):
    e = Extent(case.imx, case.jmx, case.kmx, case.n_var)
    s = Scheme()
    s.flux, s.recon, s.turb = FLUX[case.flux], RECON[case.recon], TURB[case.turb]
    s.time_local, s.time_accuracy, s.accur = case.time_local, TIME_ACC[case.time_accuracy], case.accur
    s.lim[:] = list(case.lim)
    s.tlim[:] = list(case.tlim)
    s.global_time_step = case.global_time_step
    s.transition = TRANSITION[case.transition]
    s.pbs[:] = list(case.pbs)
    s.muscl_limiter = LIMITER[case.limiter]
    f = Flow()
    f.density_inf, f.x_speed_inf, f.y_speed_inf, f.z_speed_inf = case.rho_inf, case.u_inf, case.v_inf, case.w_inf
    f.pressure_inf, f.tk_inf, f.tw_inf, f.vel_mag, f.MInf = case.p_inf, case.tk_inf, case.tw_inf, case.vel_mag, case.MInf
    f.gm, f.R_gas, f.mu_ref, f.T_ref, f.Sutherland_temp = case.gm, case.R_gas, case.mu_ref, case.T_ref, case.S_ref
    f.Pr, f.tPr, f.mu_sutherland = case.Pr, case.tPr, case.mu_sutherland
    f.tv_inf = case.tv_inf
    f.tu_inf = case.tu_inf
    f.tkl_inf = case.tkl_inf
    f.tgm_inf = case.tgm_inf
    b = BC()
    b.id[:] = list(case.bc_id)
    b.fixed_density[:] = list(case.fix_rho)
    b.fixed_x_speed[:] = list(case.fix_u)
    b.fixed_y_speed[:] = list(case.fix_v)
    b.fixed_z_speed[:] = list(case.fix_w)
    b.fixed_pressure[:] = list(case.fix_p)
    b.fixed_tk[:] = list(case.fix_tk)
    b.fixed_tw[:] = list(case.fix_tw)
    b.fixed_wall_temperature[:] = list(case.fix_twall)
    b.fixed_Tpressure[:] = list(case.fix_tp)
    b.fixed_tv[:] = list(case.fix_tv)
    b.fixed_tkl[:] = list(case.fix_tkl)
    b.fixed_tgm[:] = list(case.fix_tgm)
    return e, s, f, b


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


class GpuBlock:
    """One device-resident block."""

    def __init__(self, case: Case, precision: int = 0):
        self.case = case
        self.h = ctypes.c_void_p()
        self._structs = c_structs(case)
        e, s, f, b = self._structs
        rc = lib().fest_create(ctypes.byref(self.h), ctypes.byref(e), ctypes.byref(s), ctypes.byref(f),
                               ctypes.byref(b), precision)
        self._check(rc)

    def _check(self, rc):
        if rc != 0:
            msg = lib().fest_last_error(self.h).decode() if self.h else ""
            raise FestError(rc, msg)

    def close(self):
        if self.h:
            lib().fest_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup
    def set_grid(self, nodes: np.ndarray, wall_dist: bool = False):
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        assert nodes.shape == self.case.node_shape()
        self._check(lib().fest_set_grid(self.h, _ptr(nodes), int(bool(wall_dist))))

    def upload_geometry(self, cells4, if4, jf4, kf4, dist=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (cells4, if4, jf4, kf4)]
        d = None if dist is None else np.ascontiguousarray(dist, dtype=np.float64)
        self._check(lib().fest_upload_geometry(self.h, *[_ptr(a) for a in arrs], _ptr(d)))

    def upload_state(self, qp: np.ndarray, temp: np.ndarray = None):
        qp = np.ascontiguousarray(qp, dtype=np.float64)
        assert qp.shape == self.case.cell_shape(self.case.n_var)
        t = None if temp is None else np.ascontiguousarray(temp, dtype=np.float64)
        self._check(lib().fest_upload_state(self.h, _ptr(qp), _ptr(t)))

    # -- hot path
    def get_next_solution(self, current_iter: int, cfl: float = None):
        cfl = self.case.CFL if cfl is None else cfl
        self._check(lib().fest_get_next_solution(self.h, int(current_iter), float(cfl)))

    def residual(self, current_iter: int = 1):
        self._check(lib().fest_residual(self.h, int(current_iter)))

    def vjp(self, bar_residue: np.ndarray, nodes: bool = False):
        """fest_vjp: (dR/dqp)^T bar_residue for the inviscid residual R of
        the resident state.  bar_residue has the residue's shape
        (5, kmx-1, jmx-1, imx-1); the result the state's (5, padded cells),
        zero in ghost cells but those of interface faces.  With nodes, also
        (dR/dnodes)^T bar_residue in the node array's shape (3, padded
        nodes), zero at ghost nodes: returns (bar_qp, bar_nodes)."""
        c = self.case
        bar = np.ascontiguousarray(bar_residue, dtype=np.float64)
        assert bar.shape == self.shape_of("residue")
        out = np.zeros((5,) + c.cell_shape())
        bn = np.zeros(c.node_shape()) if nodes else None
        self._check(lib().fest_vjp(self.h, _ptr(bar), _ptr(out), _ptr(bn)))
        return (out, bn) if nodes else out

    def time_step(self, cfl: float = None):
        cfl = self.case.CFL if cfl is None else cfl
        self._check(lib().fest_time_step(self.h, float(cfl)))

    def resnorm(self) -> np.ndarray:
        out = np.zeros(16)
        self._check(lib().fest_resnorm(self.h, _ptr(out)))
        return out[: self.case.n_var + 1]

    def sync(self):
        self._check(lib().fest_sync(self.h))

    def set_option(self, name: str, value: int):
        """fest_set_option: "graphs" (0/1), "check_interval" (n >= 1)."""
        self._check(lib().fest_set_option(self.h, name.encode(), int(value)))

    KERNEL_GROUPS = ("step_begin", "bc_face", "bc_edges", "grad", "grad_bc", "flux_i", "flux_j", "flux_k",
                     "time_step", "update", "lusgs_diag", "lusgs_sweep", "lusgs_apply", "kkl_source", "vjp")

    def timing(self, groups=None):
        """Resets the event totals and brackets the launches of the named
        kernel groups with HIP events from now on ("*" all, None none)."""
        if groups is not None and not isinstance(groups, str):
            groups = ",".join(groups)
        self._check(lib().fest_timing(self.h, groups.encode() if groups else None))

    def kernel_time(self, group):
        """(summed event milliseconds, bracketed launches) of one group."""
        ms = ctypes.c_double()
        n = ctypes.c_long()
        self._check(lib().fest_kernel_time(self.h, group.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # -- downloads in the reference layout
    def shape_of(self, name):
        c = self.case
        nv, ng = c.n_var, c.n_grad
        cell = c.cell_shape()
        return {
            "qp": (nv,) + cell, "temp": cell, "mu": cell, "mut": cell, "f1": cell, "dist": cell,
            "residue": (nv, c.kmx - 1, c.jmx - 1, c.imx - 1), "dt": (c.kmx - 1, c.jmx - 1, c.imx - 1),
            "F": (nv, c.kmx - 1, c.jmx - 1, c.imx), "G": (nv, c.kmx - 1, c.jmx, c.imx - 1),
            "H": (nv, c.kmx, c.jmx - 1, c.imx - 1), "grads": (3, ng, c.kmx + 1, c.jmx + 1, c.imx + 1),
            "cells": (4,) + cell,
        }[name]

    def download(self, *names):
        order = ["qp", "temp", "residue", "F", "G", "H", "dt", "mu", "mut", "grads", "f1", "dist", "cells", "faces"]
        bufs = {}
        mask = 0
        for n in names:
            if n == "faces":
                c = self.case
                nf = [(c.kmx + 5) * (c.jmx + 5) * (c.imx + 6), (c.kmx + 5) * (c.jmx + 6) * (c.imx + 5),
                      (c.kmx + 6) * (c.jmx + 5) * (c.imx + 5)]
                bufs[n] = np.zeros(4 * sum(nf))
            else:
                bufs[n] = np.zeros(self.shape_of(n))
            mask |= DL[n]
        args = [_ptr(bufs.get(n)) for n in order]
        self._check(lib().fest_download(self.h, mask, *args))
        if "faces" in bufs:
            c = self.case
            flat, out, off = bufs["faces"], {}, 0
            shapes = [(c.kmx + 5, c.jmx + 5, c.imx + 6), (c.kmx + 5, c.jmx + 6, c.imx + 5),
                      (c.kmx + 6, c.jmx + 5, c.imx + 5)]
            for d, shp in enumerate(shapes):
                n = int(np.prod(shp))
                out[d] = flat[off: off + 4 * n].reshape((4,) + shp)
                off += 4 * n
            bufs["faces"] = out
        return bufs if len(names) > 1 else bufs[names[0]]

    def device_ptr(self, name):
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        self._check(lib().fest_device_ptr(self.h, name.encode(), ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value


class GpuGroup:
    """A multiblock grid over GpuBlocks joined by interface links
    (festgpu.multiblock.Link tuples); fest_group_* of the C ABI."""

    def __init__(self, blocks, links):
        self.blocks = list(blocks)
        arr = (FestLink * max(len(links), 1))()
        for n, L in enumerate(links):
            arr[n] = FestLink(*[int(v) for v in L])
        hs = (ctypes.c_void_p * len(self.blocks))(*[b.h for b in self.blocks])
        self.g = ctypes.c_void_p()
        rc = lib().fest_group_create(ctypes.byref(self.g), hs, len(self.blocks), arr, len(links))
        if rc != 0:
            msg = lib().fest_group_last_error(self.g).decode() if self.g else ""
            self.close()
            raise FestError(rc, msg)

    def _check(self, rc):
        if rc != 0:
            raise FestError(rc, lib().fest_group_last_error(self.g).decode())

    def vjp(self, bar_residues, nodes: bool = False):
        """fest_group_vjp: the blocks' cell adjoints (each (5, padded
        cells)) for the per-block residue adjoints bar_residues, the
        interface ghost shares added to the cells they came from; with
        nodes, also each block's node adjoint: returns (cells, nodes)."""
        ws = [np.ascontiguousarray(w, dtype=np.float64) for w in bar_residues]
        assert len(ws) == len(self.blocks)
        outs = [np.zeros((5,) + blk.case.cell_shape()) for blk in self.blocks]
        dp = ctypes.POINTER(ctypes.c_double)
        win = (dp * len(ws))(*[_ptr(w) for w in ws])
        wout = (dp * len(outs))(*[_ptr(o) for o in outs])
        bns = [np.zeros(blk.case.node_shape()) for blk in self.blocks] if nodes else None
        wn = (dp * len(bns))(*[_ptr(o) for o in bns]) if nodes else None
        self._check(lib().fest_group_vjp(self.g, win, wout, wn))
        return (outs, bns) if nodes else outs

    def wall_dist(self):
        """Wall distance of every block to the walls of all blocks."""
        self._check(lib().fest_group_wall_dist(self.g))

    def exchange(self):
        self._check(lib().fest_group_exchange(self.g))

    def resnorm(self) -> np.ndarray:
        """Res_abs(0:n_var) of the whole group (find_resnorm)."""
        out = np.zeros(16)
        self._check(lib().fest_group_resnorm(self.g, _ptr(out)))
        return out[: self.blocks[0].case.n_var + 1]

    def set_option(self, name: str, value: int):
        """fest_group_set_option: "graphs" (0/1), "check_interval" (n >= 1),
        the schedule "batched" (0/1), "grid_cap" (work-groups per batched
        launch, 0 one per tile), "order" (0/1)."""
        self._check(lib().fest_group_set_option(self.g, name.encode(), int(value)))

    AUTOTUNE_CANDIDATES = ["per-block"] + [f"batched cap={c} order={o} tile={t}x{256 // t}"
                                           for t in (64, 32) for o in (0, 1)
                                           for c in ("tiles", "16/CU", "8/CU", "4/CU")]

    def autotune(self, steps: int = 4):
        """fest_group_autotune: times every schedule on a snapshot of the
        state (restored afterwards) and keeps the fastest.  Returns
        (chosen index, [ms per step of each candidate])."""
        ms = np.zeros(32)
        nc, ch = ctypes.c_int(), ctypes.c_int()
        self._check(lib().fest_group_autotune(self.g, int(steps), _ptr(ms), 32, ctypes.byref(nc), ctypes.byref(ch)))
        return ch.value, list(ms[: nc.value])

    def get_next_solution(self, current_iter: int, cfl: float = None):
        c = self.blocks[0].case.CFL if cfl is None else cfl
        self._check(lib().fest_group_get_next_solution(self.g, int(current_iter), float(c)))

    def close(self):
        if getattr(self, "g", None):
            lib().fest_group_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
