"""Python mirror of the reference's solver interface over the C ABI.

``LBMSolver`` follows ``class LBMSolver`` (src/sim/lbm_solver.h:133-514):
init -> [reset] -> step_n, force / field readouts, exact-restore upload.
Units follow src/sim/units.h and the flag builders src/geom/voxelizer.h.
All device work happens in the HIP library; this module only marshals
host arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import (FcDims, FcForceReadout, FcInitOptions, FcLayout, FcMlpParams, FcPatchBox,
                   FcPerfStats, FcScaling, FcWallModelReadout, check, load)

Q = 19
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0], dtype=np.int64)
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1], dtype=np.int64)
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1], dtype=np.int64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)

FLUID, SOLID, INLET, OUTLET = 0, 1, 2, 3
SLIP_TOP, SLIP_BOTTOM, SLIP_FRONT, SLIP_BACK, INTERFACE = 4, 5, 6, 7, 8


def _u8(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _f32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ---------------------------------------------------------------- units ----
def compute_scaling(chord_m=1.2, airspeed_ms=30.0, nu_m2s=1.48e-5, chord_cells=256,
                    u_lat=0.08) -> FcScaling:
    """computeScaling (units.h:192-219)."""
    s = FcScaling()
    check(load().fc_compute_scaling(chord_m, airspeed_ms, nu_m2s, chord_cells, u_lat,
                                    C.byref(s)), "fc_compute_scaling")
    return s


def scaling_for_tau(tau: float, u_lat: float, chord_cells: int) -> FcScaling:
    """A self-consistent scaling whose lattice viscosity gives exactly ``tau``
    (chord 1 m, airspeed = u_lat m/s so dx = 1/Nc and dt = dx)."""
    nu_lat = (tau - 0.5) / 3.0
    dx = 1.0 / chord_cells
    nu_phys = nu_lat * dx  # nu = nu_lat dx^2 / dt with dt = dx
    s = compute_scaling(1.0, u_lat, nu_phys, chord_cells, u_lat)
    s.tau = tau
    s.nu_lat = nu_lat
    s.nu_lat_target = nu_lat
    return s


def refined_scaling(parent: FcScaling, factor: int) -> FcScaling:
    out = FcScaling()
    check(load().fc_refined_scaling(C.byref(parent), factor, C.byref(out)), "fc_refined_scaling")
    return out


def ramped_tau(s: FcScaling, step: int, nx: int) -> float:
    return load().fc_ramped_tau(C.byref(s), step, nx)


def ramped_u(s: FcScaling, step: int, nx: int) -> float:
    return load().fc_ramped_u(C.byref(s), step, nx)


def force_to_coefficient(s: FcScaling, force: float, span: int) -> float:
    return load().fc_force_to_coefficient(C.byref(s), force, span)


def flow_through_steps(s: FcScaling, nx: int) -> float:
    return load().fc_flow_through_steps(C.byref(s), nx)


# ------------------------------------------------------------- geometry ----
def generate_naca4(digits: str, points_per_surface: int = 200):
    """generateNACA4 (airfoil.cpp:177-266) -> (points (n,2) float32, le_index)."""
    lib = load()
    cap = 2 * max(points_per_surface, 20)
    xy = np.zeros((cap, 2), dtype=np.float32)
    le = C.c_int(0)
    err = C.create_string_buffer(256)
    n = lib.fc_generate_naca4(digits.encode(), points_per_surface, _f32(xy), cap, C.byref(le),
                              err, 256)
    if n < 0:
        raise ValueError(err.value.decode())
    return xy[:n].copy(), le.value


def make_layout(nx, ny, nz, chord_cells, anchor_x_frac=0.3, anchor_y_frac=0.5,
                anchor_x_cells=-1.0, anchor_y_cells=-1.0, anchor_absolute=False) -> FcLayout:
    return FcLayout(FcDims(nx, ny, nz), chord_cells, anchor_x_frac, anchor_y_frac,
                    anchor_x_cells, anchor_y_cells, 1 if anchor_absolute else 0)


def build_boundary_flags(nx, ny, nz) -> np.ndarray:
    out = np.zeros(nx * ny * nz, dtype=np.uint8)
    check(load().fc_build_boundary_flags(FcDims(nx, ny, nz), _u8(out)), "fc_build_boundary_flags")
    return out


def voxelize_airfoil(points: np.ndarray, aoa_deg: float, layout: FcLayout,
                     flags: np.ndarray) -> None:
    pts = np.ascontiguousarray(points, dtype=np.float32)
    check(load().fc_voxelize_airfoil(_f32(pts), len(pts), aoa_deg, C.byref(layout), _u8(flags)),
          "fc_voxelize_airfoil")


def close_trailing_edge_gaps(nx, ny, nz, flags: np.ndarray) -> int:
    return load().fc_close_trailing_edge_gaps(FcDims(nx, ny, nz), _u8(flags))


def build_clean_foil_flags(points: np.ndarray, aoa_deg: float, layout: FcLayout) -> np.ndarray:
    d = layout.dims
    out = np.zeros(d.nx * d.ny * d.nz, dtype=np.uint8)
    pts = np.ascontiguousarray(points, dtype=np.float32)
    check(load().fc_build_clean_foil_flags(_f32(pts), len(pts), aoa_deg, C.byref(layout),
                                           _u8(out)), "fc_build_clean_foil_flags")
    return out


def derive_patch_box(nx, ny, nz, flags, mx0, mx1, my0, my1, clearance=3) -> FcPatchBox:
    b = FcPatchBox()
    check(load().fc_derive_patch_box(FcDims(nx, ny, nz), _u8(flags), mx0, mx1, my0, my1,
                                     clearance, C.byref(b)), "fc_derive_patch_box")
    return b


def make_fine_layout(coarse: FcLayout, box: FcPatchBox, factor: int) -> FcLayout:
    out = FcLayout()
    check(load().fc_make_fine_layout(C.byref(coarse), C.byref(box), factor, C.byref(out)),
          "fc_make_fine_layout")
    return out


def build_fine_patch_flags(points, aoa_deg, coarse: FcLayout, box: FcPatchBox,
                           factor: int) -> np.ndarray:
    fl = make_fine_layout(coarse, box, factor)
    d = fl.dims
    out = np.zeros(d.nx * d.ny * d.nz, dtype=np.uint8)
    pts = np.ascontiguousarray(points, dtype=np.float32)
    check(load().fc_build_fine_patch_flags(_f32(pts), len(pts), aoa_deg, C.byref(coarse),
                                           C.byref(box), factor, _u8(out)),
          "fc_build_fine_patch_flags")
    return out


def stamp_interface_shell(nx, ny, nz, flags: np.ndarray) -> None:
    check(load().fc_stamp_interface_shell(FcDims(nx, ny, nz), _u8(flags)),
          "fc_stamp_interface_shell")


def build_wall_cell_list(dims, active, clean=None, sample_cells=2):
    """buildWallCellList (wall_model.cpp:169-238) -> (dict of SoA arrays, stats)."""
    lib = load()
    nx, ny, nz = dims
    a = np.ascontiguousarray(active, dtype=np.uint8).reshape(-1)
    c = None if clean is None else np.ascontiguousarray(clean, dtype=np.uint8).reshape(-1)
    st = (C.c_int * 3)()
    n = lib.fc_build_wall_cell_list(FcDims(nx, ny, nz), _u8(a), None if c is None else _u8(c),
                                    sample_cells, None, None, None, None, None, None, None, 0, st)
    out = {"cell": np.zeros(n, np.int64), "sample": np.zeros(n, np.int64),
           "nx": np.zeros(n, np.float32), "ny": np.zeros(n, np.float32),
           "nz": np.zeros(n, np.float32), "ys": np.zeros(n, np.float32),
           "mask": np.zeros(n, np.uint32)}
    ll = C.POINTER(C.c_longlong)
    lib.fc_build_wall_cell_list(FcDims(nx, ny, nz), _u8(a), None if c is None else _u8(c),
                                sample_cells, out["cell"].ctypes.data_as(ll),
                                out["sample"].ctypes.data_as(ll), _f32(out["nx"]), _f32(out["ny"]),
                                _f32(out["nz"]), _f32(out["ys"]),
                                out["mask"].ctypes.data_as(C.POINTER(C.c_uint32)), n, st)
    return out, {"listed": st[0], "excluded_vg": st[1], "degenerate": st[2]}


WALL_OFF, WALL_NEWTON, WALL_MLP = 0, 1, 2

VG_SINGLE, VG_PAIR, VG_COROTATING, VG_RAMP = 0, 1, 2, 3


def vg_params(type=VG_PAIR, x_c=0.07, height_c=0.01, length_h=3.0, beta_deg=16.0, pitch_c=0.06,
              gap_h=2.5, count=4, common_flow_down=True, enabled=True):
    """VGParams (vg.h:24-36; defaults = defaultVGParams, vg.h:38-50)."""
    from ._abi import FcVGParams
    return FcVGParams(type, x_c, height_c, length_h, beta_deg, pitch_c, gap_h, count,
                      1 if common_flow_down else 0, 1 if enabled else 0)


def surface_frame_at(points, le_index, x_c, upper=True):
    pts = np.ascontiguousarray(points, np.float32)
    out = np.zeros(6, np.float32)
    ok = load().fc_surface_frame_at(_f32(pts), len(pts), le_index, x_c, 1 if upper else 0,
                                    _f32(out))
    return (out[0:2], out[2:4], out[4:6]) if ok else None


def build_flags_with_vgs(points, le_index, aoa_deg, layout, vgs, clean):
    """buildFlagsWithVGs (vg.cpp:246-261) -> (flags, list of FcVaneSlab)."""
    from ._abi import FcVaneSlab, FcVGParams
    pts = np.ascontiguousarray(points, np.float32)
    cl = np.ascontiguousarray(clean, np.uint8).reshape(-1)
    out = np.zeros_like(cl)
    arr = (FcVGParams * max(len(vgs), 1))(*vgs)
    cap = 4096
    slabs = (FcVaneSlab * cap)()
    ns = C.c_int(0)
    check(load().fc_build_flags_with_vgs(_f32(pts), len(pts), le_index, aoa_deg, C.byref(layout),
                                         arr, len(vgs), _u8(cl), _u8(out), slabs, cap,
                                         C.byref(ns)), "fc_build_flags_with_vgs")
    return out, [slabs[i] for i in range(min(ns.value, cap))]


def build_vane_qlinks(dims, active, clean, slabs):
    """buildVaneQLinks + densifyQLinks (vg.cpp:300-397) -> (qfrac, ffmask, links)."""
    from ._abi import FcVaneSlab
    nx, ny, nz = dims
    n = nx * ny * nz
    a = np.ascontiguousarray(active, np.uint8).reshape(-1)
    c = np.ascontiguousarray(clean, np.uint8).reshape(-1)
    qf = np.zeros(n * Q, np.uint8)
    ff = np.zeros(n, np.uint32)
    arr = (FcVaneSlab * max(len(slabs), 1))(*slabs)
    links = load().fc_build_vane_qlinks(FcDims(nx, ny, nz), _u8(a), _u8(c), arr, len(slabs),
                                        _u8(qf), ff.ctypes.data_as(C.POINTER(C.c_uint32)))
    if links < 0:
        raise RuntimeError("fc_build_vane_qlinks failed")
    return qf, ff, links


# --------------------------------------------------------------- solver ----
@dataclass
class ForceReadout:
    cl: float
    cd: float
    lift_to_drag: float
    valid: bool
    flow_throughs: float
    cl_avg: float
    cd_avg: float
    cl_min: float = 0.0
    cl_max: float = 0.0
    cl_median: float = 0.0
    cd_min: float = 0.0
    cd_max: float = 0.0
    cd_median: float = 0.0
    ld_avg: float = 0.0
    ld_min: float = 0.0
    ld_max: float = 0.0
    ld_median: float = 0.0


class ForceFold:
    """The solver's force EMA / trailing-average fold on caller-supplied
    samples (fc_force_fold_*; reference lbm_solver.cpp:633-654, :1414-1509)."""

    def __init__(self):
        self._lib = load()
        self._h = self._lib.fc_force_fold_create()
        if not self._h:
            raise MemoryError("fc_force_fold_create failed")

    def close(self):
        if self._h:
            self._lib.fc_force_fold_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        check(self._lib.fc_force_fold_clear(self._h), "fc_force_fold_clear")

    def add(self, at_step: int, d_steps: float, window_steps: float, fx: float, fy: float):
        check(self._lib.fc_force_fold_add(self._h, int(at_step), float(d_steps), float(window_steps),
                                          float(fx), float(fy)), "fc_force_fold_add")

    def readout(self, scaling: FcScaling, span: int, one_ft_steps: int, gate_open: bool) -> ForceReadout:
        r = FcForceReadout()
        check(self._lib.fc_force_fold_readout(self._h, C.byref(scaling), int(span), int(one_ft_steps),
                                              1 if gate_open else 0, C.byref(r)), "fc_force_fold_readout")
        return ForceReadout(r.cl, r.cd, r.lift_to_drag, bool(r.valid), r.flow_throughs,
                            r.cl_avg, r.cd_avg, r.cl_min, r.cl_max, r.cl_median, r.cd_min, r.cd_max,
                            r.cd_median, r.ld_avg, r.ld_min, r.ld_max, r.ld_median)


class LBMSolver:
    """Owner of one device lattice (reference LBMSolver)."""

    def __init__(self):
        self._lib = load()
        self._h = self._lib.fc_create()
        if not self._h:
            raise MemoryError("fc_create failed")
        self.dims = None
        self.ncells = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.fc_destroy(h)
            self._h = None

    def close(self):
        self.__del__()

    # lifecycle -------------------------------------------------------------
    def init(self, dims, scaling: FcScaling, flags: np.ndarray, storage="f32",
             periodic_xy=False) -> None:
        nx, ny, nz = dims
        flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        if flags.size != nx * ny * nz:
            raise ValueError("flag field size does not match grid dimensions")
        st = _abi.STORAGE_NAMES[storage] if isinstance(storage, str) else int(storage)
        opt = FcInitOptions(st, 1 if periodic_xy else 0, 1 if periodic_xy else 0)
        err = C.create_string_buffer(512)
        rc = self._lib.fc_init(self._h, FcDims(nx, ny, nz), C.byref(scaling), _u8(flags),
                               C.byref(opt), err, 512)
        check(rc, "fc_init", err.value.decode())
        self.dims = (nx, ny, nz)
        self.ncells = nx * ny * nz
        self.storage = st

    def reset(self):
        check(self._lib.fc_reset(self._h), "fc_reset")

    def set_flags(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        check(self._lib.fc_set_flags(self._h, _u8(f)), "fc_set_flags")

    def apply_edited_flags(self, flags) -> int:
        """LBMSolver::applyEditedFlags (warm-restart flag edit); returns the
        number of cells fixed."""
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        if f.size != self.ncells:
            raise ValueError("flag field size does not match grid dimensions")
        check(self._lib.fc_apply_edited_flags(self._h, _u8(f)), "fc_apply_edited_flags", self.last_error())
        return int(self._lib.fc_last_edit_count(self._h))

    def restore_from_macroscopic(self, rho, u, v, w) -> None:
        """LBMSolver::restoreFromMacroscopic: f = feq(rho, u), compact restore."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in (rho, u, v, w)]
        if any(a.size != self.ncells for a in arrs):
            raise ValueError("macroscopic field size does not match grid dimensions")
        check(self._lib.fc_restore_from_macroscopic(self._h, *[_f32(a) for a in arrs]),
              "fc_restore_from_macroscopic", self.last_error())

    def last_error(self) -> str:
        buf = C.create_string_buffer(512)
        self._lib.fc_last_error(self._h, buf, 512)
        return buf.value.decode(errors="replace")

    def set_startup_ramp_enabled(self, on: bool):
        check(self._lib.fc_set_startup_ramp(self._h, 1 if on else 0), "fc_set_startup_ramp")

    def set_smagorinsky_cs(self, cs: float):
        check(self._lib.fc_set_smagorinsky_cs(self._h, cs), "fc_set_smagorinsky_cs")

    # refinement -------------------------------------------------------------
    def init_refinement(self, box, factor: int, fine_flags: np.ndarray) -> None:
        """LBMSolver::initRefinement: box = (x0, y0, x1, y1) in coarse cells."""
        b = FcPatchBox(*[int(v) for v in box])
        ff = np.ascontiguousarray(fine_flags, dtype=np.uint8).reshape(-1)
        err = C.create_string_buffer(512)
        rc = self._lib.fc_init_refinement(self._h, C.byref(b), int(factor), _u8(ff), err, 512)
        check(rc, "fc_init_refinement", err.value.decode())

    def shutdown_refinement(self) -> None:
        check(self._lib.fc_shutdown_refinement(self._h), "fc_shutdown_refinement")

    def set_refined_flags(self, fine_flags: np.ndarray) -> None:
        ff = np.ascontiguousarray(fine_flags, dtype=np.uint8).reshape(-1)
        check(self._lib.fc_set_refined_flags(self._h, _u8(ff)), "fc_set_refined_flags")

    def refinement_info(self) -> dict:
        r = _abi.FcRefinementInfo()
        check(self._lib.fc_get_refinement_info(self._h, C.byref(r)), "fc_get_refinement_info")
        return {"active": bool(r.active), "factor": r.factor,
                "box": (r.box.x0, r.box.y0, r.box.x1, r.box.y1),
                "fine_dims": (r.fine_dims.nx, r.fine_dims.ny, r.fine_dims.nz),
                "fine_tau": r.fine_scaling.tau, "fine_scaling": r.fine_scaling,
                "device_bytes": r.device_bytes, "forces_from_fine": bool(r.forces_from_fine)}

    def download_fine_f(self, f64: bool = True) -> np.ndarray:
        info = self.refinement_info()
        n = Q * int(np.prod(info["fine_dims"]))
        out = np.empty(n, dtype=np.float64 if f64 else np.float32)
        check(self._lib.fc_download_fine_f(self._h, out.ctypes.data_as(C.c_void_p), 1 if f64 else 0),
              "fc_download_fine_f")
        return out

    def upload_fine_f(self, f: np.ndarray) -> None:
        a = np.ascontiguousarray(f)
        is64 = a.dtype == np.float64
        a = a if is64 else a.astype(np.float32)
        check(self._lib.fc_upload_fine_f(self._h, a.ctypes.data_as(C.c_void_p), 1 if is64 else 0),
              "fc_upload_fine_f")

    def init_finer_refinement(self, box, factor: int, finer_flags: np.ndarray) -> None:
        """LBMSolver::initFinerRefinement: nested patch, box in fine cells."""
        b = FcPatchBox(*[int(v) for v in box])
        gf = np.ascontiguousarray(finer_flags, dtype=np.uint8).reshape(-1)
        err = C.create_string_buffer(512)
        rc = self._lib.fc_init_finer_refinement(self._h, C.byref(b), int(factor), _u8(gf), err, 512)
        check(rc, "fc_init_finer_refinement", err.value.decode())

    def shutdown_finer_refinement(self) -> None:
        check(self._lib.fc_shutdown_finer_refinement(self._h), "fc_shutdown_finer_refinement")

    def finer_refinement_info(self) -> dict:
        r = _abi.FcRefinementInfo()
        check(self._lib.fc_get_finer_refinement_info(self._h, C.byref(r)), "fc_get_finer_refinement_info")
        return {"active": bool(r.active), "factor": r.factor,
                "box": (r.box.x0, r.box.y0, r.box.x1, r.box.y1),
                "finer_dims": (r.fine_dims.nx, r.fine_dims.ny, r.fine_dims.nz),
                "finer_tau": r.fine_scaling.tau, "device_bytes": r.device_bytes}

    def download_finer_f(self, f64: bool = True) -> np.ndarray:
        info = self.finer_refinement_info()
        n = Q * int(np.prod(info["finer_dims"]))
        out = np.empty(n, dtype=np.float64 if f64 else np.float32)
        check(self._lib.fc_download_finer_f(self._h, out.ctypes.data_as(C.c_void_p), 1 if f64 else 0),
              "fc_download_finer_f")
        return out

    def upload_finer_f(self, f: np.ndarray) -> None:
        a = np.ascontiguousarray(f)
        is64 = a.dtype == np.float64
        a = a if is64 else a.astype(np.float32)
        check(self._lib.fc_upload_finer_f(self._h, a.ctypes.data_as(C.c_void_p), 1 if is64 else 0),
              "fc_upload_finer_f")

    def set_step_variant(self, variant: str):
        """"auto" (library default), "single" (the default cell kernel on every
        sub-step, restriction records from the separate pass), "buf" (one
        cell per lane with buffer addressing) or "wide" ("auto" with the
        64-bit population index of very large grids forced)."""
        v = {"auto": 0, "single": 1, "buf": 2, "wide": 3}[variant]
        check(self._lib.fc_set_step_variant(self._h, v), "fc_set_step_variant")

    def set_stochastic_rounding(self, on: bool):
        """Stochastic (True, the default) or round-to-nearest rounding of the
        16-bit population stores; no effect on FP32 / FP64 storage."""
        check(self._lib.fc_set_stochastic_rounding(self._h, 1 if on else 0), "fc_set_stochastic_rounding")

    def stochastic_rounding(self) -> bool:
        return bool(self._lib.fc_get_stochastic_rounding(self._h))

    def set_force_ema_window(self, ft: float):
        check(self._lib.fc_set_force_ema_window(self._h, ft), "fc_set_force_ema_window")

    # stepping ---------------------------------------------------------------
    def step_n(self, n: int):
        check(self._lib.fc_step(self._h, int(n)), "fc_step")

    def synchronize(self):
        check(self._lib.fc_synchronize(self._h), "fc_synchronize")

    # readouts ---------------------------------------------------------------
    def adaptive_steps_for_budget(self, budget_ms: float = 10.0) -> int:
        """LBMSolver::adaptiveStepsForBudget (frame pacer)."""
        return int(self._lib.fc_adaptive_steps_for_budget(self._h, float(budget_ms)))

    # mesh sequencing ---------------------------------------------------------
    def seed_from_coarse(self, presolver: "LBMSolver") -> None:
        """LBMSolver::seedFromCoarse: upsampled presolve field as the state."""
        check(self._lib.fc_seed_from_coarse(self._h, presolver._h), "fc_seed_from_coarse", self.last_error())

    def preconverge(self, pre_dims, pre_scaling: FcScaling, pre_flags, storage="f32") -> int:
        """runPreconverge (src/main.cpp:756-817) on a presolve grid built by
        the caller; returns the presolver's step count."""
        st = _abi.STORAGE_NAMES[storage] if isinstance(storage, str) else int(storage)
        fl = np.ascontiguousarray(pre_flags, dtype=np.uint8).reshape(-1)
        opt = FcInitOptions(st, 0, 0)
        steps = C.c_longlong(0)
        err = C.create_string_buffer(512)
        rc = self._lib.fc_preconverge(self._h, FcDims(*pre_dims), C.byref(pre_scaling), _u8(fl), C.byref(opt),
                                      C.byref(steps), err, 512)
        check(rc, "fc_preconverge", err.value.decode(errors="replace"))
        return int(steps.value)

    # boundary-layer readouts -----------------------------------------------
    def extract_suction_delta99(self, stations_xc):
        """LBMSolver::extractSuctionDelta99: list of dicts (x_c, delta99_c,
        ue_edge, valid), one per station."""
        from ._abi import FcDelta99Sample
        xs = np.ascontiguousarray(stations_xc, dtype=np.float32).reshape(-1)
        out = (FcDelta99Sample * max(1, xs.size))()
        check(self._lib.fc_extract_suction_delta99(self._h, _f32(xs), xs.size, out),
              "fc_extract_suction_delta99", self.last_error())
        return [{"x_c": o.x_c, "delta99_c": o.delta99_c, "ue_edge": o.ue_edge, "valid": bool(o.valid)}
                for o in out[:xs.size]]

    def separation_onset_xc(self) -> float:
        r = C.c_float(0.0)
        check(self._lib.fc_separation_onset_xc(self._h, C.byref(r)), "fc_separation_onset_xc", self.last_error())
        return float(r.value)

    def download_crossflow_plane(self, ix: int):
        """(v, w) at column ix as (nz, ny) arrays (flat index y + ny*z)."""
        _, ny, nz = self.dims
        v = np.empty(ny * nz, dtype=np.float32)
        w = np.empty(ny * nz, dtype=np.float32)
        check(self._lib.fc_download_crossflow_plane(self._h, int(ix), _f32(v), _f32(w)),
              "fc_download_crossflow_plane", self.last_error())
        return v, w

    def suction_surface_y(self, ix: int) -> int:
        return int(self._lib.fc_suction_surface_y(self._h, int(ix)))

    def lattice_x_for_chord_station(self, xc: float) -> int:
        return int(self._lib.fc_lattice_x_for_chord_station(self._h, float(xc)))

    def band_circulation_live(self, ix, y_min, y_max, z0, z1) -> dict:
        from ._abi import FcBandCirculation
        b = FcBandCirculation()
        check(self._lib.fc_band_circulation_live(self._h, int(ix), int(y_min), int(y_max), int(z0), int(z1),
                                                 C.byref(b)), "fc_band_circulation_live", self.last_error())
        return {"zone": (b.zone.pos, b.zone.neg), "ambient": (b.ambient.pos, b.ambient.neg),
                "zone_cols": b.zone_cols, "ambient_cols": b.ambient_cols}

    def read_flags(self) -> np.ndarray:
        """Host copy of the cell flags (LBMSolver::hostFlags)."""
        out = np.empty(self.ncells, dtype=np.uint8)
        check(self._lib.fc_read_flags(self._h, _u8(out)), "fc_read_flags")
        return out

    def forces(self) -> ForceReadout:
        r = FcForceReadout()
        check(self._lib.fc_forces(self._h, C.byref(r)), "fc_forces")
        return ForceReadout(r.cl, r.cd, r.lift_to_drag, bool(r.valid), r.flow_throughs,
                            r.cl_avg, r.cd_avg, r.cl_min, r.cl_max, r.cl_median, r.cd_min, r.cd_max,
                            r.cd_median, r.ld_avg, r.ld_min, r.ld_max, r.ld_median)

    def force_raw(self):
        out = (C.c_double * 3)()
        step = C.c_longlong(0)
        check(self._lib.fc_force_raw(self._h, out, C.byref(step)), "fc_force_raw")
        return np.array(out[:]), step.value

    def force_now(self) -> np.ndarray:
        out = (C.c_double * 3)()
        check(self._lib.fc_force_now(self._h, out), "fc_force_now")
        return np.array(out[:])

    def force_now_noalloc(self) -> np.ndarray:
        """fc_force_now through the library's static device scratch (no allocation)."""
        out = (C.c_double * 3)()
        check(self._lib.fc_force_now_noalloc(self._h, out), "fc_force_now_noalloc")
        return np.array(out[:])

    def mass(self) -> float:
        out = C.c_double(0)
        check(self._lib.fc_mass(self._h, C.byref(out)), "fc_mass")
        return out.value

    def read_macro(self):
        n = self.ncells
        rho, u, v, w = (np.empty(n, dtype=np.float32) for _ in range(4))
        check(self._lib.fc_read_macro(self._h, _f32(rho), _f32(u), _f32(v), _f32(w)),
              "fc_read_macro")
        return rho, u, v, w

    def download_f(self, f64=False) -> np.ndarray:
        out = np.empty(Q * self.ncells, dtype=np.float64 if f64 else np.float32)
        check(self._lib.fc_download_f(self._h, out.ctypes.data_as(C.c_void_p), 1 if f64 else 0),
              "fc_download_f")
        return out

    def upload_f(self, f: np.ndarray):
        f = np.ascontiguousarray(f).reshape(-1)
        if f.size != Q * self.ncells:
            raise ValueError("population array size mismatch")
        is64 = f.dtype == np.float64
        if not is64:
            f = f.astype(np.float32, copy=False)
        check(self._lib.fc_upload_f(self._h, f.ctypes.data_as(C.c_void_p), 1 if is64 else 0),
              "fc_upload_f")

    def nan_detected(self) -> bool:
        return bool(self._lib.fc_nan_detected(self._h))

    def nan_kind(self) -> int:
        return self._lib.fc_nan_kind(self._h)

    def perf_stats(self) -> FcPerfStats:
        p = FcPerfStats()
        check(self._lib.fc_get_perf_stats(self._h, C.byref(p)), "fc_get_perf_stats")
        return p

    def step_count(self) -> int:
        return self._lib.fc_step_count(self._h)

    def current_tau(self) -> float:
        return self._lib.fc_current_tau(self._h)

    def device_bytes(self) -> int:
        return self._lib.fc_device_bytes(self._h)

    # wall model -------------------------------------------------------------
    def set_wall_model(self, mode):
        m = {"off": WALL_OFF, "newton": WALL_NEWTON, "mlp": WALL_MLP}.get(mode, mode)
        check(self._lib.fc_set_wall_model(self._h, int(m)), "fc_set_wall_model")

    def set_surface_reference(self, clean):
        c = None if clean is None else np.ascontiguousarray(clean, dtype=np.uint8).reshape(-1)
        check(self._lib.fc_set_surface_reference(self._h, None if c is None else _u8(c)),
              "fc_set_surface_reference")

    def wall_model_readout(self) -> FcWallModelReadout:
        r = FcWallModelReadout()
        check(self._lib.fc_get_wall_model_readout(self._h, C.byref(r)), "fc_get_wall_model_readout")
        return r

    def wall_cell_count(self) -> int:
        return self._lib.fc_wall_cell_count(self._h)

    def read_wall_slip(self):
        n = self.ncells
        u16 = C.POINTER(C.c_uint16)
        ux, uy, uz = (np.zeros(n, np.uint16) for _ in range(3))
        utau = np.zeros(max(self.wall_cell_count(), 1), np.float32)
        check(self._lib.fc_read_wall_slip(self._h, ux.ctypes.data_as(u16), uy.ctypes.data_as(u16),
                                          uz.ctypes.data_as(u16), _f32(utau)), "fc_read_wall_slip")
        return ux, uy, uz, utau[: self.wall_cell_count()]

    def wall_cell_count_level(self, level: int) -> int:
        return int(self._lib.fc_wall_cell_count_level(self._h, int(level)))

    def read_wall_slip_level(self, level: int):
        """(uwx, uwy, uwz half bits, u_tau) of level 0 coarse / 1 fine / 2 finer."""
        n = {0: self.ncells}.get(level)
        if n is None:
            info = self.refinement_info() if level == 1 else self.finer_refinement_info()
            fd = info["fine_dims"] if level == 1 else info["finer_dims"]
            n = fd[0] * fd[1] * fd[2]
        cnt = self.wall_cell_count_level(level)
        u16 = C.POINTER(C.c_uint16)
        ux, uy, uz = (np.empty(n, np.uint16) for _ in range(3))
        ut = np.empty(max(cnt, 1), np.float32)
        check(self._lib.fc_read_wall_slip_level(self._h, int(level), ux.ctypes.data_as(u16), uy.ctypes.data_as(u16),
                                                uz.ctypes.data_as(u16), _f32(ut)),
              "fc_read_wall_slip_level", self.last_error())
        return ux, uy, uz, ut[:cnt]

    def set_refined_surface_reference(self, clean_fine, finer: bool = False):
        c = None if clean_fine is None else np.ascontiguousarray(clean_fine, dtype=np.uint8).reshape(-1)
        fn = self._lib.fc_set_refined_finer_surface_reference if finer else self._lib.fc_set_refined_surface_reference
        check(fn(self._h, None if c is None else _u8(c)), "fc_set_refined_surface_reference", self.last_error())

    def wall_update_now(self):
        check(self._lib.fc_wall_update_now(self._h), "fc_wall_update_now")

    def set_wall_harvest(self, on: bool):
        check(self._lib.fc_set_wall_harvest(self._h, 1 if on else 0), "fc_set_wall_harvest")

    def wall_harvest(self) -> np.ndarray:
        n = self._lib.fc_wall_harvest(self._h, None, 0)
        if n < 0:
            raise RuntimeError("fc_wall_harvest failed")
        out = np.zeros((max(n, 1), 4), np.float32)
        got = self._lib.fc_wall_harvest(self._h, _f32(out), n)
        if got < 0:
            raise RuntimeError("fc_wall_harvest failed")
        return out[:got]

    def set_wall_mlp(self, params):
        check(self._lib.fc_set_wall_mlp(self._h, None if params is None else C.byref(params)),
              "fc_set_wall_mlp")

    # q-LIBB ------------------------------------------------------------------
    def set_qlibb_enabled(self, on: bool):
        check(self._lib.fc_set_qlibb_enabled(self._h, 1 if on else 0), "fc_set_qlibb_enabled")

    def set_qlinks(self, qfrac, ffmask, links=0, fallback=0):
        qf = np.ascontiguousarray(qfrac, np.uint8).reshape(-1)
        ff = np.ascontiguousarray(ffmask, np.uint32).reshape(-1)
        if qf.size != Q * self.ncells or ff.size != self.ncells:
            raise ValueError("q-link field size mismatch")
        check(self._lib.fc_set_qlinks(self._h, _u8(qf), ff.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      links, fallback), "fc_set_qlinks")

    def set_fine_qlinks(self, qfrac, ffmask, links=0, fallback=0, level=1):
        """setFineQLinks / setFinerQLinks (lbm_solver.cpp:1049-1063): q-link
        fields over the fine patch (level 1) or the nested finer patch
        (level 2), sized by that level's cells."""
        if level == 1:
            info, key, fn = self.refinement_info(), "fine_dims", "fc_set_fine_qlinks"
        elif level == 2:
            info, key, fn = self.finer_refinement_info(), "finer_dims", "fc_set_finer_qlinks"
        else:
            raise ValueError("level must be 1 or 2")
        n = int(np.prod(info[key]))
        qf = np.ascontiguousarray(qfrac, np.uint8).reshape(-1)
        ff = np.ascontiguousarray(ffmask, np.uint32).reshape(-1)
        # an inactive level is refused by the library (FC_ERR_STATE)
        if info["active"] and (qf.size != Q * n or ff.size != n):
            raise ValueError("q-link field size mismatch")
        check(getattr(self._lib, fn)(self._h, _u8(qf), ff.ctypes.data_as(C.POINTER(C.c_uint32)), links,
                                     fallback), fn)

    def qlibb_readout(self):
        v = [C.c_int(0) for _ in range(4)]
        check(self._lib.fc_qlibb_readout(self._h, *[C.byref(x) for x in v]), "fc_qlibb_readout")
        return {"enabled": v[0].value, "links": v[1].value, "fallback": v[2].value,
                "active": v[3].value}

    # tracer particles ---------------------------------------------------------
    def particles_seed(self, count: int, seed: int = 0xF01DCAFE):
        check(self._lib.fc_particles_seed(self._h, int(count), seed & 0xFFFFFFFF),
              "fc_particles_seed")

    def particles_advect(self, dt_steps: float, seed: int):
        check(self._lib.fc_particles_advect(self._h, dt_steps, seed & 0xFFFFFFFF),
              "fc_particles_advect")

    def particles_read(self) -> np.ndarray:
        n = self._lib.fc_particles_count(self._h)
        out = np.zeros((n, 4), np.float32)
        check(self._lib.fc_particles_read(self._h, _f32(out)), "fc_particles_read")
        return out

    def particles_write(self, pos: np.ndarray):
        pos = np.ascontiguousarray(pos, np.float32)
        check(self._lib.fc_particles_write(self._h, _f32(pos)), "fc_particles_write")

    def particles_set_params(self, age_rate=0.002, inlet_fraction=0.3):
        check(self._lib.fc_particles_set_params(self._h, age_rate, inlet_fraction),
              "fc_particles_set_params")

    # measurement --------------------------------------------------------------
    def set_kernel_timing(self, on: bool):
        check(self._lib.fc_set_kernel_timing(self._h, 1 if on else 0), "fc_set_kernel_timing")

    def kernel_timing(self):
        n = C.c_int(0)
        ms = C.c_double(0)
        check(self._lib.fc_kernel_timing(self._h, C.byref(n), C.byref(ms)), "fc_kernel_timing")
        return n.value, ms.value

    def kernel_timing_level(self, level: int):
        """(step launches, summed ms) of level 0 coarse / 1 fine / 2 finer."""
        n = C.c_int(0)
        ms = C.c_double(0.0)
        check(self._lib.fc_kernel_timing_level(self._h, int(level), C.byref(n), C.byref(ms)),
              "fc_kernel_timing_level")
        return n.value, ms.value

    def time_steps(self, calls: int, steps_per_call: int) -> float:
        ms = C.c_double(0)
        check(self._lib.fc_time_steps(self._h, calls, steps_per_call, C.byref(ms)),
              "fc_time_steps")
        return ms.value
