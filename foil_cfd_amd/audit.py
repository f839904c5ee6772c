"""VG vortex-strength audit (src/geom/vg_audit.h): the Wendt correlation, the
signed streamwise-circulation integrals over a crossflow band (host planes),
and the audit against a live solver (band integral on the device)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import FcBandCirculation, FcPlaneCirculation, FcVGAuditReadout, check, load

_fp = C.POINTER(C.c_float)


def wendt_circulation_lattice(beta_rad, ue_lat, h_cells, vane_chord_cells, delta99_cells) -> float:
    return float(load().fc_wendt_circulation_lattice(beta_rad, ue_lat, h_cells, vane_chord_cells, delta99_cells))


def _planes(v, w):
    v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1)
    w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
    return v, w


def integrate_streamwise_circulation(v, w, ny, nz, y_min, y_max):
    """PlaneCirculation (pos, neg) of a host plane indexed y + ny*z."""
    v, w = _planes(v, w)
    if v.size < ny * nz or w.size < ny * nz:
        return 0.0, 0.0
    o = FcPlaneCirculation()
    check(load().fc_integrate_streamwise_circulation(v.ctypes.data_as(_fp), w.ctypes.data_as(_fp), ny, nz,
                                                     y_min, y_max, C.byref(o)), "fc_integrate_streamwise_circulation")
    return o.pos, o.neg


def integrate_band_circulation(v, w, ny, nz, y_min, y_max, z0, z1) -> dict:
    v, w = _planes(v, w)
    b = FcBandCirculation()
    if v.size >= ny * nz and w.size >= ny * nz:
        check(load().fc_integrate_band_circulation(v.ctypes.data_as(_fp), w.ctypes.data_as(_fp), ny, nz, y_min,
                                                   y_max, z0, z1, C.byref(b)), "fc_integrate_band_circulation")
    return {"zone": (b.zone.pos, b.zone.neg), "ambient": (b.ambient.pos, b.ambient.neg),
            "zone_cols": b.zone_cols, "ambient_cols": b.ambient_cols}


def audit_vg_vortex_strength(solver, vg, chord_cells, ue_lat, delta99_cells) -> dict:
    r = FcVGAuditReadout()
    check(load().fc_audit_vg_vortex_strength(solver._h, C.byref(vg), int(chord_cells), float(ue_lat),
                                             float(delta99_cells), C.byref(r)), "fc_audit_vg_vortex_strength")
    return {"valid": bool(r.valid), "gamma_measured": r.gamma_measured, "gamma_predicted": r.gamma_predicted,
            "ratio": r.ratio, "plane_xc": r.plane_xc}
