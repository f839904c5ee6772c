"""Boundary-layer readouts and the VG vortex audit on the GPU against the
oracle's restatement of the reference (lbm_solver.cpp:1511-1655,
vg_audit.cpp:12-187):

* delta99 samples and the separation onset are identical (same FP32
  arithmetic, same float hypot) for planted fields with closed-form answers
  and for a developed tunnel flow;
* the crossflow plane is the column of the macro fields;
* the live band integral (FP64 sums on the device) matches the oracle's FP64
  sums to 1e-6 and recovers a planted Lamb-Oseen pair within the reference
  test's 3 %; the audit matches the oracle's audit."""
import sys

import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd import audit, cases
from foil_cfd_amd import solver as fsolver
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
STATIONS = list(np.linspace(-0.05, 1.05, 23).astype(np.float32))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tunnel(nz=8):
    return cases.naca_tunnel(96, 48, nz, 32, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56, ramp=False)


def planted_profile(case, ue=0.05, height=6.0, seed=0):
    """u rising linearly to ue over `height` cells above each column's
    surface row at mid-span (all z), small random v, reversed near-wall flow
    on a run of columns; rho = 1, w = 0."""
    nx, ny, nz = case.dims
    top, xle, xte, _ = orc.surface_cache(case.flags, case.dims)
    rng = np.random.default_rng(seed)
    u = np.full((nz, ny, nx), ue, np.float32)
    for x in range(nx):
        if top[x] >= 0:
            for y in range(top[x] + 1, ny):
                u[:, y, x] = ue * min(1.0, (y - top[x]) / height)
    v = (rng.standard_normal((nz, ny, nx)) * 1e-4).astype(np.float32)
    x0 = xle + (xte - xle) * 2 // 3
    for x in range(x0, xte + 1):
        if top[x] >= 0 and top[x] + 2 < ny:
            u[:, top[x] + 2, x] = -0.01
            u[:, top[x] + 1, x] = -0.005
    return np.ones(u.size, np.float32), u.reshape(-1), v.reshape(-1), np.zeros(u.size, np.float32), x0


def same_samples(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert np.float32(g["x_c"]) == np.float32(w["x_c"])
        assert np.float32(g["ue_edge"]) == np.float32(w["ue_edge"]), (g, w)
        assert np.float32(g["delta99_c"]) == np.float32(w["delta99_c"]), (g, w)
        assert g["valid"] == w["valid"], (g, w)


def test_readouts_on_planted_fields(gpu):
    say("case: delta99 / separation / crossflow on planted fields, tunnel 96x48x8 f32")
    case = tunnel()
    s = cases.make_solver(case, "f32")
    rho, u, v, w, x0 = planted_profile(case)
    s.restore_from_macroscopic(rho, u, v, w)
    got = s.extract_suction_delta99(STATIONS)
    want = orc.suction_delta99(u, v, case.flags, case.dims, STATIONS, case.scaling.chord_cells, case.scaling.u_lat)
    same_samples(got, want)
    assert sum(o["valid"] for o in got) >= 8 and sum(not o["valid"] and o["delta99_c"] > 0 for o in got) >= 2
    sep = s.separation_onset_xc()
    assert np.float32(sep) == np.float32(orc.separation_onset_xc(u, case.flags, case.dims, case.scaling.u_lat))
    top, xle, xte, _ = orc.surface_cache(case.flags, case.dims)
    assert sep == pytest.approx((x0 - xle) / (xte - xle), abs=1e-6)
    for ix in (0, 40, 95):
        assert s.suction_surface_y(ix) == top[ix]
    assert s.lattice_x_for_chord_station(0.5) == int(np.floor(xle + 0.5 * (xte - xle) + 0.5))
    nx, ny, nz = case.dims
    pv, pw = s.download_crossflow_plane(37)
    assert np.array_equal(pv, v.reshape(nz, ny, nx)[:, :, 37].reshape(-1))
    assert np.array_equal(pw, w.reshape(nz, ny, nx)[:, :, 37].reshape(-1))
    with pytest.raises(RuntimeError):
        s.download_crossflow_plane(nx)
    # no separation once the reversal is gone
    u2 = u.copy().reshape(nz, ny, nx)
    u2[u2 < 0] = 0.01
    s.restore_from_macroscopic(rho, u2.reshape(-1), v, w)
    assert s.separation_onset_xc() == -1.0
    s.close()


def test_readouts_on_developed_flow(gpu):
    say("case: delta99 / separation on a developed tunnel flow, 96x48x8 f32, 1500 steps")
    case = tunnel()
    s = cases.make_solver(case, "f32")
    s.step_n(1500)
    s.synchronize()
    rho, u, v, w = s.read_macro()
    got = s.extract_suction_delta99(STATIONS)
    want = orc.suction_delta99(u, v, case.flags, case.dims, STATIONS, case.scaling.chord_cells, case.scaling.u_lat)
    same_samples(got, want)
    say("  " + ", ".join(f"{o['x_c']:.2f}:{o['delta99_c']:.3f}{'' if o['valid'] else '*'}" for o in got))
    sep = s.separation_onset_xc()
    assert np.float32(sep) == np.float32(orc.separation_onset_xc(u, case.flags, case.dims, case.scaling.u_lat))
    say(f"  separation onset x/c {sep:.3f}")
    # the clean-foil reference moves the surface: a vane above the foil in the
    # live flags must not change the readouts
    vg = case.flags.copy().reshape(case.dims[2], case.dims[1], case.dims[0])
    top, xle, xte, _ = orc.surface_cache(case.flags, case.dims)
    xm = (xle + xte) // 2
    vg[:, top[xm] + 3:top[xm] + 5, xm:xm + 2] = fc.SOLID
    s.set_surface_reference(case.flags)
    s.apply_edited_flags(vg.reshape(-1))
    assert s.suction_surface_y(xm) == top[xm]
    s.set_surface_reference(None)
    assert s.suction_surface_y(xm) == top[xm] + 4
    s.close()


def test_band_circulation_and_audit(gpu):
    say("case: live band circulation and VG audit, Lamb-Oseen pair planted, tunnel 96x48x64 f32")
    case = tunnel(nz=64)
    nx, ny, nz = case.dims
    s = cases.make_solver(case, "f32")
    top, xle, xte, _ = orc.surface_cache(case.flags, case.dims)
    vgp = dict(type=1, x_c=0.2, height_c=0.1, length_h=3.0, beta_deg=16.0, pitch_c=0.3, gap_h=2.5, count=1)
    pxc = np.float32(vgp["x_c"]) + np.float32(vgp["height_c"]) * np.float32(6.0)
    ix = int(np.floor(xle + pxc * (xte - xle) + 0.5))
    ys = int(top[ix])
    rng = np.random.default_rng(2)
    u = np.full((nz, ny, nx), 0.05, np.float32)
    v = (rng.standard_normal((nz, ny, nx)) * 1e-6).astype(np.float32)
    w = (rng.standard_normal((nz, ny, nx)) * 1e-6).astype(np.float32)
    gamma, rc = 0.3, 2.5
    pv = np.zeros(ny * nz, np.float32)
    pw = np.zeros(ny * nz, np.float32)
    from test_bl import lamb_oseen
    lamb_oseen(pv, pw, ny, nz, ys + 6.0, nz / 2 - 8.0, +gamma, rc)
    lamb_oseen(pv, pw, ny, nz, ys + 6.0, nz / 2 + 8.0, -gamma, rc)
    v[:, :, ix] += pv.reshape(nz, ny)
    w[:, :, ix] += pw.reshape(nz, ny)
    s.restore_from_macroscopic(np.ones(u.size, np.float32), u.reshape(-1), v.reshape(-1), w.reshape(-1))
    y0, y1 = ys + 1, min(ny - 2, ys + 14)
    b = s.band_circulation_live(ix, y0, y1, nz // 2 - 14, nz // 2 + 14)
    pv2, pw2 = s.download_crossflow_plane(ix)
    ref = orc.band_circulation(pv2, pw2, ny, nz, y0, y1, nz // 2 - 14, nz // 2 + 14, f64=True)
    for k in ("zone", "ambient"):
        for a, r in zip(b[k], ref[k]):
            assert abs(a - r) <= 1e-6 * max(1.0, abs(r)), (k, a, r)
    assert (b["zone_cols"], b["ambient_cols"]) == (ref["zone_cols"], ref["ambient_cols"]) == (29, nz - 29)
    # the planted pair within the reference test's 3 %
    say(f"  zone +{b['zone'][0]:.4f} {b['zone'][1]:.4f} (planted +-{gamma})")
    assert abs(b["zone"][0] - gamma) < 0.03 * gamma and abs(-b["zone"][1] - gamma) < 0.03 * gamma
    # empty or inverted bands read zero
    assert s.band_circulation_live(ix, 10, 10, 0, nz - 1)["zone_cols"] == 0
    vg = fsolver.vg_params(type=vgp["type"], x_c=vgp["x_c"], height_c=vgp["height_c"], length_h=vgp["length_h"],
                      beta_deg=vgp["beta_deg"], pitch_c=vgp["pitch_c"], gap_h=vgp["gap_h"], count=vgp["count"])
    ue, d99 = 0.05, 4.0
    got = audit.audit_vg_vortex_strength(s, vg, case.scaling.chord_cells, ue, d99)
    want = orc.audit_vg(v.reshape(-1), w.reshape(-1), case.flags, case.dims, vgp, case.scaling.chord_cells, ue, d99)
    say(f"  audit: measured {got['gamma_measured']:.4f} predicted {got['gamma_predicted']:.4f} "
        f"ratio {got['ratio']:.3f} plane x/c {got['plane_xc']:.3f}")
    assert got["valid"] and want["valid"]
    assert np.float32(got["plane_xc"]) == np.float32(want["plane_xc"])
    assert got["gamma_predicted"] == pytest.approx(want["gamma_predicted"], rel=1e-6)
    assert got["gamma_measured"] == pytest.approx(want["gamma_measured"], rel=1e-5)
    assert got["gamma_measured"] == pytest.approx(gamma, rel=0.03)
    bad = audit.audit_vg_vortex_strength(s, vg, case.scaling.chord_cells, 0.0, d99)
    assert not bad["valid"]
    s.close()
