"""Boundary-layer readouts and the VG vortex audit on the CPU: the
reference's own known answers (tests/test_vgaudit.cpp: Lamb-Oseen pair
recovery within 3 %, zone / ambient split, Wendt spot value 0.2793 with
monotonicity, saturation and guards) for the library's host functions and
the oracle; the library's host integrals bit-identical to the oracle's FP32
restatement; and the oracle's delta99 / separation restatement pinned by
synthetic fields with closed-form answers."""
import numpy as np
import pytest

from foil_cfd_amd import audit
from oracle import oracle as O

PI = np.float32(3.14159265)


def lamb_oseen(v, w, ny, nz, y0, z0, gamma, rc):
    """Superpose a Lamb-Oseen vortex (circulation gamma, core radius rc) on a
    plane indexed y + ny*z: v_theta = gamma / (2 pi r) (1 - exp(-r^2 / rc^2))."""
    z, y = np.meshgrid(np.arange(nz, dtype=np.float32), np.arange(ny, dtype=np.float32), indexing="ij")
    dy, dz = y - np.float32(y0), z - np.float32(z0)
    r2 = np.maximum(dy * dy + dz * dz, np.float32(1e-6))
    swirl = np.float32(gamma) / (np.float32(2) * PI * r2) * (np.float32(1) - np.exp(-r2 / np.float32(rc * rc)))
    v += (-swirl * dz).reshape(-1)
    w += (swirl * dy).reshape(-1)


def pair_plane(ny=64, nz=64, gamma=0.8, rc=4.0):
    v = np.zeros(ny * nz, np.float32)
    w = np.zeros(ny * nz, np.float32)
    lamb_oseen(v, w, ny, nz, 24.0, 16.0, +gamma, rc)
    lamb_oseen(v, w, ny, nz, 24.0, 48.0, -gamma, rc)
    return v, w


def rel(a, b):
    return abs(a - b) / abs(b)


@pytest.mark.parametrize("impl", ["library", "oracle"])
def test_lamb_oseen_pair_recovery(impl):
    ny = nz = 64
    gamma = 0.8
    v, w = pair_plane()
    if impl == "library":
        full = audit.integrate_streamwise_circulation(v, w, ny, nz, 2, ny - 3)
        far = audit.integrate_streamwise_circulation(v, w, ny, nz, 44, ny - 3)
        both = audit.integrate_band_circulation(v, w, ny, nz, 2, ny - 3, 8, 56)
        half = audit.integrate_band_circulation(v, w, ny, nz, 2, ny - 3, 8, 24)
    else:
        b = O.band_circulation(v, w, ny, nz, 2, ny - 3, 0, nz - 1)
        full = b["zone"]
        far = O.band_circulation(v, w, ny, nz, 44, ny - 3, 0, nz - 1)["zone"]
        both = O.band_circulation(v, w, ny, nz, 2, ny - 3, 8, 56)
        half = O.band_circulation(v, w, ny, nz, 2, ny - 3, 8, 24)
    assert rel(full[0], gamma) < 0.03 and rel(-full[1], gamma) < 0.03
    assert far[0] < 0.05 * gamma and -far[1] < 0.05 * gamma
    assert rel(both["zone"][0], gamma) < 0.03 and rel(-both["zone"][1], gamma) < 0.03
    assert both["ambient"][0] < 0.03 * gamma and -both["ambient"][1] < 0.03 * gamma
    assert both["zone_cols"] == 49 and both["ambient_cols"] == nz - 49
    assert rel(half["zone"][0], gamma) < 0.03 and rel(-half["ambient"][1], gamma) < 0.03


def test_library_band_integral_bit_identical_to_fp32_restatement():
    rng = np.random.default_rng(5)
    ny, nz = 40, 24
    v = rng.standard_normal(ny * nz).astype(np.float32) * 0.01
    w = rng.standard_normal(ny * nz).astype(np.float32) * 0.01
    for y0, y1, z0, z1 in ((3, 30, 5, 15), (0, 39, 0, 23), (10, 11, 20, 40), (-4, 50, -3, 2), (12, 12, 0, 5),
                           (5, 9, 9, 4)):
        got = audit.integrate_band_circulation(v, w, ny, nz, y0, y1, z0, z1)
        want = O.band_circulation(v, w, ny, nz, y0, y1, z0, z1)
        assert [np.float32(x) for x in got["zone"]] == [np.float32(x) for x in want["zone"]]
        assert [np.float32(x) for x in got["ambient"]] == [np.float32(x) for x in want["ambient"]]
        assert (got["zone_cols"], got["ambient_cols"]) == (want["zone_cols"], want["ambient_cols"])
    # too few span planes: nothing
    assert audit.integrate_streamwise_circulation(v[:ny * 2], w[:ny * 2], ny, 2, 0, ny - 1) == (0.0, 0.0)


def test_wendt_prediction_known_answers():
    beta = 16.0 * 3.14159265 / 180.0
    for fn in (audit.wendt_circulation_lattice, O.wendt_circulation):
        g = fn(beta, 0.05, 8.0, 24.0, 10.0)
        assert abs(g - 0.2793) < 2e-3  # hand-computed spot value of the reference test
        assert fn(1.2 * beta, 0.05, 8, 24, 10) > g
        assert fn(beta, 0.05, 10, 24, 10) > g
        deep = fn(beta, 0.05, 8, 24, 8.0 / 5.0)
        deeper = fn(beta, 0.05, 8, 24, 8.0 / 50.0)
        assert rel(deep, deeper) < 0.01
        assert fn(beta, 0.0, 8, 24, 10) == 0.0 and fn(beta, 0.05, 0, 24, 10) == 0.0
    assert abs(audit.wendt_circulation_lattice(beta, 0.05, 8, 24, 10) - O.wendt_circulation(beta, 0.05, 8, 24, 10)) \
        < 1e-7


def test_speed_is_glibc_hypotf():
    rng = np.random.default_rng(9)
    a = (rng.standard_normal(200000) * np.exp(rng.uniform(-8, 2, 200000))).astype(np.float32)
    b = (rng.standard_normal(200000) * np.exp(rng.uniform(-8, 2, 200000))).astype(np.float32)
    assert np.array_equal(O.speed32(a, b), np.hypot(a, b))


# ---- oracle delta99 / separation pinned by synthetic fields --------------------
def plate(nx=64, ny=40, nz=4, ys=5, slope=0):
    """A solid slab up to row ys (+ slope * (x - x0) // 4 rise) for x in [8, 56)."""
    flags = np.zeros((nz, ny, nx), np.uint8)
    for x in range(8, 56):
        top = ys + (slope * (x - 8)) // 4
        flags[:, :top + 1, x] = O.SOLID_FLAG
    return flags.reshape(-1), (nx, ny, nz)


def profile_fields(flags, dims, ue=0.05, height=10.0, near=None):
    """u rising linearly from the wall to ue at `height` cells above each
    column's top solid row; v = 0."""
    nx, ny, nz = dims
    top, *_ = O.surface_cache(flags, dims)
    u = np.zeros((nz, ny, nx), np.float32)
    for x in range(nx):
        ys = top[x] if top[x] >= 0 else -1
        for y in range(ys + 1, ny):
            u[:, y, x] = ue * min(1.0, (y - ys) / height)
        if near is not None and top[x] >= 0:
            u[:, top[x] + 1, x] = near
    return u.reshape(-1), np.zeros_like(u).reshape(-1)


def test_oracle_delta99_linear_profile_closed_form():
    flags, dims = plate()
    u, v = profile_fields(flags, dims, ue=0.05, height=10.0)
    top, xle, xte, ok = O.surface_cache(flags, dims)
    assert ok and (xle, xte) == (8, 55) and top[20] == 5
    s = O.suction_delta99(u, v, flags, dims, [0.25, 0.5], chord_cells=48, u_lat=0.05)
    # first row with (y - ys) / 10 >= 0.99: y - ys = 10 -> 9.5 cells
    for o in s:
        assert o["valid"] and abs(o["ue_edge"] - 0.05) < 1e-9
        assert o["delta99_c"] == pytest.approx(9.5 / 48, abs=1e-7)
    # dead air: ue below 5 % of u_lat
    u2, v2 = profile_fields(flags, dims, ue=0.001)
    assert not any(o["valid"] for o in O.suction_delta99(u2, v2, flags, dims, [0.5], 48, 0.05))
    # reversed near-wall flow: thickness reported, station invalid
    u3, v3 = profile_fields(flags, dims, near=-0.01)
    o = O.suction_delta99(u3, v3, flags, dims, [0.5], 48, 0.05)[0]
    assert not o["valid"] and o["delta99_c"] == pytest.approx(9.5 / 48, abs=1e-7)
    # a station off the body (left of the leading edge after the clamp): no surface
    o = O.suction_delta99(u, v, flags, dims, [-0.5], 48, 0.05)[0]
    assert o["ue_edge"] == 0.0 and not o["valid"]


def test_oracle_delta99_slope_correction():
    flags, dims = plate(slope=2)  # rise of 2 rows every 4 columns: slope 0.5
    u, v = profile_fields(flags, dims, ue=0.05, height=10.0)
    o = O.suction_delta99(u, v, flags, dims, [0.5], chord_cells=48, u_lat=0.05)[0]
    top, *_ = O.surface_cache(flags, dims)
    ix = int(round(8 + 0.5 * 47))
    slope = (top[ix + 2] - top[ix - 2]) / 4.0
    assert slope == 0.5
    assert o["delta99_c"] == pytest.approx(9.5 / np.sqrt(1 + slope ** 2) / 48, rel=1e-6)


def test_oracle_separation_onset():
    flags, dims = plate()
    nx, ny, nz = dims
    top, xle, xte, _ = O.surface_cache(flags, dims)
    u = np.full((nz, ny, nx), 0.05, np.float32)
    assert O.separation_onset_xc(u.reshape(-1), flags, dims, 0.05) == -1.0
    # an isolated reversed column and a run of two before the real onset
    for x in (15, 20, 21):
        u[:, top[x] + 2, x] = -0.01
    for x in range(30, 40):
        u[:, top[x] + 2, x] = -0.01
    xc = O.separation_onset_xc(u.reshape(-1), flags, dims, 0.05)
    assert xc == pytest.approx((30 - 8) / 47, abs=1e-7)
    # weaker than 2 % of u_lat is jitter, not separation
    u[:, :, :] = 0.05
    for x in range(30, 40):
        u[:, top[x] + 2, x] = -0.0009
    assert O.separation_onset_xc(u.reshape(-1), flags, dims, 0.05) == -1.0
    # reversal inside the first 3 % of chord (stagnation region) is skipped
    u[:, :, :] = 0.05
    for x in (8, 9, 10, 11):
        u[:, top[x] + 2, x] = -0.01
    assert O.separation_onset_xc(u.reshape(-1), flags, dims, 0.05) == -1.0
