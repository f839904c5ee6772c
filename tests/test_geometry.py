"""Flag builders (reference src/geom/airfoil.cpp, src/geom/voxelizer.cpp)
through the C ABI: the reference's own known answers
(tests/test_voxelizer.cpp, tests/test_airfoil.cpp:202-232), bit-exact NACA
points against the reference's compiled generator (tests/golden/naca_ref.npz)
and an independent brute-force point-in-polygon classification of every
voxelized cell centre."""
import os

import numpy as np
import pytest

import foil_cfd_amd as fc

GOLD = os.path.join(os.path.dirname(__file__), "golden", "naca_ref.npz")


def idx(d, x, y, z):
    return x + d[0] * (y + d[1] * z)


def diamond():
    corners = [(1.0, 0.0), (0.5, 0.25), (0.0, 0.0), (0.5, -0.25), (1.0, 0.0)]
    pts = []
    for e in range(4):
        a = np.array(corners[e], dtype=np.float32)
        b = np.array(corners[e + 1], dtype=np.float32)
        for s in range(6):
            t = np.float32(s) / np.float32(6)
            pts.append(a + (b - a) * t)
    pts.append(np.array(corners[4], dtype=np.float32))
    return np.array(pts, dtype=np.float32)


# ---- tests/test_voxelizer.cpp ----------------------------------------------
def test_boundary_flags_exact():
    d = (8, 6, 3)
    f = fc.build_boundary_flags(*d)
    for z in range(d[2]):
        for y in range(d[1]):
            for x in range(d[0]):
                e = fc.FLUID
                if y == 0:
                    e = fc.SLIP_BOTTOM
                if y == d[1] - 1:
                    e = fc.SLIP_TOP
                if x == 0:
                    e = fc.INLET
                if x == d[0] - 1 and y not in (0, d[1] - 1):
                    e = fc.OUTLET
                assert f[idx(d, x, y, z)] == e, (x, y, z)


def diamond_s(px, py, cx=13.6, cy=12.0, ax=8.0, ay=4.0):
    return abs(px - cx) / ax + abs(py - cy) / ay


def test_diamond_parity():
    d = (32, 24, 3)
    layout = fc.make_layout(*d, chord_cells=16)
    f = fc.build_boundary_flags(*d)
    fc.voxelize_airfoil(diamond(), 0.0, layout, f)
    solid = 0
    for j in range(1, d[1] - 1):
        for i in range(1, d[0] - 1):
            s = f[idx(d, i, j, 0)] == fc.SOLID
            solid += s
            sa, sb = diamond_s(i, j), diamond_s(i + 0.5, j + 0.5)
            if sa <= 0.75 and sb <= 0.75:
                assert s, (i, j)
            if sa >= 1.25 and sb >= 1.25:
                assert not s, (i, j)
    assert 40 <= solid <= 90
    g = f.reshape(d[2], d[1], d[0])
    assert all((g[z] == g[0]).all() for z in range(1, d[2]))
    assert f[idx(d, 0, 12, 0)] == fc.INLET


def test_trailing_edge_closure():
    d = (8, 8, 2)
    f = np.zeros(8 * 8 * 2, dtype=np.uint8)
    for z in range(d[2]):
        f[idx(d, 3, 2, z)] = fc.SOLID
        f[idx(d, 3, 4, z)] = fc.SOLID
        f[idx(d, 6, 2, z)] = fc.SOLID
    assert fc.close_trailing_edge_gaps(*d, f) == 2
    for z in range(d[2]):
        assert f[idx(d, 3, 3, z)] == fc.SOLID
        assert f[idx(d, 6, 1, z)] == fc.FLUID
        assert f[idx(d, 6, 3, z)] == fc.FLUID
        assert f[idx(d, 3, 2, z)] == fc.SOLID and f[idx(d, 3, 4, z)] == fc.SOLID


@pytest.mark.parametrize("aoa", [25.0, -5.0])
def test_aoa_margins(aoa):
    nc = 128
    d = (3 * nc, nc + nc // 4, 8)
    layout = fc.make_layout(*d, chord_cells=nc)
    f = fc.build_boundary_flags(*d)
    fc.voxelize_airfoil(diamond(), aoa, layout, f)
    g = f.reshape(d[2], d[1], d[0])[0]
    ys, xs = np.nonzero(g == fc.SOLID)
    assert xs.size > 0
    k = 8
    assert ys.min() >= k and ys.max() < d[1] - k
    assert xs.min() >= k and xs.max() < d[0] - k
    box = fc.derive_patch_box(*d, f, nc // 5, nc // 2, nc // 10, nc // 5)
    assert (box.x1 - box.x0) > 8 and (box.y1 - box.y0) > 8


def test_patch_local_anchor_above_chord():
    d = (32, 24, 3)
    coarse = fc.make_layout(*d, chord_cells=16)
    box = fc._abi.FcPatchBox(10, 13, 20, 22)
    fine = fc.make_fine_layout(coarse, box, 2)
    assert fine.anchor_absolute == 1
    assert abs(fine.anchor_x_cells - (-0.8)) < 1e-4
    assert abs(fine.anchor_y_cells - (-2.0)) < 1e-4
    fd = (fine.dims.nx, fine.dims.ny, fine.dims.nz)
    f = np.zeros(fd[0] * fd[1] * fd[2], dtype=np.uint8)
    fc.voxelize_airfoil(diamond(), 0.0, fine, f)
    inside = 0
    for j in range(1, fd[1] - 1):
        for i in range(1, fd[0] - 1):
            s = f[idx(fd, i, j, 0)] == fc.SOLID
            sa = diamond_s(i, j, 7.2, -2.0, 16.0, 8.0)
            sb = diamond_s(i + 0.5, j + 0.5, 7.2, -2.0, 16.0, 8.0)
            if sa <= 0.75 and sb <= 0.75:
                inside += 1
                assert s
            if sa >= 1.25 and sb >= 1.25:
                assert not s
    assert inside > 10


# ---- tests/test_airfoil.cpp:202-232 ------------------------------------------
def surface_y(pts, le, x, upper):
    seg = pts[: le + 1][::-1] if upper else pts[le:]
    for a, b in zip(seg[:-1], seg[1:]):
        lo, hi = min(a[0], b[0]), max(a[0], b[0])
        if x < lo or x > hi or hi - lo < 1e-9:
            continue
        t = (x - a[0]) / (b[0] - a[0])
        return a[1] + (b[1] - a[1]) * t
    return 0.0


def test_naca2412_shape():
    pts, le = fc.generate_naca4("2412")
    assert len(pts) >= 150
    assert pts[le][0] == 0.0 and pts[le][1] == 0.0
    max_t = max_t_at = max_m = max_m_at = 0.0
    for x in np.arange(0.02, 0.98, 0.005):
        yu, yl = surface_y(pts, le, x, True), surface_y(pts, le, x, False)
        t, m = yu - yl, 0.5 * (yu + yl)
        if t > max_t:
            max_t, max_t_at = t, x
        if m > max_m:
            max_m, max_m_at = m, x
    assert 0.113 < max_t < 0.127 and 0.24 < max_t_at < 0.36
    assert 0.017 < max_m < 0.023 and 0.34 < max_m_at < 0.46
    for bad in ("241", "24x2", "24122"):
        with pytest.raises(ValueError):
            fc.generate_naca4(bad)


# ---- bit-exact against the reference's compiled generator --------------------
def test_naca_bit_exact_vs_reference():
    gold = np.load(GOLD)
    for key in gold.files:
        if key.endswith("_le"):
            continue
        digits, pps = key[4:8], int(key.split("_")[1])
        pts, le = fc.generate_naca4(digits, pps)
        assert pts.tobytes() == gold[key].tobytes(), key
        assert le == int(gold[key + "_le"][0])


# ---- independent classification of every voxelized cell ---------------------
def brute_inside(poly, px, py):
    """Even-odd ray cast in FP64 for points (px, py)."""
    inside = np.zeros(px.shape, dtype=bool)
    n = len(poly)
    for i in range(n):
        ax, ay = poly[i]
        bx, by = poly[(i + 1) % n]
        cond = (ay > py) != (by > py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xc = ax + (py - ay) * (bx - ax) / (by - ay)
        inside ^= cond & (px < xc)
    return inside


@pytest.mark.parametrize("aoa", [0.0, 4.0, 8.0, 12.0])
def test_naca_voxelization_matches_point_in_polygon(aoa):
    nx, ny, nz, nc = 192, 96, 2, 64
    pts, _ = fc.generate_naca4("2412")
    layout = fc.make_layout(nx, ny, nz, nc)
    f = fc.build_clean_foil_flags(pts, aoa, layout)
    # lattice polygon in FP64 (same transform)
    a = np.deg2rad(aoa)
    p = pts.astype(np.float64) - [0.25, 0.0]
    q = np.stack([p[:, 0] * np.cos(-a) - p[:, 1] * np.sin(-a),
                  p[:, 0] * np.sin(-a) + p[:, 1] * np.cos(-a)], axis=1)
    poly = np.stack([0.3 * nx + q[:, 0] * nc, 0.5 * ny + q[:, 1] * nc], axis=1)
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    inside = brute_inside(poly, xx + 0.5, yy + 0.5)
    solid = f.reshape(nz, ny, nx)[0] == fc.SOLID
    interior = np.zeros_like(solid)
    interior[1:-1, 1:-1] = True
    # Cells more than 1e-3 cell from the outline must classify identically;
    # the only allowed extra solids are TE-gap closures (Fluid between two
    # solids in y).
    diff = (inside != solid) & interior
    ys, xs = np.nonzero(diff)
    for y, x in zip(ys, xs):
        closed = solid[y, x] and not inside[y, x] and solid[y - 1, x] and solid[y + 1, x]
        near = min(np.hypot(poly[:, 0] - (x + 0.5), poly[:, 1] - (y + 0.5)).min(), 1.0) < 0.6
        assert closed or near, (x, y, aoa)
    assert diff.sum() <= 0.02 * solid.sum()
    g = f.reshape(nz, ny, nx)
    assert (g[1] == g[0]).all()


def test_config_flag_fields_shape():
    pts, _ = fc.generate_naca4("2412")
    layout = fc.make_layout(384, 128, 8, 128)
    f = fc.build_clean_foil_flags(pts, 8.0, layout)
    g = f.reshape(8, 128, 384)
    assert (g[:, 0, 1:] == fc.SLIP_BOTTOM).all() and (g[:, -1, 1:] == fc.SLIP_TOP).all()
    assert (g[:, :, 0] == fc.INLET).all()
    assert (g[:, 1:-1, -1] == fc.OUTLET).all()
    # chord 128 at 12% thickness: solid area ~ 0.082 c^2 per slice
    area = (g[0] == fc.SOLID).sum()
    assert 0.07 * 128 ** 2 < area < 0.095 * 128 ** 2
