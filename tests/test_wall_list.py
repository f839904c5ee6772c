"""Wall-model cell list (reference src/sim/wall_model.cpp) through the C ABI:
the reference's own checks (tests/test_wallmodel.cpp:44-284) -- flat-slab
masks, normals and sampling distance, 45-degree facet and sphere normal
quality, periodic z, determinism, vane exclusion, sampling fallback."""
import numpy as np

import foil_cfd_amd as fc
from foil_cfd_amd.solver import build_wall_cell_list

Q = fc.Q


def idx(d, x, y, z):
    return x + d[0] * (y + d[1] * z)


def err_deg(l, i, e):
    n = np.array([l["nx"][i], l["ny"][i], l["nz"][i]], np.float64)
    e = np.asarray(e, np.float64)
    c = np.clip(n @ e / np.linalg.norm(e), -1, 1)
    return np.degrees(np.arccos(c))


def xyz(d, c):
    return c % d[0], (c // d[0]) % d[1], c // (d[0] * d[1])


def test_flat_slab():
    d = (32, 32, 8)
    f = np.zeros(d[::-1], np.uint8)
    f[:, :16, :] = fc.SOLID
    l, st = build_wall_cell_list(d, f)
    assert st == {"listed": 32 * 8, "excluded_vg": 0, "degenerate": 0}
    want = sum(1 << q for q in (3, 7, 11, 14, 17))
    for i, c in enumerate(l["cell"]):
        x, y, z = xyz(d, c)
        if x < 2 or x >= d[0] - 2:
            continue
        assert l["mask"][i] == want
        assert err_deg(l, i, (0, 1, 0)) < 0.5
        assert abs(l["ys"][i] - 2.5) < 1e-5
        assert l["sample"][i] == c + 2 * d[0]


def test_wedge45():
    d = (64, 64, 8)
    z, y, x = np.meshgrid(range(8), range(64), range(64), indexing="ij")
    f = np.where(x + y < 32, fc.SOLID, fc.FLUID).astype(np.uint8)
    l, _ = build_wall_cell_list(d, f)
    worst = 0.0
    for i, c in enumerate(l["cell"]):
        x, y, _ = xyz(d, c)
        if x < 4 or y < 4 or x >= 30 or y >= 30:
            continue
        worst = max(worst, err_deg(l, i, (1, 1, 0)))
    assert len(l["cell"]) > 0 and worst < 5.0


def sphere(d, c, r):
    z, y, x = np.meshgrid(range(d[2]), range(d[1]), range(d[0]), indexing="ij")
    x = x.astype(np.float32); y = y.astype(np.float32); z = z.astype(np.float32)
    inside = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 < np.float32(r * r)
    return np.where(inside, fc.SOLID, fc.FLUID).astype(np.uint8)


def test_sphere_normals():
    d = (64, 64, 64)
    f = sphere(d, (32, 32, 32), 14)
    l, _ = build_wall_cell_list(d, f)
    errs = [err_deg(l, i, np.array(xyz(d, c), np.float64) - 32) for i, c in enumerate(l["cell"])]
    assert max(errs) < 15.0 and np.mean(errs) < 8.0


def test_periodic_z():
    d = (16, 16, 8)
    f = np.zeros(d[::-1], np.uint8)
    f[0] = fc.SOLID
    l, _ = build_wall_cell_list(d, f)
    assert len(l["cell"]) == 2 * 16 * 16
    for i, c in enumerate(l["cell"]):
        x, y, z = xyz(d, c)
        assert z in (1, d[2] - 1)
        if x < 2 or x >= 14 or y < 2 or y >= 14:
            continue
        assert err_deg(l, i, (0, 0, 1 if z == 1 else -1)) < 0.5


def test_determinism():
    d = (48, 48, 16)
    f = sphere(d, (24, 24, 8), 9)
    a, _ = build_wall_cell_list(d, f)
    b, _ = build_wall_cell_list(d, f)
    for k in a:
        assert np.array_equal(a[k], b[k])


def test_vane_exclusion():
    d = (64, 32, 8)
    clean = np.zeros(d[::-1], np.uint8)
    clean[:, :16, :] = fc.SOLID
    active = clean.copy()
    active[3:5, 16:24, 32] = fc.SOLID
    l, st = build_wall_cell_list(d, active, clean)
    assert st["excluded_vg"] > 0 and st["listed"] > 0
    cl = clean.reshape(-1)
    for i, c in enumerate(l["cell"]):
        x, y, z = xyz(d, c)
        for q in range(1, Q):
            if l["mask"][i] & (1 << q):
                n = idx(d, x - fc.CX[q], y - fc.CY[q], (z - fc.CZ[q]) % d[2])
                assert cl[n] == fc.SOLID


def test_sample_fallback():
    d = (32, 32, 8)
    f = np.zeros(d[::-1], np.uint8)
    f[:, :16, :] = fc.SOLID
    f[:, 18:20, :] = fc.SOLID
    l, _ = build_wall_cell_list(d, f)
    for i, c in enumerate(l["cell"]):
        x, y, _ = xyz(d, c)
        if y != 16 or x < 2 or x >= 30:
            continue
        assert l["sample"][i] == c + d[0]
        assert abs(l["ys"][i] - 1.5) < 1e-5


def test_naca_tunnel_list_nonempty():
    pts, _ = fc.generate_naca4("2412")
    layout = fc.make_layout(192, 64, 4, 64)
    f = fc.build_clean_foil_flags(pts, 8.0, layout)
    l, st = build_wall_cell_list((192, 64, 4), f)
    assert st["listed"] > 100 and st["degenerate"] < 0.05 * st["listed"]
    # every listed cell is fluid with at least one solid pull link
    assert (f[l["cell"]] == fc.FLUID).all() and (l["mask"] != 0).all()
