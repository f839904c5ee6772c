"""VG geometry + q-link cut fractions (reference src/geom/vg.cpp) through the
C ABI: the reference's tests/test_qlibb.cpp checks (cut 0.8 within 2 %, the
0.05 clamp, the two-node flag, Fluid-only origins, densify round trip) plus
geometric properties of the stamped vanes."""
import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd._abi import FcVaneSlab
from foil_cfd_amd.solver import (build_flags_with_vgs, build_vane_qlinks, surface_frame_at,
                                 vg_params)

Q = fc.Q
D = (16, 8, 4)


def idx(x, y, z, d=D):
    return x + d[0] * (y + d[1] * z)


def wall(d, wx):
    a = np.zeros(d[::-1], np.uint8)
    a[:, :, wx:] = fc.SOLID
    return a.reshape(-1), np.zeros(np.prod(d), np.uint8)


def xslab(plane_x, half):
    s = FcVaneSlab()
    s.center[:] = (plane_x + half, 50.0, 50.0)
    s.b_axis[:] = (1, 0, 0)
    s.n_axis[:] = (0, 1, 0)
    s.d_axis[:] = (0, 0, 1)
    s.half_thick, s.half_height, s.half_length = half, 1000.0, 1000.0
    return s


def q_of(qf, cell, q):
    b = qf[cell * Q + q]
    return b / 255.0 if b else -1.0


def test_exact_cut_fraction():
    a, c = wall(D, 10)
    qf, ff, n = build_vane_qlinks(D, a, c, [xslab(10.3, 0.5)])
    assert n > 0
    assert q_of(qf, idx(9, 4, 2), 2) == pytest.approx(0.8, rel=0.02)


def test_clamp_drops_small_cut():
    a, c = wall(D, 10)
    qf, ff, n = build_vane_qlinks(D, a, c, [xslab(9.51, 0.5)])
    assert q_of(qf, idx(9, 4, 2), 2) < 0


def test_second_node_flag():
    a = np.zeros(D[::-1], np.uint8)
    a[:, :, 10] = fc.SOLID
    qf, ff, n = build_vane_qlinks(D, a.reshape(-1), np.zeros(np.prod(D), np.uint8),
                                  [xslab(10.3, 0.5)])
    assert qf[idx(9, 4, 2) * Q + 2] != 0
    assert ff[idx(9, 4, 2)] & (1 << 2)


def test_origins_are_fluid():
    a, c = wall(D, 10)
    a = a.reshape(D[::-1])
    a[:, :, 10] = fc.INTERFACE
    qf, ff, n = build_vane_qlinks(D, a.reshape(-1), c, [xslab(11.3, 0.5)])
    cells = np.nonzero(qf.reshape(-1, Q).any(axis=1))[0]
    assert (a.reshape(-1)[cells] == fc.FLUID).all()


def test_densify_round_trip():
    a, c = wall(D, 10)
    qf, ff, n = build_vane_qlinks(D, a, c, [xslab(10.4, 0.5)])
    assert n > 0
    b = qf.reshape(-1, Q)
    assert (b[b != 0] / 255.0 >= 0.05).all()
    # every link decodes to ~0.9 (the +x wall link) or the diagonal cut
    assert np.count_nonzero(b) == n


def test_ray_obb_entry():
    s = xslab(10.3, 0.5)
    o = np.array([9.5, 4.5, 2.5], np.float32)
    d = np.array([1.0, 0.0, 0.0], np.float32)
    C = __import__("ctypes")
    t = fc.load().fc_ray_obb_entry(o.ctypes.data_as(C.POINTER(C.c_float)),
                                   d.ctypes.data_as(C.POINTER(C.c_float)), C.byref(s))
    assert t == pytest.approx(0.8, abs=1e-6)
    d2 = np.array([0.0, 1.0, 0.0], np.float32)  # parallel, outside the slab
    t2 = fc.load().fc_ray_obb_entry(o.ctypes.data_as(C.POINTER(C.c_float)),
                                    d2.ctypes.data_as(C.POINTER(C.c_float)), C.byref(s))
    assert t2 == -1.0


def test_surface_frame_on_naca():
    pts, le = fc.generate_naca4("2412")
    p, t, nrm = surface_frame_at(pts, le, 0.3, True)
    assert 0.29 < p[0] < 0.31 and p[1] > 0.07        # upper surface near max thickness
    assert t[0] > 0.99 and abs(np.dot(t, nrm)) < 1e-6 and nrm[1] > 0.99
    p2, _, n2 = surface_frame_at(pts, le, 0.3, False)
    assert p2[1] < 0 and n2[1] < -0.99
    assert surface_frame_at(pts, le, 1.5, True) is None


def config4_geometry():
    nx, ny, nz, nc = 256, 96, 120, 128
    pts, le = fc.generate_naca4("2412")
    layout = fc.make_layout(nx, ny, nz, nc)
    clean = fc.build_clean_foil_flags(pts, 10.0, layout)
    h = 0.008 * 2  # doubled at this half-size chord so vanes are ~2 cells
    vg = vg_params(fc.solver.VG_PAIR, x_c=0.2, height_c=h, length_h=2.5, beta_deg=16.0,
                   pitch_c=6 * h, gap_h=2.5, count=12)
    flags, slabs = build_flags_with_vgs(pts, le, 10.0, layout, [vg], clean)
    return (nx, ny, nz), clean, flags, slabs


def test_vg_stamping_properties():
    d, clean, flags, slabs = config4_geometry()
    assert len(slabs) == 24
    added = (flags == fc.SOLID) & (clean != fc.SOLID)
    assert added.sum() > 24 * 4
    # vanes only ever turn Fluid into Solid; faces untouched
    assert ((clean == fc.SOLID) <= (flags == fc.SOLID)).all()
    g = flags.reshape(d[2], d[1], d[0])
    assert (g[:, :, 0] == fc.INLET).all() and (g[:, 1:-1, -1] == fc.OUTLET).all()
    # counter-rotating pairs: mirrored yaw -> z centroids symmetric about mid-span
    zs = sorted(s.center[2] for s in slabs)
    assert np.allclose(np.array(zs) + np.array(zs[::-1]), d[2], atol=1e-3)
    # vane slabs sit on the suction side near x/c = 0.2 of the rotated foil
    xs = np.array([s.center[0] for s in slabs])
    assert np.all(np.abs(xs - xs.mean()) < 1.0)
    # every stamped vane cell is inside (or within a cell of) some slab
    zz, yy, xx = np.nonzero(added.reshape(d[2], d[1], d[0]))
    pts = np.stack([xx + 0.5, yy + 0.5, zz + 0.5], axis=1)
    best = np.full(len(pts), np.inf)
    for s in slabs:
        for shift in (-d[2], 0, d[2]):  # the array wraps the periodic span
            p = pts - np.array(s.center[:]) - np.array([0, 0, shift])
            e = [np.abs(p @ np.array(ax[:])) - h for ax, h in
                 ((s.n_axis, s.half_height), (s.d_axis, s.half_length), (s.b_axis, s.half_thick))]
            best = np.minimum(best, np.max(e, axis=0))
    assert (best < 1.0).all()


def test_vane_qlinks_on_config4_geometry():
    d, clean, flags, slabs = config4_geometry()
    qf, ff, n = build_vane_qlinks(d, flags, clean, slabs)
    assert n > 100
    b = qf.reshape(-1, Q)
    cells = np.nonzero(b.any(axis=1))[0]
    assert (flags[cells] == fc.FLUID).all()
    assert ((b[b != 0] / 255.0) >= 0.05 - 1e-6).all()
