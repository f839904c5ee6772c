"""The oracle's restatement of the mesh-sequencing upsample
(upsampleMacroKernel, lbm_refine.cu:437-501) pinned by what trilinear
interpolation fixes: a constant field stays exact, equal grids copy bit for
bit, a field linear in x / y / z is reproduced at the destination centres
mapped by (i + 1/2) s / d - 1/2 (an off-by-half or transposed axis fails),
z wraps periodically, and x / y clamp to the border sample."""
import numpy as np

from oracle import oracle as O


def grid(d):
    nx, ny, nz = d
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return x.reshape(-1), y.reshape(-1), z.reshape(-1)


def test_constant_and_identity():
    s = (7, 5, 4)
    n = 7 * 5 * 4
    rng = np.random.default_rng(1)
    fields = [np.full(n, 1.25, np.float32), rng.standard_normal(n).astype(np.float32),
              rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)]
    out = O.upsample_macro(fields, s, (28, 20, 16))
    assert np.all(out[0] == np.float32(1.25))
    same = O.upsample_macro(fields, s, s)
    for a, b in zip(same, fields):
        assert np.array_equal(a, b)


def test_linear_field_reproduced_at_mapped_centres():
    s, d = (8, 6, 5), (24, 18, 15)
    sx, sy, sz = grid(s)
    a, b = 0.01, -0.02
    f_x = (0.1 + a * sx).astype(np.float32)
    f_y = (0.3 + b * sy).astype(np.float32)
    out = O.upsample_macro([f_x, f_y, f_x, f_y], s, d)
    dx, dy, dz = grid(d)
    gx = (dx + 0.5) * s[0] / d[0] - 0.5
    gy = (dy + 0.5) * s[1] / d[1] - 0.5
    inx = (gx >= 0) & (gx <= s[0] - 1)
    iny = (gy >= 0) & (gy <= s[1] - 1)
    assert np.allclose(out[0][inx], 0.1 + a * gx[inx], atol=2e-7)
    assert np.allclose(out[1][iny], 0.3 + b * gy[iny], atol=2e-7)
    # clamped outside the first / last source centre: nearest border value
    assert np.allclose(out[0][gx < 0], 0.1, atol=1e-7)
    assert np.allclose(out[0][gx > s[0] - 1], 0.1 + a * (s[0] - 1), atol=2e-7)


def test_z_periodic_wrap():
    s, d = (4, 4, 6), (4, 4, 24)
    _, _, sz = grid(s)
    k = 2 * np.pi / s[2]
    f = np.cos(k * sz).astype(np.float32)
    out = O.upsample_macro([f, f, f, f], s, d)[0]
    _, _, dz = grid(d)
    gz = (dz + 0.5) * s[2] / d[2] - 0.5
    # linear interpolation of the periodic samples, wrapping the last interval
    z0 = np.floor(gz).astype(int)
    t = gz - z0
    want = (1 - t) * np.cos(k * (z0 % s[2])) + t * np.cos(k * ((z0 + 1) % s[2]))
    assert np.allclose(out, want, atol=3e-7)
    assert np.all(out[dz == 0] != f[sz == 0][0])  # z = 0 blends with the last source plane
