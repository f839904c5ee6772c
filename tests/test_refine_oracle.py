"""The oracle's two-level refinement restatement (lbm_refine.cu, stepN
sequencing lbm_solver.cpp:1172-1290) pinned by closed forms and invariants:
the Latt-Chopard projection, the fill's coordinate mapping and time blend,
the Dupuis-Chopard rescale round trip, the restriction's child average and
edge blend, and the uniform-freestream invariance of the whole coupled step
(the reference's m3 gate pattern)."""
import numpy as np
import pytest

import foil_cfd_amd as fc
from oracle import oracle as orc

Q = fc.Q
CX, CY, CZ, W = fc.CX, fc.CY, fc.CZ, fc.W


def hermite_fneq(pi):
    """fneq_q = w_q / (2 cs^4) (c_q c_q - cs^2 I) : Pi for a symmetric Pi."""
    out = np.empty(Q)
    c = np.stack([CX, CY, CZ], 1).astype(float)
    for q in range(Q):
        qq = np.outer(c[q], c[q]) - np.eye(3) / 3.0
        out[q] = 4.5 * W[q] * np.sum(qq * pi)
    return out


def test_regularize_projection_properties():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 3)) * 1e-3
    pi = a + a.T
    h = hermite_fneq(pi)
    # a pure second-order (stress) deviation is reproduced exactly
    assert np.allclose(orc.regularize_fneq(h), h, atol=1e-15)
    # arbitrary input: the result carries no mass / momentum, is idempotent,
    # and keeps the input's second moment (the stress) exactly
    x = rng.normal(size=Q) * 1e-3
    r = orc.regularize_fneq(x)
    assert abs(r.sum()) < 1e-15
    for cvec in (CX, CY, CZ):
        assert abs((cvec * r).sum()) < 1e-15
    assert np.allclose(orc.regularize_fneq(r), r, atol=1e-16)
    for ca, cb in ((CX, CX), (CY, CY), (CZ, CZ), (CX, CY), (CX, CZ), (CY, CZ)):
        assert abs((ca * cb * r).sum() - (ca * cb * x).sum()) < 1e-15
    # the non-hydrodynamic ghost (c_x^2 c_y^2 pattern minus its stress part) is removed
    ghost = np.where((CX != 0) & (CY != 0), 1.0, 0.0) * 1e-3
    assert np.abs(orc.regularize_fneq(ghost) - orc.regularize_fneq(orc.regularize_fneq(ghost))).max() < 1e-16


def small_patch(cdims=(24, 20, 4), box=(6, 5, 18, 15), m=2):
    nx, ny, nz = cdims
    cflags = np.zeros(nx * ny * nz, np.uint8)
    fd = orc.fine_dims(box, cdims, m)
    fflags = np.zeros(fd[0] * fd[1] * fd[2], np.uint8)
    fc.stamp_interface_shell(*fd, fflags)
    return cflags, fflags, fd


def test_fill_maps_fine_centres_linearly():
    """A coarse rest field with density linear in x and y: the full-volume fill
    gives w_q rho(x_c, y_c) at every fine cell centre, x_c = x0 + (xf + 1/2)/m
    - 1/2 in coarse cell-centre coordinates (equilibrium at rest is linear in
    rho, so interpolation and the feq / fneq split are exact)."""
    cdims, box, m = (24, 20, 4), (6, 5, 18, 15), 2
    cflags, fflags, fd = small_patch(cdims, box, m)
    nx, ny, nz = cdims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rho = (1.0 + 1e-3 * x + 2e-3 * y).reshape(-1)
    cf = np.stack([W[q] * rho for q in range(Q)]).reshape(-1)
    ff = np.zeros(Q * fd[0] * fd[1] * fd[2])
    orc.fine_fill(cf, cf, cflags, cdims, fflags, ff, box, m, 0.0, 1.0, full_volume=True)
    zf, yf, xf = np.meshgrid(np.arange(fd[2]), np.arange(fd[1]), np.arange(fd[0]), indexing="ij")
    xc = box[0] + (xf + 0.5) / m - 0.5
    yc = box[1] + (yf + 0.5) / m - 0.5
    want = (1.0 + 1e-3 * xc + 2e-3 * yc).reshape(-1)
    got = ff.reshape(Q, -1)
    for q in range(Q):
        assert np.abs(got[q] - W[q] * want).max() < 1e-14


def test_fill_time_blend_and_shell_only():
    """Shell-only fill between two uniform rest states of density r0 and r1 at
    weight k/m writes w_q (r0 + k/m (r1 - r0)) on Interface cells only."""
    cdims, box, m = (24, 20, 4), (6, 5, 18, 15), 2
    cflags, fflags, fd = small_patch(cdims, box, m)
    n = cdims[0] * cdims[1] * cdims[2]
    c0 = np.concatenate([np.full(n, W[q] * 1.00) for q in range(Q)])
    c1 = np.concatenate([np.full(n, W[q] * 1.02) for q in range(Q)])
    ff = np.full(Q * fd[0] * fd[1] * fd[2], -1.0)
    orc.fine_fill(c0, c1, cflags, cdims, fflags, ff, box, m, 0.5, 1.0)
    got = ff.reshape(Q, -1)
    shell = fflags == fc.INTERFACE
    assert shell.sum() > 0 and (~shell).sum() > 0
    for q in range(Q):
        assert np.abs(got[q, shell] - W[q] * 1.01).max() < 1e-15
        assert np.all(got[q, ~shell] == -1.0)


def test_uniform_stress_round_trips_through_fill_and_restrict():
    """Uniform rho, u and a uniform Hermite stress: the fill rescales fneq by
    s = tau_f / (m tau_c), the restriction by 1/s, so the coarse interior is
    reproduced exactly (child mean, 19-point filter, regularisation and blend
    are all identities on a uniform field)."""
    cdims, box, m = (32, 28, 4), (6, 5, 26, 23), 2
    cflags, fflags, fd = small_patch(cdims, box, m)
    n = cdims[0] * cdims[1] * cdims[2]
    feq = orc.equilibrium(1.01, 0.03, -0.01, 0.005)
    rng = np.random.default_rng(5)
    a = rng.normal(size=(3, 3)) * 1e-4
    fneq = hermite_fneq(a + a.T)
    cell = feq + fneq
    cf = np.repeat(cell[:, None], n, axis=1).reshape(-1)
    tau_c = 0.53
    tau_f = orc.fine_tau_for(tau_c, m)
    s = tau_f / (m * tau_c)
    ff = np.zeros(Q * fd[0] * fd[1] * fd[2])
    orc.fine_fill(cf, cf, cflags, cdims, fflags, ff, box, m, 0.0, s, full_volume=True)
    g = ff.reshape(Q, -1)
    assert np.abs(g - (feq + s * fneq)[:, None]).max() < 1e-15
    out = np.zeros_like(cf)
    orc.fine_restrict(ff, fflags, out, cflags, cdims, box, m, 1.0 / s)
    got = out.reshape(Q, cdims[2], cdims[1], cdims[0])
    x0, y0, x1, y1 = box
    inner = got[:, :, y0 + 2:y1 - 2, x0 + 2:x1 - 2]
    # existing coarse value was 0: the blend ramp scales the edge cells, the
    # cells >= 3 inside the restricted box carry the full restricted value
    core = got[:, :, y0 + 5:y1 - 5, x0 + 5:x1 - 5]
    assert np.abs(core - cell[:, None, None, None]).max() < 1e-14
    edge = got[:, :, y0 + 2, x0 + 6]
    assert np.abs(edge - 0.25 * cell[:, None]).max() < 1e-14
    # outside the restricted box nothing is written
    assert np.all(got[:, :, :y0 + 2, :] == 0.0) and np.all(got[:, :, :, :x0 + 2] == 0.0)
    assert inner.shape[2] > 0


def test_restriction_averages_children_at_coarse_centre():
    """A fine rest field with density linear in x, y and z: each restricted
    coarse cell receives the mean of its m^3 children, which for a linear field
    is the value at the coarse cell centre (blend weight 1 in the core)."""
    cdims, box, m = (24, 20, 8), (6, 5, 20, 17), 2
    cflags, fflags, fd = small_patch(cdims, box, m)
    zf, yf, xf = np.meshgrid(np.arange(fd[2]), np.arange(fd[1]), np.arange(fd[0]), indexing="ij")
    rho = (1.0 + 1e-3 * xf + 2e-3 * yf + 5e-4 * zf).reshape(-1)
    ff = np.stack([W[q] * rho for q in range(Q)]).reshape(-1)
    n = cdims[0] * cdims[1] * cdims[2]
    out = np.zeros(Q * n)
    orc.fine_restrict(ff, fflags, out, cflags, cdims, box, m, 1.0)
    got = out.reshape(Q, cdims[2], cdims[1], cdims[0])
    x0, y0, x1, y1 = box
    for zc in (1, 5):
        for yc in range(y0 + 5, y1 - 5):
            for xc in range(x0 + 5, x1 - 5):
                fxm = m * (xc - x0) + 0.5 * (m - 1)
                fym = m * (yc - y0) + 0.5 * (m - 1)
                fzm = m * zc + 0.5 * (m - 1)
                r = 1.0 + 1e-3 * fxm + 2e-3 * fym + 5e-4 * fzm
                assert abs(got[0, zc, yc, xc] - W[0] * r) < 1e-14


def test_fine_tau_composition():
    for tau in (0.5005, 0.52, 0.6, 0.9):
        a = orc.fine_tau_for(orc.fine_tau_for(tau, 2), 2)
        b = orc.fine_tau_for(tau, 4)
        assert a == pytest.approx(b, rel=1e-6)
        assert fc.refined_scaling(fc.scaling_for_tau(tau, 0.08, 64), 2).tau == pytest.approx(
            orc.fine_tau_for(tau, 2), rel=1e-6)


def test_uniform_freestream_stays_uniform_through_coupled_steps():
    """Tunnel without a body (inlet / zero-gradient outlet / slip walls) at the
    uniform equilibrium of the inlet speed, with a factor-2 patch: 20 coupled
    steps (coarse + 2 fine sub-steps + restriction each) leave both levels
    uniform to rounding (the reference's m3 uniform-freestream gate)."""
    cdims, box, m = (40, 24, 4), (12, 6, 28, 18), 2
    nx, ny, nz = cdims
    cflags = fc.build_boundary_flags(*cdims)
    fd = orc.fine_dims(box, cdims, m)
    fflags = np.zeros(fd[0] * fd[1] * fd[2], np.uint8)
    fc.stamp_interface_shell(*fd, fflags)
    u0 = 0.05
    feq = orc.equilibrium(1.0, u0, 0.0, 0.0)
    n = nx * ny * nz
    cf = np.repeat(feq[:, None], n, axis=1).reshape(-1).copy()
    tau = 0.56
    ff = orc.seed_fine(cf, cflags, cdims, fflags, box, m, tau)
    steps = 20
    rho, u, v, w = orc.run_refined(cf, ff, cflags, fflags, cdims, box, m,
                                   np.full(steps, tau), np.full(steps, u0), cs=0.1)
    fluid = cflags == fc.FLUID
    assert np.abs(u[fluid] - u0).max() < 1e-13
    assert np.abs(v[fluid]).max() < 1e-13 and np.abs(rho[fluid] - 1.0).max() < 1e-13
    g = ff.reshape(Q, -1)[:, fflags == fc.FLUID]
    assert np.abs(g - feq[:, None]).max() < 1e-14


def test_uniform_freestream_stays_uniform_through_three_levels():
    """The nested third level (reference m4_finer's 3-level uniform-freestream
    gate): coarse tunnel without a body, a factor-2 fine patch and inside it a
    factor-2 finer patch; 10 coupled steps (coarse + 2 x (fine + 2 finer
    sub-steps + finer restriction) + fine restriction) keep all three levels
    at the uniform equilibrium to rounding."""
    cdims, box, m = (40, 24, 4), (10, 5, 30, 19), 2
    nx, ny, nz = cdims
    cflags = fc.build_boundary_flags(*cdims)
    fd = orc.fine_dims(box, cdims, m)
    fflags = np.zeros(fd[0] * fd[1] * fd[2], np.uint8)
    fc.stamp_interface_shell(*fd, fflags)
    gbox, m2 = (10, 8, 30, 20), 2  # in fine cells, 3+ cells inside the fine shell
    gd = orc.fine_dims(gbox, fd, m2)
    gflags = np.zeros(gd[0] * gd[1] * gd[2], np.uint8)
    fc.stamp_interface_shell(*gd, gflags)
    u0, tau = 0.05, 0.56
    feq = orc.equilibrium(1.0, u0, 0.0, 0.0)
    cf = np.repeat(feq[:, None], nx * ny * nz, axis=1).reshape(-1).copy()
    ff = orc.seed_fine(cf, cflags, cdims, fflags, box, m, tau)
    gf = orc.seed_fine(ff, fflags, fd, gflags, gbox, m2, orc.fine_tau_for(tau, m))
    steps = 10
    rho, u, v, w = orc.run_refined3(cf, ff, gf, cflags, fflags, gflags, cdims, box, m, gbox, m2,
                                    np.full(steps, tau), np.full(steps, u0), cs=0.1)
    fluid = cflags == fc.FLUID
    assert np.abs(u[fluid] - u0).max() < 1e-13 and np.abs(rho[fluid] - 1.0).max() < 1e-13
    for arr, fl in ((ff, fflags), (gf, gflags)):
        g = arr.reshape(Q, -1)[:, fl == fc.FLUID]
        assert g.shape[1] > 0 and np.abs(g - feq[:, None]).max() < 1e-13


def test_refined_run_slip_wiring():
    """The slip-aware refined runner: zero slip fields equal the plain run bit
    for bit; a coarse slip field changes the coarse step outside the patch
    exactly as the single-level step with that field does."""
    import numpy as np
    from oracle import oracle as O
    dims, box, m = (48, 32, 4), (14, 10, 30, 22), 2
    nx, ny, nz = dims
    n = nx * ny * nz
    import foil_cfd_amd as fc
    flags = fc.build_boundary_flags(*dims).reshape(nz, ny, nx)
    flags[:, 4:6, 4:8] = 1  # a small solid block outside the patch
    flags = flags.reshape(-1)
    fd = O.fine_dims(box, dims, m)
    fine = np.zeros(fd[0] * fd[1] * fd[2], np.uint8)
    fc.stamp_interface_shell(*fd, fine)
    rng = np.random.default_rng(0)
    f0 = O.init_equilibrium(flags, dims, 0.05, f64=True)
    f0 += 1e-4 * rng.standard_normal(f0.size)
    g0 = O.init_equilibrium(fine, fd, 0.05, f64=True)
    taus, uins = np.full(2, 0.56), np.full(2, 0.05)
    z3 = [np.zeros(n, np.uint16)] * 3
    zf = [np.zeros(fd[0] * fd[1] * fd[2], np.uint16)] * 3
    a_c, a_f = f0.copy(), g0.copy()
    O.run_refined(a_c, a_f, flags, fine, dims, box, m, taus, uins)
    b_c, b_f = f0.copy(), g0.copy()
    O.run_refined(b_c, b_f, flags, fine, dims, box, m, taus, uins, coarse_slip=z3, fine_slip=zf)
    assert np.array_equal(a_c, b_c) and np.array_equal(a_f, b_f)
    # coarse slip on the cells around the block
    sl = np.zeros(n, np.float16)
    sl.reshape(nz, ny, nx)[:, 3:7, 3:9] = np.float16(0.01)
    slip = [sl.view(np.uint16), np.zeros(n, np.uint16), np.zeros(n, np.uint16)]
    c_c, c_f = f0.copy(), g0.copy()
    O.run_refined(c_c, c_f, flags, fine, dims, box, m, taus[:1], uins[:1], coarse_slip=slip)
    d = f0.copy()
    O.run(d, flags, dims, taus[:1].astype(np.float64), uins[:1].astype(np.float64), slip=tuple(slip), cs=0.1733)
    x, y = np.meshgrid(np.arange(nx), np.arange(ny))
    outside = ((x < box[0]) | (x >= box[2]) | (y < box[1]) | (y >= box[3])).reshape(-1)
    outside = np.tile(outside, nz)
    fl = flags == 0
    sel = outside & fl
    cc = c_c.reshape(19, -1)[:, sel]
    dd = d.reshape(19, -1)[:, sel]
    assert np.array_equal(cc, dd)
    e_c, _ = f0.copy(), g0.copy()
    O.run_refined(e_c, g0.copy(), flags, fine, dims, box, m, taus[:1], uins[:1])
    assert not np.array_equal(e_c.reshape(19, -1)[:, sel], cc)


def test_refined_run_qlinks_wiring():
    """The q-LIBB-aware refined runner: zero cut fields equal the plain run
    bit for bit; a coarse q field changes the coarse step outside the patch
    exactly as the single-level step with that field does; a fine q field
    around a solid block inside the patch changes the fine state."""
    import numpy as np
    from oracle import oracle as O
    import foil_cfd_amd as fc
    dims, box, m = (48, 32, 4), (14, 10, 30, 22), 2
    nx, ny, nz = dims
    n = nx * ny * nz
    flags = fc.build_boundary_flags(*dims).reshape(nz, ny, nx)
    flags[:, 4:6, 4:8] = 1  # a small solid block outside the patch
    flags = flags.reshape(-1)
    fd = O.fine_dims(box, dims, m)
    nf = fd[0] * fd[1] * fd[2]
    fine = np.zeros(nf, np.uint8)
    fc.stamp_interface_shell(*fd, fine)
    fine.reshape(fd[2], fd[1], fd[0])[:, 10:13, 12:16] = 1  # a solid block inside the patch
    rng = np.random.default_rng(1)
    f0 = O.init_equilibrium(flags, dims, 0.05, f64=True)
    f0 += 1e-4 * rng.standard_normal(f0.size)
    g0 = O.init_equilibrium(fine, fd, 0.05, f64=True)
    taus, uins = np.full(2, 0.56), np.full(2, 0.05)

    def qfield(fl, d):
        x, y, z = d
        g = fl.reshape(z, y, x)
        qf = np.zeros((z, y, x, 19), np.uint8)
        ff = np.zeros((z, y, x), np.uint32)
        for q in range(1, 19):
            sh = (int(O.CZ[q]), int(O.CY[q]), int(O.CX[q]))
            link = (g == 0) & (np.roll(g, sh, axis=(0, 1, 2)) == 1)
            qf[..., q][link] = 200
            ff[link & (np.roll(g, tuple(-v for v in sh), axis=(0, 1, 2)) == 0)] |= np.uint32(1 << q)
        return qf.reshape(-1), ff.reshape(-1)

    a_c, a_f = f0.copy(), g0.copy()
    O.run_refined(a_c, a_f, flags, fine, dims, box, m, taus, uins)
    b_c, b_f = f0.copy(), g0.copy()
    zero = (np.zeros(n * 19, np.uint8), np.zeros(n, np.uint32))
    O.run_refined(b_c, b_f, flags, fine, dims, box, m, taus, uins, coarse_q=zero,
                  fine_q=(np.zeros(nf * 19, np.uint8), np.zeros(nf, np.uint32)))
    assert np.array_equal(a_c, b_c) and np.array_equal(a_f, b_f)
    cq = qfield(flags, dims)
    assert cq[0].any()
    c_c, c_f = f0.copy(), g0.copy()
    O.run_refined(c_c, c_f, flags, fine, dims, box, m, taus[:1], uins[:1], coarse_q=cq)
    d = f0.copy()
    O.run(d, flags, dims, taus[:1], uins[:1], qfrac=cq[0], ffmask=cq[1], cs=0.1733)
    x, y = np.meshgrid(np.arange(nx), np.arange(ny))
    outside = np.tile(((x < box[0] - 2) | (x >= box[2] + 2) | (y < box[1] - 2) | (y >= box[3] + 2)).reshape(-1), nz)
    sel = outside & (flags == 0)
    assert np.array_equal(c_c.reshape(19, -1)[:, sel], d.reshape(19, -1)[:, sel])
    p_c, p_f = f0.copy(), g0.copy()
    O.run_refined(p_c, p_f, flags, fine, dims, box, m, taus[:1], uins[:1])
    assert not np.array_equal(c_c, p_c)  # the coarse links took effect
    e_c, e_f = f0.copy(), g0.copy()
    O.run_refined(e_c, e_f, flags, fine, dims, box, m, taus, uins, fine_q=qfield(fine, fd))
    assert not np.array_equal(e_f, a_f)


def test_shell_fill_reads_only_inward_populations():
    """The shell fill's read set. Each level's step pulls f_q of a shell cell
    x only into x + c_q, so only the populations whose target lies inside the
    shell's interior can matter; the HIP fill stores only those. With the
    oracle's poison switch every other shell population is NaN after each
    fill. Two- and three-level runs over a foil (solid cells in every level,
    wall-adjacent fluid, restriction, the finer level's band from the fine
    grid) must stay finite and equal the unpoisoned runs bit for bit. A fill
    that dropped one inward population, or a step, band or restriction that
    read an outward one, would fail here."""
    from foil_cfd_amd import cases
    case = cases.naca_tunnel(96, 48, 4, 24, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56)
    box, fine = cases.refinement_patch(case)
    fd = orc.fine_dims(box, case.dims, 2)
    gbox = (6, max(3, fd[1] // 2 - 10), min(fd[0] - 3, 6 + 30), min(fd[1] - 3, fd[1] // 2 + 10))
    finer = cases.finer_patch(case, box, 2, gbox, 2)
    gd = orc.fine_dims(gbox, fd, 2)
    assert (fine == fc.SOLID).any() and (finer == fc.SOLID).any()
    rng = np.random.default_rng(3)
    c0 = orc.init_equilibrium(case.flags, case.dims, 0.05, f64=True)
    c0 += 1e-4 * rng.standard_normal(c0.size)
    f0 = orc.seed_fine(c0, case.flags, case.dims, fine, box, 2, 0.56)
    g0 = orc.seed_fine(f0, fine, fd, finer, gbox, 2, orc.fine_tau_for(0.56, 2))
    taus, uins = np.full(3, 0.56), np.full(3, 0.05)
    runs = {}
    try:
        for poison in (False, True):
            orc.set_shell_poison(poison)
            c, f = c0.copy(), f0.copy()
            orc.run_refined(c, f, case.flags, fine, case.dims, box, 2, taus, uins, cs=0.1)
            c3, f3, g3 = c0.copy(), f0.copy(), g0.copy()
            orc.run_refined3(c3, f3, g3, case.flags, fine, finer, case.dims, box, 2, gbox, 2,
                             taus, uins, cs=0.1)
            runs[poison] = (c, f, c3, f3, g3)
    finally:
        orc.set_shell_poison(False)
    masks = [case.flags, fine, case.flags, fine, finer]
    for a, b, fl in zip(runs[False], runs[True], masks):
        live = (fl == fc.FLUID)
        x, y = a.reshape(Q, -1)[:, live], b.reshape(Q, -1)[:, live]
        assert np.isfinite(y).all()
        assert np.array_equal(x, y)
    # the switch is live: the poisoned fine array holds NaN in its shell slots
    assert np.isnan(runs[True][1].reshape(Q, -1)[:, fine == fc.INTERFACE]).any()
