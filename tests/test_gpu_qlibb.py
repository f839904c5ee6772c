"""GPU q-LIBB (sparse pre-pass on the in-place store) against the oracle's
in-loop restatement of lbm_core.cu:338-364, with synthetic cut fractions
uniform in [0,1) from a fixed-seed lowbias32 stream (the config-4 recipe)."""
import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd import cases
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
Q = fc.Q


def lowbias32(x):
    x = np.uint32(x)
    x ^= x >> np.uint32(16)
    x = np.uint32((int(x) * 0x7FEB352D) & 0xFFFFFFFF)
    x ^= x >> np.uint32(15)
    x = np.uint32((int(x) * 0x846CA68B) & 0xFFFFFFFF)
    x ^= x >> np.uint32(16)
    return int(x)


def synthetic_qlinks(flags, dims, seed=1234):
    nx, ny, nz = dims
    n = nx * ny * nz
    qfrac = np.zeros((n, Q), np.uint8)
    ffm = np.zeros(n, np.uint32)
    g = flags.reshape(nz, ny, nx)
    state = seed
    links = 0
    for c in np.nonzero(flags == fc.FLUID)[0]:
        x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
        for q in range(1, Q):
            sx, sy, sz = x - fc.CX[q], y - fc.CY[q], (z - fc.CZ[q]) % nz
            if g[sz, sy, sx] != fc.SOLID:
                continue
            state = lowbias32(state)
            u = (state & 0xFFFFFF) / 16777216.0
            qc = min(max(u, 0.05), 1.0)
            b = int(np.floor(qc * 255.0 + 0.5)) or 1
            qfrac[c, q] = b
            fx, fy, fz = x + fc.CX[q], y + fc.CY[q], (z + fc.CZ[q]) % nz
            if g[fz, fy, fx] == fc.FLUID:
                ffm[c] |= np.uint32(1 << q)
            links += 1
    return qfrac.reshape(-1), ffm, links


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_qlibb_parity(gpu, storage):
    """FP64 storage: within 1e-12 of the FP64 oracle. FP32 storage (which
    keeps f - w): against the FP64 oracle, no further from it than the FP32
    oracle (which keeps f, as the reference does) and within 1e-6."""
    print(f"case: q-LIBB tunnel {storage}", flush=True)
    case = cases.naca_tunnel(128, 48, 8, 40, aoa=8.0, cs=0.1, ramp=False)
    qfrac, ffm, links = synthetic_qlinks(case.flags, case.dims)
    s = cases.make_solver(case, storage)
    s.set_qlibb_enabled(True)
    s.set_qlinks(qfrac, ffm, links, 0)
    r = s.qlibb_readout()
    assert r["enabled"] == 1 and r["links"] == links and r["active"] > 0
    f = s.download_f(f64=(storage == "f64"))
    steps = 25
    s.step_n(steps)
    s.synchronize()
    got = s.download_f(f64=True).reshape(Q, -1)
    fluid = case.flags == fc.FLUID

    def oracle(dt):
        g = np.array(f, dtype=dt)
        taus = np.full(steps, case.scaling.tau, dt)
        uins = np.full(steps, case.scaling.u_lat, dt)
        orc.run(g, case.flags, case.dims, taus, uins, cs=case.cs, qfrac=qfrac, ffmask=ffm)
        return g.reshape(Q, -1)[:, fluid].astype(np.float64)

    a = got[:, fluid]
    b = oracle(np.float64)
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"q-LIBB {storage} rel L2 vs FP64 oracle after {steps} steps: {err:.3e}", flush=True)
    if storage == "f64":
        tol = 1e-12
        assert err < tol
    else:
        tol = 1e-6
        e32 = np.linalg.norm(oracle(np.float32) - b) / np.linalg.norm(b)
        print(f"FP32 oracle rel L2 vs FP64 oracle: {e32:.3e}", flush=True)
        assert err < tol and err <= max(2.0 * e32, 1e-7)
    # and the q-links actually changed the flow vs plain bounce-back
    s2 = cases.make_solver(case, storage)
    s2.step_n(steps)
    s2.synchronize()
    plain = s2.download_f(f64=True).reshape(Q, -1)[:, fluid]
    assert np.linalg.norm(plain - a) / np.linalg.norm(a) > 100 * tol
    s.close()
    s2.close()


def synthetic_qlinks_fast(flags, dims, seed):
    """Cut bytes in [13, 255] from a seeded generator on every Fluid cell
    link whose upstream neighbour is Solid, two-node bit where the
    downstream neighbour is Fluid (x / y borders of a patch level are
    Interface cells, so the wrap of np.roll never pairs two fluid cells)."""
    nx, ny, nz = dims
    g = np.asarray(flags, np.uint8).reshape(nz, ny, nx)
    rng = np.random.default_rng(seed)
    qfrac = np.zeros((nz, ny, nx, Q), np.uint8)
    ffm = np.zeros((nz, ny, nx), np.uint32)
    fluid = g == fc.FLUID
    links = 0
    for q in range(1, Q):
        sh = (int(fc.CZ[q]), int(fc.CY[q]), int(fc.CX[q]))
        up = np.roll(g, sh, axis=(0, 1, 2))
        down = np.roll(g, tuple(-v for v in sh), axis=(0, 1, 2))
        link = fluid & (up == fc.SOLID)
        qfrac[..., q][link] = rng.integers(13, 256, size=int(link.sum()), dtype=np.uint8)
        ffm[link & (down == fc.FLUID)] |= np.uint32(1 << q)
        links += int(link.sum())
    return qfrac.reshape(-1), ffm.reshape(-1), links


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_fine_level_qlibb_parity(gpu, storage):
    """setFineQLinks (lbm_solver.cpp:1049-1055): q-links on the fine patch.
    The coupled run matches the oracle's refined run with the fine level's
    q fields at the refinement tolerance (the coupling computes in FP32), the
    links change the fine flow, the readout sums the levels, and switching
    q-LIBB off frees every level's links."""
    print(f"case: refined tunnel 160x64x8 with fine-level q-links, {storage}", flush=True)
    case = cases.naca_tunnel(160, 64, 8, 32, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56, ramp=False)
    box, fine = cases.refinement_patch(case)
    s = cases.make_solver(case, storage)
    s.init_refinement(box, 2, fine)
    fd = s.refinement_info()["fine_dims"]
    qfrac, ffm, links = synthetic_qlinks_fast(fine, fd, seed=99)
    s.set_qlibb_enabled(True)
    with pytest.raises(RuntimeError):
        s.set_fine_qlinks(qfrac, ffm, links, 0, level=2)  # no nested level
    s.set_fine_qlinks(qfrac, ffm, links, 3)
    r = s.qlibb_readout()
    assert r["enabled"] == 1 and r["links"] == links and r["fallback"] == 3 and r["active"] > 0
    cf = s.download_f(f64=True)
    ff = s.download_fine_f(f64=True)
    steps = 12
    s.step_n(steps)
    s.synchronize()
    got_c = s.download_f(f64=True).reshape(Q, -1)
    got_f = s.download_fine_f(f64=True).reshape(Q, -1)
    sc = case.scaling
    taus, uins = orc.ramp_schedule(sc.tau, sc.nu_lat, sc.u_lat, case.dims[0], steps, ramp=case.ramp)
    orc.run_refined(cf, ff, case.flags, fine, case.dims, box, 2, taus.astype(np.float64), uins.astype(np.float64),
                    cs=case.cs, fine_q=(qfrac, ffm))
    cfl = case.flags == fc.FLUID
    ffl = fine == fc.FLUID
    ec = np.linalg.norm(got_c[:, cfl] - cf.reshape(Q, -1)[:, cfl]) / np.linalg.norm(cf.reshape(Q, -1)[:, cfl])
    ef = np.linalg.norm(got_f[:, ffl] - ff.reshape(Q, -1)[:, ffl]) / np.linalg.norm(ff.reshape(Q, -1)[:, ffl])
    print(f"  {r['active']} active links; coarse {ec:.3e}, fine {ef:.3e} rel L2 vs the oracle", flush=True)
    assert ec < 1e-6 and ef < 1e-6
    plain = cases.make_solver(case, storage)
    plain.init_refinement(box, 2, fine)
    plain.step_n(steps)
    plain.synchronize()
    pf = plain.download_fine_f(f64=True).reshape(Q, -1)[:, ffl]
    dq = np.linalg.norm(pf - got_f[:, ffl]) / np.linalg.norm(got_f[:, ffl])
    print(f"  plain bounce-back vs q-LIBB on the fine level: {dq:.3e}", flush=True)
    assert dq > 3e-5  # the links' effect is far above the parity tolerance
    s.set_qlibb_enabled(False)
    r = s.qlibb_readout()
    assert r["links"] == 0 and r["active"] == 0
    s.close()
    plain.close()


def test_vg_vanes_on_fine_level(gpu):
    """The vortex-generator path on the patch level, as the reference app
    runs it (main.cpp:520-560): vanes voxelized into the fine flags
    (buildFlagsWithVGs on the fine layout), their cut links (buildVaneQLinks
    + densifyQLinks) uploaded with setFineQLinks, the wall model on both
    levels against the clean references. The run stays finite with every
    link active, and the record-writing fine step (vanes are Solid, so the
    restricted box stays Fluid / Solid) equals the separate child-average
    pass bit for bit."""
    from foil_cfd_amd import solver as S
    print("case: refined tunnel 192x72x16 at 10 deg, VG pair on the fine level, f16c", flush=True)
    case = cases.naca_tunnel(192, 72, 16, 48, aoa=10.0, cs=0.1, u_lat=0.05, tau=0.56, ramp=False)
    box, clean = cases.refinement_patch(case, aoa=10.0)
    layout = S.make_layout(*case.dims, case.scaling.chord_cells)
    fl = S.make_fine_layout(layout, S.FcPatchBox(*box), 2)
    pts, le = S.generate_naca4("2412", 200)
    vg = S.vg_params(x_c=0.2, height_c=0.04, length_h=2.5, beta_deg=16.0, pitch_c=0.24, gap_h=2.5, count=2)
    flags, slabs = S.build_flags_with_vgs(pts, le, 10.0, fl, [vg], clean)
    fd = (fl.dims.nx, fl.dims.ny, fl.dims.nz)
    assert ((flags == fc.SOLID) & (clean != fc.SOLID)).sum() > 100
    qf, ff, links = S.build_vane_qlinks(fd, flags, clean, slabs)
    assert links > 100
    out = []
    for variant in ("auto", "single"):
        s = cases.make_solver(case, "f16c")
        s.set_surface_reference(case.flags)
        s.init_refinement(box, 2, flags)
        s.set_refined_surface_reference(clean)
        s.set_wall_model("newton")
        s.set_qlibb_enabled(True)
        s.set_fine_qlinks(qf, ff, links, 0)
        s.set_step_variant(variant)
        r = s.qlibb_readout()
        assert r["links"] == links and r["active"] > 0
        for _ in range(4):
            s.step_n(25)
        s.synchronize()
        assert not s.nan_detected()
        assert s.wall_cell_count_level(1) > 0
        out.append((s.download_f(f64=True), s.download_fine_f(f64=True)))
        print(f"  {variant}: {r['active']} active fine links, fine wall cells {s.wall_cell_count_level(1)}",
              flush=True)
        s.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert not np.count_nonzero((a - b)[~np.isnan(a)])
