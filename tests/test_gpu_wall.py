"""GPU wall model (iMEM / Reichardt) against the CPU oracle and the
reference's planted-profile gate (tests/m4_wallmodel.cpp:106-190)."""
import sys

import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd import cases
from foil_cfd_amd.solver import build_wall_cell_list
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
Q = fc.Q


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def half(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float32)


def developed_tunnel(storage="f32", steps=300):
    case = cases.naca_tunnel(192, 64, 16, 48, aoa=8.0, cs=0.1, ramp=False)
    s = cases.make_solver(case, storage)
    s.step_n(steps)
    s.synchronize()
    return case, s


def test_wall_update_parity(gpu):
    say("case: wall update parity, tunnel 192x64x16")
    case, s = developed_tunnel()
    f = s.download_f(f64=False)
    s.set_wall_model("newton")
    n = s.wall_cell_count()
    assert n > 100
    s.wall_update_now()
    ux, uy, uz, utau = s.read_wall_slip()
    wl, st = build_wall_cell_list(case.dims, case.flags)
    assert len(wl["cell"]) == n
    nu = (case.scaling.tau - 0.5) / 3.0
    ut_o = np.zeros(n, np.float32)
    ox, oy, oz, ost = orc.wall_update(f, case.dims, wl, ut_o, nu, 1e-3 * case.scaling.u_lat)
    live = ut_o > 0
    assert live.sum() > 0.5 * n
    assert np.allclose(utau[live], ut_o[live], rtol=2e-4, atol=1e-9)
    assert np.array_equal(utau > 0, live)
    for a, b in ((ux, ox), (uy, oy), (uz, oz)):
        ga, gb = half(a), half(b)
        assert np.allclose(ga, gb, rtol=2e-3, atol=2e-6)
    r = s.wall_model_readout()
    assert r.cells == int(ost[2])
    assert r.mean_yplus == pytest.approx(ost[0] / max(ost[2], 1), rel=1e-3)
    s.close()


def test_wall_model_stepping_parity(gpu):
    """Collision with the moving-wall slip term and the force with the Ladd
    correction, against the oracle fed the same slip field."""
    say("case: wall model stepping parity")
    case, s = developed_tunnel()
    s.set_wall_model("newton")
    s.step_n(5)          # batch 1: ends with an update -> slip field u1
    s.synchronize()
    f = s.download_f(f64=False)
    ux, uy, uz, _ = s.read_wall_slip()
    assert np.abs(half(ux)).max() > 0
    s.step_n(10)         # batch 2 runs with u1; its force sample uses u1
    s.synchronize()
    got = s.download_f(f64=True).reshape(Q, -1)
    raw, at = s.force_raw()
    taus = np.full(10, case.scaling.tau, np.float32)
    uins = np.full(10, case.scaling.u_lat, np.float32)
    orc.run(f, case.flags, case.dims, taus, uins, cs=case.cs, slip=(ux, uy, uz))
    fluid = case.flags == fc.FLUID
    a, b = got[:, fluid], f.reshape(Q, -1)[:, fluid]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6
    Fo = orc.force(f, case.flags, case.dims, slip=(ux, uy, uz))
    assert np.abs(raw - Fo).max() < 1e-4 * np.abs(Fo).max()
    s.close()


def test_planted_reichardt_profile_recovered(gpu):
    """Plant u(y) = u_tau u+((y - 1/2) u_tau / nu) over a flat wall with the
    sample at y+ = 20; the update must recover it (reference gate: within 15%)."""
    say("case: planted Reichardt profile")
    nx, ny, nz = 64, 40, 8
    tau = 0.5005
    nu = (tau - 0.5) / 3.0
    ys = 2.5
    utau = 20.0 * nu / ys
    flags = np.zeros((nz, ny, nx), np.uint8)
    flags[:, 0, :] = fc.SOLID
    flags[:, -1, :] = fc.SOLID
    flags = flags.reshape(-1)
    y = np.arange(ny, dtype=np.float64)
    dist = np.minimum(y - 0.5, (ny - 1.5) - y)  # wall planes at 0.5 and ny-1.5
    u_row = np.array([utau * orc.reichardt_uplus(float(max(d, 0.0) * utau / nu)) for d in dist])
    u = np.broadcast_to(u_row[None, :, None], (nz, ny, nx)).reshape(-1)
    rho = np.ones(nx * ny * nz)
    f0 = cases.equilibrium_field(rho, u, np.zeros_like(u), np.zeros_like(u))
    sc = fc.scaling_for_tau(tau, 0.05, 64)
    s = fc.LBMSolver()
    s.set_startup_ramp_enabled(False)
    s.init((nx, ny, nz), sc, flags, storage="f32", periodic_xy=True)
    s.upload_f(f0)
    s.set_wall_model("newton")
    s.wall_update_now()
    r = s.wall_model_readout()
    say(f"planted y+ 20 -> mean {r.mean_yplus:.3f} max {r.max_yplus:.3f} clamped {r.clamped_frac:.3f}")
    assert r.cells == 2 * nx * nz
    assert r.mean_yplus == pytest.approx(20.0, rel=0.15)   # the reference's gate
    assert r.clamped_frac < 0.2
    _, _, _, ut = s.read_wall_slip()
    wl, _ = build_wall_cell_list((nx, ny, nz), flags)
    x = wl["cell"] % nx
    inner = (x >= 2) & (x < nx - 2)  # the normal stencil clips at the x faces
    assert np.allclose(ut[inner], utau, rtol=0.02)
    assert np.allclose(wl["ys"][inner] * ut[inner] / nu, 20.0, rtol=0.02)
    s.close()


def test_wall_model_fp16_runs(gpu):
    say("case: wall model f16 stepping")
    case, s = developed_tunnel("f16", 200)
    s.set_wall_model("newton")
    s.step_n(100)
    s.synchronize()
    r = s.wall_model_readout()
    assert r.cells > 0 and np.isfinite(r.mean_yplus)
    assert not s.nan_detected()
    s.close()


def test_fine_level_wall_model(gpu):
    """Wall model on the refinement patch (rebuildWallModelFine and the fine
    update, lbm_solver.cpp:425-438, :1368-1377): the fine list at sample
    distance 3, its update at the fine viscosity against the oracle, the fine
    readout, and one coupled step with both levels' slip fields against the
    oracle's refined run fed the same fields."""
    say("case: fine-level wall model, refined tunnel 160x64x8, f64")
    case = cases.naca_tunnel(160, 64, 8, 32, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56, ramp=False)
    box, fine = cases.refinement_patch(case)
    s = cases.make_solver(case, "f64")
    s.init_refinement(box, 2, fine)
    s.step_n(120)
    s.synchronize()
    s.set_wall_model("newton")
    fd = orc.fine_dims(box, case.dims, 2)
    nf = s.wall_cell_count_level(1)
    wl, _ = build_wall_cell_list(fd, fine, sample_cells=3)
    assert nf == len(wl["cell"]) and nf > 100
    assert s.wall_cell_count_level(2) == -1
    ff = s.download_fine_f(f64=True)
    s.wall_update_now()
    ux, uy, uz, ut = s.read_wall_slip_level(1)
    tau_f = orc.fine_tau_for(case.scaling.tau, 2)
    nu = (tau_f - 0.5) / 3.0
    ut_o = np.zeros(nf, np.float32)
    ox, oy, oz, ost = orc.wall_update(ff.astype(np.float32), fd, wl, ut_o, nu, 1e-3 * case.scaling.u_lat)
    live = ut_o > 0
    say(f"  fine wall cells {nf}, live {int(live.sum())}")
    assert live.sum() > 0.5 * nf
    assert np.allclose(ut[live], ut_o[live], rtol=2e-4, atol=1e-9)
    for a, b in ((ux, ox), (uy, oy), (uz, oz)):
        assert np.allclose(half(a), half(b), rtol=2e-3, atol=2e-6)
    r = s.wall_model_readout()
    assert r.from_fine == 1 and r.listed == nf and r.cells == int(ost[2])
    # one coupled step with the current slip fields of both levels
    cux, cuy, cuz, _ = s.read_wall_slip_level(0)
    cf = s.download_f(f64=True)
    ff = s.download_fine_f(f64=True)
    s.step_n(1)
    s.synchronize()
    got_c = s.download_f(f64=True).reshape(Q, -1)
    got_f = s.download_fine_f(f64=True).reshape(Q, -1)
    taus = np.full(1, case.scaling.tau)
    uins = np.full(1, case.scaling.u_lat)
    orc.run_refined(cf, ff, case.flags, fine, case.dims, box, 2, taus, uins, cs=case.cs,
                    coarse_slip=(cux, cuy, cuz), fine_slip=(ux, uy, uz))
    cfl, ffl = case.flags == fc.FLUID, fine == fc.FLUID
    ec = np.linalg.norm(got_c[:, cfl] - cf.reshape(Q, -1)[:, cfl]) / np.linalg.norm(cf.reshape(Q, -1)[:, cfl])
    ef = np.linalg.norm(got_f[:, ffl] - ff.reshape(Q, -1)[:, ffl]) / np.linalg.norm(ff.reshape(Q, -1)[:, ffl])
    say(f"  coupled step with slip: coarse {ec:.2e}, fine {ef:.2e}")
    assert ec < 1e-6 and ef < 1e-6
    # the fine force carries the fine slip (forcesFromFine)
    if s.refinement_info()["forces_from_fine"]:
        raw, _ = s.force_raw()
        assert np.all(np.isfinite(raw))
    s.close()
