"""GPU two-level refinement (band snapshot, shell fill, LDS-tiled restriction
on the in-place store) against the oracle's restatement of the reference
coupling (lbm_refine.cu, stepN sequencing lbm_solver.cpp:1172-1290).

The coupling kernels compute in FP32 for every storage format, so the FP64
storage run is compared at FP32 tolerance; the tolerances are far below the
O(1e-3) a wrong slot, layout or time level produces."""
import sys

import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd import cases
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
Q = fc.Q


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def refined_case():
    case = cases.naca_tunnel(160, 64, 8, 32, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56)
    box, fine = cases.refinement_patch(case)
    return case, box, fine


def schedule(case, steps):
    s = case.scaling
    return orc.ramp_schedule(s.tau, s.nu_lat, s.u_lat, case.dims[0], steps, ramp=case.ramp)


def test_refined_uniform_freestream_stays_uniform(gpu):
    """Body-free tunnel at the inlet's uniform equilibrium with a factor-2
    patch (the reference's m3 gate): coarse and fine stay uniform."""
    say("case: refined uniform freestream 96x48x8, box (20,12,60,36), f32, 40 steps")
    dims, box, m, u0 = (96, 48, 8), (20, 12, 60, 36), 2, 0.05
    flags = fc.build_boundary_flags(*dims)
    sc = fc.scaling_for_tau(0.56, u0, 32)
    feq = orc.equilibrium(1.0, u0, 0.0, 0.0)
    n = dims[0] * dims[1] * dims[2]
    case = cases.Case("uniform", dims, flags, sc, cs=0.1, ramp=False,
                      f0=np.repeat(feq[:, None], n, axis=1).reshape(-1))
    fd = orc.fine_dims(box, dims, m)
    fine = np.zeros(fd[0] * fd[1] * fd[2], np.uint8)
    fc.stamp_interface_shell(*fd, fine)
    s = cases.make_solver(case, "f32")
    s.init_refinement(box, m, fine)
    info = s.refinement_info()
    assert info["active"] and info["fine_dims"] == fd and info["factor"] == m
    s.step_n(40)
    s.synchronize()
    rho, u, v, w = s.read_macro()
    fl = flags == fc.FLUID
    say(f"coarse max |u-u0| {np.abs(u[fl] - u0).max():.2e}, max |v| {np.abs(v[fl]).max():.2e}")
    assert np.abs(u[fl] - u0).max() < 2e-6 and np.abs(v[fl]).max() < 2e-6
    assert np.abs(rho[fl] - 1.0).max() < 2e-6
    g = s.download_fine_f(f64=True).reshape(Q, -1)[:, fine == fc.FLUID]
    assert np.abs(g - feq[:, None]).max() < 2e-7
    assert not s.nan_detected()
    s.close()


def test_refined_seed_matches_oracle(gpu):
    """The fine level right after init_refinement (rest + full-volume fill of
    the current coarse state) against the oracle's seedFineFromCoarse."""
    say("case: refined tunnel 160x64x8 seeding, f64")
    case, box, fine = refined_case()
    s = cases.make_solver(case, "f64")
    s.step_n(30)
    s.synchronize()
    s.init_refinement(box, 2, fine)
    cf = s.download_f(f64=True)
    got = s.download_fine_f(f64=True).reshape(Q, -1)
    tau_c = fc.ramped_tau(case.scaling, 30, case.dims[0])
    want = orc.seed_fine(cf, case.flags, case.dims, fine, box, 2, tau_c).reshape(Q, -1)
    live = fine != fc.SOLID
    err = rel_l2(got[:, live], want[:, live])
    say(f"seed rel L2 vs oracle {err:.3e}")
    assert err < 1e-6
    s.close()


@pytest.mark.parametrize("storage,steps", [("f64", 1), ("f64", 8), ("f32", 8), ("f32", 60)])
def test_refined_step_parity(gpu, storage, steps):
    """Coupled steps (coarse step, two fine sub-steps with the shell fill,
    restriction) against the FP64 oracle from the same coarse and fine state."""
    say(f"case: refined tunnel 160x64x8 {storage} steps={steps}")
    case, box, fine = refined_case()
    s = cases.make_solver(case, storage)
    s.init_refinement(box, 2, fine)
    cf = s.download_f(f64=True)
    ff = s.download_fine_f(f64=True)
    s.step_n(steps)
    s.synchronize()
    got_c = s.download_f(f64=True).reshape(Q, -1)
    got_f = s.download_fine_f(f64=True).reshape(Q, -1)
    r2, u2, v2, _ = s.read_macro()
    taus, uins = schedule(case, steps)
    rho, u, v, w = orc.run_refined(cf, ff, case.flags, fine, case.dims, box, 2,
                                   taus.astype(np.float64), uins.astype(np.float64), cs=case.cs)
    cfl = case.flags == fc.FLUID
    ffl = fine == fc.FLUID
    ec = rel_l2(got_c[:, cfl], cf.reshape(Q, -1)[:, cfl])
    ef = rel_l2(got_f[:, ffl], ff.reshape(Q, -1)[:, ffl])
    ul = case.scaling.u_lat
    eu = float(np.linalg.norm(np.stack([u2, v2]) - np.stack([u, v])) / (ul * np.sqrt(2 * u.size)))
    say(f"coarse pops {ec:.3e}, fine pops {ef:.3e}, coarse velocity (rel to u_lat) {eu:.3e}")
    assert ec < 1e-6 and ef < 1e-6
    assert eu < 1e-5
    assert rel_l2(r2[cfl], rho[cfl]) < 1e-6
    s.close()


def test_refined_forces_on_fine_grid(gpu):
    """With the patch covering the foil the momentum exchange runs on the fine
    grid: the readout equals the oracle's force on the fine state (FP64
    storage, freestream start so the force is developed after 40 steps)."""
    say("case: refined tunnel 160x64x8 f64 forces from fine, 40 steps")
    case, box, fine = refined_case()
    case.ramp = False
    s = cases.make_solver(case, "f64")
    s.init_refinement(box, 2, fine)
    info = s.refinement_info()
    assert info["forces_from_fine"]
    ff = s.download_fine_f(f64=True)
    cf = s.download_f(f64=True)
    s.step_n(40)
    s.synchronize()
    F = s.force_now()
    taus, uins = schedule(case, 40)
    orc.run_refined(cf, ff, case.flags, fine, case.dims, box, 2, taus.astype(np.float64),
                    uins.astype(np.float64), cs=case.cs)
    fd = info["fine_dims"]
    Fo = orc.force(ff, fine, fd)
    say(f"fine force gpu {F} vs oracle {Fo}")
    assert np.abs(Fo).max() > 1e-4
    assert np.abs(F - Fo).max() < 1e-4 * np.abs(Fo).max()
    r = s.forces()
    assert np.isfinite(r.cl) and np.isfinite(r.cd)
    s.close()


def test_refined_run_stays_finite_and_counts_updates(gpu):
    """500 coupled steps of the refined tunnel in FP16C: finite fields, no
    watchdog trip, and the perf counter counts m x fine cell updates."""
    say("case: refined tunnel 160x64x8 f16c 500 steps")
    case, box, fine = refined_case()
    s = cases.make_solver(case, "f16c")
    s.init_refinement(box, 2, fine)
    for _ in range(5):
        s.step_n(100)
    s.synchronize()
    rho, u, v, w = s.read_macro()
    assert np.isfinite(u).all() and not s.nan_detected()
    g = s.download_fine_f(f64=True).reshape(Q, -1)[:, fine == fc.FLUID]
    assert np.isfinite(g).all()
    p = s.perf_stats()
    n_c = int(np.prod(case.dims))
    n_f = int(np.prod(s.refinement_info()["fine_dims"]))
    assert p.population_bytes_per_step == pytest.approx(76.0 * (n_c + 2 * n_f))
    s.close()


@pytest.mark.parametrize("storage,wall,variant", [("f32", "off", "auto"), ("f16c", "off", "auto"),
                                                  ("f16c", "newton", "auto"), ("f64", "off", "auto"),
                                                  ("f16c", "off", "wide"), ("f32", "off", "wide")])
def test_record_writing_fine_step_matches_separate_pass(gpu, storage, wall, variant):
    """With the default step variant the last fine sub-step of each coarse
    step writes the restriction's child-average records from its registers;
    the single-cell variant keeps the separate child-average pass over the
    fine array. Coarse and fine states and the macro fields must agree bit
    for bit after several coupled steps ("wide": the default with the 64-bit
    population index of very large grids forced)."""
    say(f"case: refined tunnel 160x64x8 {storage} wall={wall}, {variant} record-writing fine step vs separate pass")
    case, box, fine = refined_case()
    out = []
    for variant in (variant, "single"):
        s = cases.make_solver(case, storage)
        s.init_refinement(box, 2, fine)
        if wall != "off":
            s.set_refined_surface_reference(fine)
            s.set_wall_model(wall)
        s.set_step_variant(variant)
        for _ in range(3):
            s.step_n(5)
        s.synchronize()
        out.append((s.download_f(f64=True), s.download_fine_f(f64=True), s.read_macro()))
        assert not s.nan_detected()
        s.close()
    (ca, fa, ma), (cb, fb, mb) = out
    # slots of cells without populations (solid, Interface) read back as NaN
    for name, a, b in (("coarse", ca, cb), ("fine", fa, fb)):
        assert np.array_equal(np.isnan(a), np.isnan(b)), name
        d = np.abs(a - b)[~np.isnan(a)]
        assert not np.count_nonzero(d), f"{name}: {np.count_nonzero(d)} values differ, max {d.max():.3e}"
    for x, y in zip(ma, mb):
        assert np.array_equal(x, y, equal_nan=True)


# ---- nested third level --------------------------------------------------------
def test_three_level_uniform_freestream_stays_uniform(gpu):
    """Body-free tunnel with a factor-2 fine patch and a factor-2 finer patch
    nested in it (the reference's m4_finer 3-level gate): all three levels
    stay at the inlet's uniform equilibrium."""
    say("case: 3-level uniform freestream 96x48x8, fine (20,12,60,36), finer (20,10,60,38), f32, 40 steps")
    dims, box, m, u0 = (96, 48, 8), (20, 12, 60, 36), 2, 0.05
    gbox, m2 = (20, 10, 60, 38), 2
    flags = fc.build_boundary_flags(*dims)
    sc = fc.scaling_for_tau(0.56, u0, 32)
    feq = orc.equilibrium(1.0, u0, 0.0, 0.0)
    n = dims[0] * dims[1] * dims[2]
    case = cases.Case("uniform", dims, flags, sc, cs=0.1, ramp=False,
                      f0=np.repeat(feq[:, None], n, axis=1).reshape(-1))
    fd = orc.fine_dims(box, dims, m)
    fine = np.zeros(fd[0] * fd[1] * fd[2], np.uint8)
    fc.stamp_interface_shell(*fd, fine)
    gd = orc.fine_dims(gbox, fd, m2)
    finer = np.zeros(gd[0] * gd[1] * gd[2], np.uint8)
    fc.stamp_interface_shell(*gd, finer)
    s = cases.make_solver(case, "f32")
    s.init_refinement(box, m, fine)
    s.init_finer_refinement(gbox, m2, finer)
    info = s.finer_refinement_info()
    assert info["active"] and info["finer_dims"] == gd
    s.step_n(40)
    s.synchronize()
    rho, u, v, w = s.read_macro()
    fl = flags == fc.FLUID
    assert np.abs(u[fl] - u0).max() < 2e-6 and np.abs(v[fl]).max() < 2e-6
    for got, ff in ((s.download_fine_f(f64=True), fine), (s.download_finer_f(f64=True), finer)):
        g = got.reshape(Q, -1)[:, ff == fc.FLUID]
        say(f"  level max |f - feq| {np.abs(g - feq[:, None]).max():.2e}")
        # FP32 rounding of the coupling (2.3e-7 measured on the fine level);
        # a wrong slot, layout or time level gives O(1e-3)
        assert np.abs(g - feq[:, None]).max() < 1e-6
    assert s.perf_stats().population_bytes_per_step > 0
    assert not s.nan_detected()
    s.close()


def nested_case():
    case, box, fine = refined_case()
    fd = orc.fine_dims(box, case.dims, 2)
    # finer box around the foil's front half, 3+ fine cells inside the fine shell
    gbox = (6, max(3, fd[1] // 2 - 14), min(fd[0] - 3, 6 + 40), min(fd[1] - 3, fd[1] // 2 + 14))
    finer = cases.finer_patch(case, box, 2, gbox, 2)
    return case, box, fine, gbox, finer


@pytest.mark.parametrize("steps", [1, 6])
def test_three_level_step_parity(gpu, steps):
    """Coupled three-level steps on the refined tunnel (foil in all three
    levels) against the FP64 oracle's restatement of the nested sequencing,
    from the same coarse, fine and finer state (FP64 storage; the coupling
    kernels compute in FP32)."""
    case, box, fine, gbox, finer = nested_case()
    say(f"case: 3-level tunnel 160x64x8, fine {box}, finer {gbox}, f64 steps={steps}")
    s = cases.make_solver(case, "f64")
    s.step_n(10)
    s.init_refinement(box, 2, fine)
    s.step_n(3)
    s.init_finer_refinement(gbox, 2, finer)
    assert (finer == fc.SOLID).any()
    cf = s.download_f(f64=True)
    ff = s.download_fine_f(f64=True)
    gf = s.download_finer_f(f64=True)
    start = s.step_count()
    s.step_n(steps)
    s.synchronize()
    got = [s.download_f(f64=True), s.download_fine_f(f64=True), s.download_finer_f(f64=True)]
    sc = case.scaling
    taus, uins = orc.ramp_schedule(sc.tau, sc.nu_lat, sc.u_lat, case.dims[0], start + steps, ramp=case.ramp)
    orc.run_refined3(cf, ff, gf, case.flags, fine, finer, case.dims, box, 2, gbox, 2,
                     taus[start:].astype(np.float64), uins[start:].astype(np.float64), cs=case.cs)
    for name, g, w, fl in (("coarse", got[0], cf, case.flags), ("fine", got[1], ff, fine),
                           ("finer", got[2], gf, finer)):
        live = fl == fc.FLUID
        e = rel_l2(g.reshape(Q, -1)[:, live], w.reshape(Q, -1)[:, live])
        say(f"  {name} populations rel L2 {e:.3e}")
        assert e < 1e-6, name
    s.close()
