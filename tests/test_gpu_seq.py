"""Mesh sequencing on the GPU (seedFromCoarse, lbm_solver.cpp:2062-2102;
upsampleMacroKernel, lbm_refine.cu:437-501; runPreconverge, main.cpp:756-817):
the seeded macro fields are the oracle's FP32 upsample of the presolver's
fields bit for bit, the populations are feq of them (storage tolerance), the
startup ramp is re-armed and the gate closed, and fc_preconverge equals the
same sequence run by hand."""
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


def target_case():
    return cases.naca_tunnel(192, 72, 16, 64, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56, ramp=False)


@pytest.mark.parametrize("storage", ["f32", "f16c"])
def test_seed_from_coarse_matches_oracle(gpu, storage):
    say(f"case: seed 192x72x16 (chord 64) from a 96x36x8 presolve after 400 steps, {storage}")
    case = target_case()
    pre_case = cases.presolve_case(case, chord=32)
    pre = cases.make_solver(pre_case, storage)
    pre.step_n(400)
    pre.synchronize()
    pre_fields = pre.read_macro()
    s = cases.make_solver(case, storage)
    s.step_n(5)
    s.seed_from_coarse(pre)
    rho, u, v, w = s.read_macro()
    want = orc.upsample_macro(pre_fields, pre_case.dims, case.dims)
    for name, a, b in zip("rho u v w".split(), (rho, u, v, w), want):
        assert np.array_equal(a, b), f"{name}: max |diff| {np.abs(a - b).max():.3e}"
    f = s.download_f(f64=True).reshape(Q, -1)
    feq = orc.init_from_macro(case.flags, rho, u, v, w)
    tol = 2e-7 if storage == "f32" else 1e-6
    rel = 0.0 if storage == "f32" else 2.0 ** -12
    excess = np.abs(f - feq) - (tol + rel * np.abs(feq - orc.W[:, None]))
    assert excess.max() <= 0
    # the ramp re-arms over the seeded field: elevated viscosity at its start
    s2 = cases.make_solver(case, storage)
    s2.set_startup_ramp_enabled(True)
    s2.seed_from_coarse(pre)
    assert s2.current_tau() == pytest.approx(fc.ramped_tau(case.scaling, 0, case.dims[0]), rel=1e-6)
    assert s2.current_tau() > case.scaling.tau
    assert not s.forces().valid
    s.step_n(300)
    s.synchronize()
    assert not s.nan_detected()
    with pytest.raises(RuntimeError):
        s.seed_from_coarse(s)
    for x in (pre, s, s2):
        x.close()


def test_preconverge_equals_manual_sequence(gpu):
    say("case: fc_preconverge 192x72x16 from a chord-32 presolve vs the manual sequence, f32")
    case = target_case()
    pre_case = cases.presolve_case(case, chord=32)
    total = int(np.float32(2.0) * (np.float32(pre_case.dims[0]) / np.float32(pre_case.scaling.u_lat)))
    a = cases.make_solver(case, "f32")
    steps = a.preconverge(pre_case.dims, pre_case.scaling, pre_case.flags, "f32")
    assert steps == total
    pre = fc.LBMSolver()
    pre.init(pre_case.dims, pre_case.scaling, pre_case.flags, storage="f32")
    pre.set_smagorinsky_cs(case.cs)
    done = 0
    while done < total:
        n = min(1024, total - done)
        pre.step_n(n)
        pre.synchronize()
        assert not pre.nan_detected()
        done += n
    b = cases.make_solver(case, "f32")
    b.seed_from_coarse(pre)
    fa, fb = a.download_f(f64=True), b.download_f(f64=True)
    assert np.array_equal(fa, fb)
    ua, ub = a.read_macro()[1], b.read_macro()[1]
    say(f"  presolve {pre_case.dims} ran {steps} steps; seeded max u {ua.max():.4f}")
    assert np.array_equal(ua, ub) and ua.max() > 0.5 * case.scaling.u_lat
    for x in (a, b, pre):
        x.close()


def test_seed_with_fine_patch(gpu):
    say("case: seed a refined tunnel 160x64x8 from a chord-16 presolve, f32")
    case = cases.naca_tunnel(160, 64, 8, 32, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56, ramp=False)
    box, fine = cases.refinement_patch(case)
    pre_case = cases.presolve_case(case, chord=16)
    pre = cases.make_solver(pre_case, "f32")
    pre.step_n(200)
    s = cases.make_solver(case, "f32")
    s.init_refinement(box, 2, fine)
    s.seed_from_coarse(pre)
    cf = s.download_f(f64=True)
    got = s.download_fine_f(f64=True).reshape(Q, -1)
    tau_c = s.current_tau()
    want = orc.seed_fine(cf, case.flags, case.dims, fine, box, 2, tau_c).reshape(Q, -1)
    live = fine != fc.SOLID
    err = float(np.linalg.norm(got[:, live] - want[:, live]) / np.linalg.norm(want[:, live]))
    say(f"  fine re-seed vs oracle rel L2 {err:.2e}")
    assert err < 1e-6
    s.step_n(100)
    s.synchronize()
    assert not s.nan_detected()
    for x in (pre, s):
        x.close()


def test_adaptive_steps_pacer(gpu):
    """adaptiveStepsForBudget (lbm_solver.cpp:1392-1408): 1 until a batch has
    been timed, then budget / (EMA ms per step) with growth capped at 2x per
    call and the 4096 ceiling; 0 once the watchdog has latched."""
    say("case: frame pacer on a 96x48x8 tunnel, f32")
    case = cases.naca_tunnel(96, 48, 8, 32, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56, ramp=False)
    s = cases.make_solver(case, "f32")
    assert s.adaptive_steps_for_budget(10.0) == 1
    s.step_n(64)
    s.synchronize()
    s.step_n(64)
    s.synchronize()
    assert s.perf_stats().last_step_ms > 0
    seq = [s.adaptive_steps_for_budget(10.0) for _ in range(16)]
    say(f"  choices {seq}")
    assert seq[0] <= 2 and all(b <= 2 * a for a, b in zip(seq, seq[1:]))
    assert seq[-1] == seq[-2] and 1 <= seq[-1] <= 4096  # settled at budget / EMA
    assert s.perf_stats().last_steps_per_frame == seq[-1]
    assert s.adaptive_steps_for_budget(1e9) == min(4096, 2 * seq[-1])
    assert s.adaptive_steps_for_budget(1e-9) == 1  # shrinking is immediate
    f = s.download_f(f64=False).reshape(Q, case.dims[2], case.dims[1], case.dims[0])
    f[:, 3, 1:-1, 1:-1] = np.nan
    s.upload_f(f.reshape(-1))
    s.step_n(400)
    s.synchronize()
    assert s.nan_detected()
    assert s.adaptive_steps_for_budget(10.0) == 0
    s.close()
