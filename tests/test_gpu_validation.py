"""The reference's physics milestones M0-M2 run through the HIP path, checked
against the numbers the reference itself reports for them
(notes/VALIDATION.md) and against its own acceptance gates.

* M0 (tests/m0_taylor_green.cpp:36-157): 64^3 Taylor-Green field in the
  open tunnel domain (inlet / outlet in x, free slip in y, periodic z),
  nu 0.01, inlet 1e-3, Smagorinsky 0.1733, injected through the snapshot
  path (no ramps); kinetic energy of the fluid cells against the reported
  trace KE0 = 76.80, KE(10k) = 1.757, KE(20k) = 0.1442 (VALIDATION.md:29-34).
* M1 (tests/m1_cavity.cpp:47-205): lid-driven cavity 128 x 128 x 4 at
  Re 1000, run to the reference's convergence rule; centreline u at the 15
  Ghia stations against the reference's own u_sim (VALIDATION.md:52-66) and
  Ghia's table within the reference's 7 % gate.
* M2 (tests/m2_cylinder.cpp:31-244): cylinder D = 40 at Re 150 in a
  560 x 400 x 4 tunnel; Strouhal number from the reference's Hann-windowed
  Goertzel scan of 3000 lift samples against its reported St = 0.2020
  (VALIDATION.md:16).
"""
import sys

import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd import solver as S

pytestmark = pytest.mark.gpu

NU_AIR = np.float32(1.48e-5)  # kNuAir, units.h:32
Q = S.Q


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def open_domain_flags(nx, ny, nz):
    """openDomainFlags (tests/lbm_test_util.h:34-54): slip y faces, then the
    inlet / outlet x faces over the corners, periodic z."""
    f = np.zeros((nz, ny, nx), np.uint8)
    f[:, 0, :] = S.SLIP_BOTTOM
    f[:, ny - 1, :] = S.SLIP_TOP
    f[:, :, 0] = S.INLET
    f[:, :, nx - 1] = S.OUTLET
    return f.reshape(-1)


def feq32(u, v):
    """Host equilibrium in FP32 at rho = 1 (lbm_test_util.h:56-75)."""
    one = np.float32(1)
    usq = np.float32(1.5) * (u * u + v * v)
    f = np.empty((Q, u.size), np.float32)
    for q in range(Q):
        cu = np.float32(3) * (np.float32(S.CX[q]) * u + np.float32(S.CY[q]) * v)
        f[q] = np.float32(S.W[q]) * (one + cu + np.float32(0.5) * cu * cu - usq)
    return f.reshape(-1)


def moments32(f, n):
    """momentsFromF (lbm_test_util.h:156-188): FP32 sums in q order."""
    g = f.reshape(Q, n).astype(np.float32)
    rho = np.zeros(n, np.float32)
    jx, jy, jz = (np.zeros(n, np.float32) for _ in range(3))
    for q in range(Q):
        rho += g[q]
        jx += np.float32(S.CX[q]) * g[q]
        jy += np.float32(S.CY[q]) * g[q]
        jz += np.float32(S.CZ[q]) * g[q]
    inv = np.where(rho > 1e-12, np.float32(1) / rho, np.float32(0)).astype(np.float32)
    return rho, jx * inv, jy * inv, jz * inv


def kinetic_energy(f, flags, n):
    rho, u, v, w = moments32(f, n)
    m = flags == S.FLUID
    return 0.5 * float(np.sum(rho[m].astype(np.float64) * (u[m].astype(np.float64) ** 2
                                                            + v[m].astype(np.float64) ** 2
                                                            + w[m].astype(np.float64) ** 2)))


def solver_for(dims, scaling, flags, storage="f32", ramp=True):
    s = S.LBMSolver()
    s.set_startup_ramp_enabled(ramp)
    s.init(dims, scaling, flags, storage=storage)
    return s


def m0_case():
    n = 64
    u_lat = np.float32(1e-3)
    airspeed = float(NU_AIR * u_lat * np.float32(n) / (np.float32(0.01) * np.float32(1.0)))
    sc = S.compute_scaling(1.0, airspeed, float(NU_AIR), n, float(u_lat))
    assert not sc.tau_clamped and abs(sc.nu_lat - 0.01) < 1e-5
    flags = open_domain_flags(n, n, n)
    k = np.float32(2.0 * 3.14159265358979 / n)
    z, y, x = (a.reshape(-1).astype(np.float32) for a in np.meshgrid(np.arange(n), np.arange(n), np.arange(n),
                                                                     indexing="ij"))
    a = np.float32(0.05)
    u = a * np.sin(k * x) * np.cos(k * y) * np.cos(k * z)
    v = -a * np.cos(k * x) * np.sin(k * y) * np.cos(k * z)
    return n, sc, flags, feq32(u.astype(np.float32), v.astype(np.float32))


def test_m0_taylor_green_against_the_oracle(gpu):
    """The M0 set-up for 1000 steps against the oracle run with the
    reference's buffer contents: the uploaded field in the active buffer,
    the rest state of the cold start in the other one (the outlet corner
    cells copy slip cells, whose storage is never written)."""
    from oracle import oracle as orc
    n, sc, flags, f0 = m0_case()
    say("case: m0 Taylor-Green 64^3 open domain, f32, 1000 steps vs the FP64 oracle")
    s = solver_for((n, n, n), sc, flags, "f32")
    s.upload_f(f0)
    s.step_n(1000)
    s.synchronize()
    got = s.download_f()
    s.close()
    # the upload ends the startup ramps, as the reference's uploadF does
    # (notifySnapshotRestored, tests/lbm_test_util.h:145); FP64 oracle: the
    # GPU's FP32 store of f - w carries far less rounding than the FP32
    # oracle's store of f, so the exact arithmetic is the bar (the FP32
    # oracle drifts 2e-4 from it here)
    taus, uins = orc.ramp_schedule(sc.tau, sc.nu_lat, sc.u_lat, n, 1000, ramp=False)
    f = f0.astype(np.float64)
    other = np.repeat(np.asarray(S.W, np.float64)[:, None], n ** 3, axis=1).reshape(-1).copy()
    orc.run(f, flags, (n, n, n), taus.astype(np.float64), uins.astype(np.float64), cs=0.1733, macro=False,
            other=other)
    fl = flags == S.FLUID
    mg, mo = moments32(got, n ** 3), moments32(f, n ** 3)
    e = float(np.linalg.norm(np.stack(mg[1:])[:, fl] - np.stack(mo[1:])[:, fl]) /
              np.linalg.norm(np.stack(mo[1:])[:, fl]))
    kg, ko = kinetic_energy(got, flags, n ** 3), kinetic_energy(f, flags, n ** 3)
    say(f"  KE(1000) gpu {kg:.6f} oracle {ko:.6f}; velocity rel L2 {e:.2e}")
    # FP32 arithmetic leaves ~1e-5; running the ramps, or holding the copied
    # upload in the never-written corner storage instead of the rest state,
    # moves KE(1000) by 7e-2 and 2.5e-3
    assert e < 1e-4 and abs(kg / ko - 1) < 1e-4


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_m0_taylor_green_energy_trace(gpu, storage):
    """The M0 run to 50 000 steps with the reference's gates (monotone decay
    above 0.2 % of KE0 within 1 % jitter, no NaN, final KE < 2 % of KE0) and
    its reported trace. KE0 is exact. During the viscous decay the run
    follows the reported trace within 3 %: the open domain's corner cells
    (outlets copying never-written slip storage, inlet corners) add a
    boundary circulation of about 0.05 lattice units, which is the whole
    plateau below 0.2 % of KE0 and a few percent of KE(10k); the oracle with
    the reference's current code gives 1.791 at 10k (1.773 without the eddy
    viscosity), so the reported 1.757 is held to that band, and the plateau
    to its order of magnitude."""
    n, sc, flags, f0 = m0_case()
    ke0 = kinetic_energy(f0, flags, n ** 3)
    say(f"case: m0 Taylor-Green 64^3 open domain, {storage}, KE0 {ke0:.4f}")
    assert abs(ke0 - 76.80) < 0.005
    s = solver_for((n, n, n), sc, flags, storage)
    s.upload_f(f0)
    reported = {10000: 1.757, 20000: 0.1442, 30000: 0.08533, 40000: 0.09143, 50000: 0.09413}
    got = {}
    prev = ke0
    for step in range(1000, 50001, 1000):
        s.step_n(1000)
        s.synchronize()
        assert not s.nan_detected()
        ke = kinetic_energy(s.download_f(), flags, n ** 3)
        if prev > 2e-3 * ke0:
            assert ke <= prev * 1.01  # the reference's monotonicity gate
        prev = ke
        if step in reported:
            got[step] = ke
            say(f"step {step}: KE {ke:.5g} (reference {reported[step]})")
    s.close()
    assert prev < 0.02 * ke0  # the reference's decay gate
    assert abs(got[10000] / 1.757 - 1) < 3e-2
    for st in (20000, 30000, 40000, 50000):
        assert 0.3 < got[st] / reported[st] < 3.0


GHIA = [(0.0547, -0.18109), (0.0625, -0.20196), (0.0703, -0.22220), (0.1016, -0.29730),
        (0.1719, -0.38289), (0.2813, -0.27805), (0.4531, -0.10648), (0.5000, -0.06080),
        (0.6172, 0.05702), (0.7344, 0.18719), (0.8516, 0.33304), (0.9531, 0.46604),
        (0.9609, 0.51117), (0.9688, 0.57492), (0.9766, 0.65928)]   # m1_cavity.cpp:47-53
U_SIM_REFERENCE = [-0.16878, -0.18871, -0.20824, -0.28267, -0.37545, -0.27617, -0.10656, -0.06165,
                   0.05435, 0.18289, 0.32702, 0.46958, 0.51818, 0.58573, 0.67135]  # VALIDATION.md:52-66


def m1_case():
    nx = ny = 128
    nz = 4
    lid = np.float32(0.08)
    L = nx - 2
    sc = S.compute_scaling(1.0, float(np.float32(1000.0) * NU_AIR / np.float32(1.0)), float(NU_AIR), L,
                           float(lid))
    assert not sc.tau_clamped
    f = np.zeros((nz, ny, nx), np.uint8)
    f[:, ny - 1, :] = S.INLET   # lid first, walls over the corners
    f[:, 0, :] = S.SOLID
    f[:, :, 0] = S.SOLID
    f[:, :, nx - 1] = S.SOLID
    return (nx, ny, nz), sc, f.reshape(-1), lid, L


def test_m1_cavity_against_the_oracle(gpu):
    """The M1 cavity for 4000 steps from rest against the FP64 oracle."""
    from oracle import oracle as orc
    dims, sc, flags, lid, L = m1_case()
    n = int(np.prod(dims))
    say("case: m1 cavity 128x128x4 Re 1000, f32, 4000 steps vs the FP64 oracle")
    rest = np.repeat(np.asarray(S.W, np.float64)[:, None], n, axis=1).reshape(-1)
    s = solver_for(dims, sc, flags, "f32")
    s.upload_f(rest.astype(np.float32))
    s.step_n(4000)
    s.synchronize()
    _, u, v, _ = s.read_macro()
    s.close()
    # no ramps after the upload (lid at full speed from the first step); the
    # FP32 oracle drifts 8e-4 from the FP64 one here
    taus, uins = orc.ramp_schedule(sc.tau, sc.nu_lat, sc.u_lat, dims[0], 4000, ramp=False)
    f = rest.copy()
    _, uo, vo, _ = orc.run(f, flags, dims, taus.astype(np.float64), uins.astype(np.float64), cs=0.1733,
                           other=rest.copy())
    fl = flags == S.FLUID
    e = float(np.linalg.norm(np.stack([u, v])[:, fl] - np.stack([uo, vo])[:, fl]) /
              np.linalg.norm(np.stack([uo, vo])[:, fl]))
    say(f"  velocity rel L2 vs the FP64 oracle {e:.2e}")
    assert e < 1e-4   # the ramped lid would give 2.4e-2


def test_m1_lid_driven_cavity_ghia(gpu):
    """To the reference's convergence rule, then the 15 Ghia stations: within
    the reference's 7 % gate of Ghia's table, and within 1 % of the lid speed
    of the u_sim the reference reports. That report predates the current
    reference code: the oracle restating the current code, run the same way
    on the CPU, converges at the same step 68 000 and differs from it by up
    to 0.53 % of the lid speed (and from this run by 0.05 %)."""
    dims, sc, flags, lid, L = m1_case()
    nx, ny, nz = dims
    n = nx * ny * nz
    s = solver_for(dims, sc, flags, "f32")
    s.upload_f(np.repeat(np.asarray(S.W, np.float32)[:, None], n, axis=1).reshape(-1))

    def centreline(fpop):
        _, u, _, _ = moments32(fpop, n)
        u = u.reshape(nz, ny, nx)
        return u[:, 1:ny - 1, nx // 2 - 1: nx // 2 + 1].astype(np.float64).sum(axis=(0, 2)) / (2 * nz) / float(lid)

    prev, prof, steps = None, None, 0
    while steps < 240000:
        s.step_n(4000)
        steps += 4000
        s.synchronize()
        assert not s.nan_detected()
        prof = centreline(s.download_f())
        if prev is not None and steps >= 40000 and np.max(np.abs(prof - prev)) < 5e-4:
            break
        prev = prof
    s.close()
    say(f"case: m1 cavity 128x128x4 Re 1000 converged at step {steps} (reference: 68000)")

    def at(yn):
        sidx = yn * L + 0.5 - 1.0
        lo = int(np.floor(sidx))
        t = sidx - lo
        return prof[lo] * (1 - t) + prof[lo + 1] * t

    dev_ref = []
    for (yn, ug), ur in zip(GHIA, U_SIM_REFERENCE):
        us = at(yn)
        dev_ref.append(abs(us - ur))
        say(f"y {yn:.4f}: u {us:+.5f}  reference {ur:+.5f}  Ghia {ug:+.5f}")
        assert abs(us - ug) <= 0.07          # the reference's gate
    say(f"max |u - u_reference| = {max(dev_ref):.2e} of the lid speed")
    assert steps == 68000
    assert max(dev_ref) < 1e-2


def goertzel_power(x, cycles):
    w = 2.0 * np.pi * cycles
    coeff = 2.0 * np.cos(w)
    s1 = s2 = 0.0
    for v in x:
        s0 = v + coeff * s1 - s2
        s2, s1 = s1, s0
    return s1 * s1 + s2 * s2 - coeff * s1 * s2


def test_m2_cylinder_strouhal(gpu):
    nx, ny, nz = 560, 400, 4
    d_nom, cx, cy, u_lat = 40.0, 160.0, 200.0, np.float32(0.08)
    sc = S.compute_scaling(1.0, float(np.float32(150.0) * NU_AIR / np.float32(1.0)), float(NU_AIR), int(d_nom),
                           float(u_lat))
    assert not sc.tau_clamped
    flags = open_domain_flags(nx, ny, nz).reshape(nz, ny, nx)
    jj, ii = np.meshgrid(np.arange(ny, dtype=np.float32), np.arange(nx, dtype=np.float32), indexing="ij")
    disk = (ii - np.float32(cx)) ** 2 + (jj - np.float32(cy)) ** 2 <= np.float32(0.25 * d_nom * d_nom)
    flags[:, disk] = S.SOLID
    row = np.nonzero(disk[int(cy)])[0]
    d_meas = float(row.max() - row.min() + 1)
    flags = flags.reshape(-1)
    n = nx * ny * nz
    s = solver_for((nx, ny, nz), sc, flags, "f32", ramp=False)
    solid = np.tile(disk.reshape(-1), nz)
    i = np.tile(ii.reshape(-1), nz)
    j = np.tile(jj.reshape(-1), nz)
    two_pi = np.float32(2.0 * 3.14159265358979)
    u = np.where(solid, np.float32(0), u_lat).astype(np.float32)
    v = np.where(solid, np.float32(0), np.float32(0.1) * u_lat * np.sin(two_pi * i / np.float32(nx))
                 * np.sin(np.float32(3.14159265) * j / np.float32(ny))).astype(np.float32)
    s.upload_f(feq32(u, v))
    for _ in range(25):
        s.step_n(1000)
    s.synchronize()
    assert not s.nan_detected()
    lift = np.empty(3000)
    for k in range(3000):
        s.step_n(16)
        lift[k] = s.force_now()[1]
    assert not s.nan_detected()
    s.close()
    centred = lift - lift.mean()
    assert centred.std() > 1e-3
    hann = 0.5 * (1.0 - np.cos(2.0 * np.pi * np.arange(3000) / 2999.0))
    sig = centred * hann
    sts = 0.10 + 0.0005 * np.arange(401)
    power = np.array([goertzel_power(sig, st * float(u_lat) / d_meas * 16.0) for st in sts])
    best = float(sts[int(np.argmax(power))])
    ratio = float(power.max() / power.mean())
    say(f"case: m2 cylinder D {d_meas:.0f} Re 150: St {best:.4f} (reference 0.2020), peak {ratio:.1f}x scan mean")
    assert ratio > 5.0 and 0.17 <= best <= 0.21   # the reference's gates
    assert abs(best - 0.2020) <= 0.0025          # within five scan steps of the reference's peak
