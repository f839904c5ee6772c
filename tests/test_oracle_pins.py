"""Closed-form pins for the oracle functions that conservation and fixed
points do not see: the Smagorinsky effective relaxation time, the TRT odd
rate (magic parameter 3/16), the speed cap and density floor, and q-LIBB as
the reference writes it. Each check compares the oracle's output with what
the model's definition fixes, computed independently of the code's own
formula, and each is chosen so that a plausible slip (a wrong prefactor, a
dropped factor 2 on the off-diagonal stresses, a wrong odd-rate expression, a
cap applied per component, a swapped q-LIBB operand) fails it. Where a test
shows that it has teeth, the mistake is planted in the oracle's inputs, on
the CPU.

Reference lines: Smagorinsky closure lbm_core.cu:420-444 with the prefactor
18 sqrt(2) Cs^2 of lbm_core.cu:785-787; TRT rates lbm_core.cu:446-452; density
floor lbm_core.cu:393-398; speed cap lbm_core.cu:400-415 (kMaxSimSpeed 0.30,
lbm_core.cuh:80); q-LIBB lbm_core.cu:338-364.
"""
import numpy as np
import pytest

from oracle import oracle as orc

Q, CX, CY, CZ, W = orc.Q, orc.CX, orc.CY, orc.CZ, orc.W
OPP = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]
FLUID, SOLID, INLET, OUTLET = 0, 1, 2, 3


def eq_field(rho, u, v, w):
    usq = 1.5 * (u * u + v * v + w * w)
    f = np.empty((Q, np.size(rho)))
    for q in range(Q):
        cu = 3.0 * (CX[q] * u + CY[q] * v + CZ[q] * w)
        f[q] = W[q] * rho * (1 + cu + 0.5 * cu * cu - usq)
    return f.reshape(-1)


def moments(f, n):
    g = f.reshape(Q, n)
    rho = g.sum(0)
    return rho, (CX[:, None] * g).sum(0), (CY[:, None] * g).sum(0), (CZ[:, None] * g).sum(0)


# ---------------------------------------------------------------- Smagorinsky
def couette(cs, tau0=0.52, ny=24, u_lid=0.1, steps=60000):
    """Steady plane Couette flow: x / z periodic (one cell each, the flow is
    invariant along them), a bounce-back wall row at y = 0 and the reference's
    moving-lid Inlet row at y = ny - 1 (tests/m1_cavity.cpp:55-79 drives its
    lid the same way). Returns the bulk shear rate (a least-squares slope of
    u(y) over rows away from both walls), the wall shear stress from the
    momentum exchange and the mean density."""
    d = (1, ny, 1)
    flags = np.zeros(ny, dtype=np.uint8)
    flags[0] = SOLID
    flags[ny - 1] = INLET
    y = np.arange(ny, dtype=np.float64)
    u0 = np.clip(u_lid * (y - 0.5) / (ny - 1.5), 0.0, u_lid)  # start near the answer
    f = eq_field(np.ones(ny), u0, np.zeros(ny), np.zeros(ny))
    _, u, _, _ = orc.run(f, flags, d, np.full(steps, tau0), np.full(steps, u_lid), cs=cs,
                         periodic_xy=True)
    rows = np.arange(3, ny - 4)
    gamma = np.polyfit(rows, u[rows], 1)[0]
    rho, _, _, _ = moments(f, ny)
    fx = orc.force(f, flags, d, periodic_xy=True)[0]  # wall area 1 x 1
    return gamma, fx, float(rho[rows].mean()), u


@pytest.mark.parametrize("cs", [0.0, 0.5, 1.0])
def test_smagorinsky_couette_effective_viscosity(cs):
    """In steady linear shear the momentum flux through every plane is the
    wall shear stress sigma = rho nu_eff gamma. The Smagorinsky model sets
    nu_eff = nu_0 + (Cs Delta)^2 |S| with |S| = sqrt(2 S:S) = gamma for simple
    shear, Delta = 1 cell, nu_0 = (tau_0 - 1/2) / 3. The oracle's wall force
    and bulk shear rate must satisfy that, whatever closed form the code uses
    to reach it (Hou et al. 1996)."""
    tau0 = 0.52
    gamma, sigma, rho, u = couette(cs, tau0)
    nu_meas = sigma / (rho * gamma)
    nu_model = (tau0 - 0.5) / 3.0 + cs * cs * gamma
    # the profile is linear in the bulk (constant eddy viscosity)
    rows = np.arange(3, 20)
    resid = u[rows] - np.polyval(np.polyfit(rows, u[rows], 1), rows)
    assert np.max(np.abs(resid)) < 1e-6
    assert abs(nu_meas / nu_model - 1) < 2e-3, (cs, nu_meas, nu_model)
    if cs == 1.0:  # the eddy viscosity is a large part of the total
        assert cs * cs * gamma > 0.5 * (tau0 - 0.5) / 3.0


def test_smagorinsky_pin_has_teeth():
    """A prefactor of 18 Cs^2 (the sqrt 2 dropped), planted through the
    constant the oracle receives, or the off-diagonal stresses counted once
    (the same as Cs^2 / sqrt 2 for pure shear), moves the measured viscosity
    off the model by several percent."""
    tau0, cs = 0.52, 1.0
    for planted in (cs / 2 ** 0.25, cs * 2 ** -0.25):
        gamma, sigma, rho, _ = couette(planted, tau0)
        nu_meas = sigma / (rho * gamma)
        nu_model = (tau0 - 0.5) / 3.0 + cs * cs * gamma
        assert abs(nu_meas / nu_model - 1) > 2e-2


# ------------------------------------------------------------------- TRT odd
def relaxation_rates(tau, lam, seed=5):
    """Rates at which one collision relaxes an even and an odd
    non-equilibrium mode of a planted uniform state (a periodic 2^3 box:
    streaming leaves a uniform state unchanged). The modes carry no mass and
    no momentum, so the equilibrium is that of the base state and
    f* - feq = (1 - w+) E + (1 - w-) O exactly for a two-relaxation-time
    collision; E and O are orthogonal, so each rate is one projection."""
    rng = np.random.default_rng(seed)
    base = eq_field(np.array([1.02]), np.array([0.04]), np.array([-0.03]), np.array([0.02]))
    r = rng.random(Q) - 0.5
    ev = 0.5 * (r + r[OPP])
    ev -= ev.mean()                        # no mass
    od = 0.5 * (r - r[OPP])
    for c in (CX, CY, CZ):                 # no momentum (stays odd)
        od -= (od @ c) / (c @ c) * c
    ev *= 1e-4 / np.abs(ev).max()
    od *= 1e-4 / np.abs(od).max()
    n = 8
    f = np.repeat((base + ev + od)[:, None], n, axis=1).reshape(-1)
    orc.run(f, np.zeros(n, np.uint8), (2, 2, 2), np.full(1, tau), np.zeros(1), lam=lam, cs=0.0,
            periodic_xy=True)
    dev = f.reshape(Q, n)[:, 0] - base
    return 1 - (dev @ ev) / (ev @ ev), 1 - (dev @ od) / (od @ od), dev - (dev @ ev) / (ev @ ev) * ev \
        - (dev @ od) / (od @ od) * od


def test_trt_odd_rate_is_the_magic_parameter():
    """Two-relaxation-time collision (Ginzburg 2005): even modes relax at
    w+ = 1/tau, odd modes at w- with (1/w+ - 1/2)(1/w- - 1/2) = Lambda, the
    magic parameter, 3/16 in the reference. Measured on planted modes at
    several tau, so a single-rate (BGK) collision, an odd rate from
    Lambda / tau+ instead of Lambda / (tau+ - 1/2), or a swapped pair fails.

    The reference's boundary vocabulary cannot drive the Poiseuille flow
    whose wall position would show Lambda too: it has no body force, its
    equilibrium inlet pins rho = 1 so a no-slip channel behind it stagnates,
    and an equilibrium lid relieves the return-flow pressure of a shallow
    cavity (both checked on this oracle while writing this test)."""
    for tau in (0.5005, 0.52, 0.8, 1.5):
        wp, wm, rest = relaxation_rates(tau, 3.0 / 16.0)
        assert abs(1 / wp - tau) < 1e-9 * tau
        lam = (1 / wp - 0.5) * (1 / wm - 0.5)
        assert abs(lam - 3.0 / 16.0) < 1e-7, (tau, lam)
        assert np.abs(rest).max() < 1e-15     # nothing but the two modes moves
    # the measurement sees the planted parameter, not a constant
    wp, wm, _ = relaxation_rates(0.8, 0.25)
    assert abs((1 / wp - 0.5) * (1 / wm - 0.5) - 0.25) < 1e-7


# ----------------------------------------------------- speed cap, rho floor
def one_step_uniform(rho, u, tau=1.0, lam=0.25):
    """A uniform periodic 2^3 box (streaming leaves a uniform state as it is)
    with BGK-equivalent rates tau+ = tau- = 1 (Lambda = 1/4 at tau 1): one
    step replaces the populations by the equilibrium the collision targets."""
    d = (2, 2, 2)
    n = 8
    f = eq_field(np.full(n, rho), np.full(n, u[0]), np.full(n, u[1]), np.full(n, u[2]))
    mrho, mu, mv, mw = orc.run(f, np.zeros(n, np.uint8), d, np.full(1, tau), np.zeros(1), lam=lam,
                               cs=0.0, periodic_xy=True)
    r, jx, jy, jz = moments(f, n)
    return (mrho[0], np.array([mu[0], mv[0], mw[0]])), (r[0], np.array([jx[0], jy[0], jz[0]])), f


def test_speed_cap_rescales_the_vector_onto_0_30():
    u = np.array([0.40, -0.30, 0.12])  # |u| = 0.514
    (mr, mu), (r, j), _ = one_step_uniform(1.0, u)
    cap = float(np.float32(0.30)) * u / np.linalg.norm(u)  # kMaxSimSpeed is the FP32 constant 0.30f
    assert abs(mr - 1.0) < 1e-14
    assert np.allclose(mu, cap, atol=1e-14, rtol=0)        # direction kept, |u| = 0.30
    assert np.allclose(j, 1.0 * cap, atol=1e-14, rtol=0)   # relaxed onto feq(rho, capped u)
    assert abs(r - 1.0) < 1e-14
    # below the cap nothing happens
    u2 = np.array([0.2, 0.1, -0.15])  # |u| = 0.269
    (_, mu2), (_, j2), _ = one_step_uniform(1.0, u2)
    assert np.allclose(mu2, u2, atol=1e-14) and np.allclose(j2, u2, atol=1e-14)


def test_density_floor_keeps_momentum_and_raises_mass():
    u = np.array([0.1, 0.0, -0.05])
    (mr, mu), (r, j), _ = one_step_uniform(0.02, u)
    j_in = 0.02 * u
    floor = float(np.float32(0.05))                      # the FP32 constant 0.05f
    assert abs(mr - floor) < 1e-15                       # rho = max(rho, 0.05)
    assert np.allclose(mu, j_in / floor, atol=1e-15)     # u = j / floored rho
    assert abs(r - floor) < 1e-15                        # feq(0.05, u) carries 0.05
    assert np.allclose(j, j_in, atol=1e-15)              # and the same momentum
    (mr2, _), (r2, _), _ = one_step_uniform(0.06, u)     # above the floor: untouched
    assert abs(mr2 - 0.06) < 1e-15 and abs(r2 - 0.06) < 1e-15


# ------------------------------------------------------------------ q-LIBB
def pulled_population(byte, ff_bit, q, seed=3):
    """Population q pulled by fluid cell x = (0, 1, 1) of a periodic 3^3 box
    whose centre (1, 1, 1) is solid, with c_q = (-1, 0, 0) so that x - c_q is
    the solid centre. One step from a random planted post-collision state;
    the population is recovered exactly from the cell's density: only the q
    link differs between a run with the cut byte and a plain bounce-back run."""
    d = (3, 3, 3)
    n = 27
    flags = np.zeros(n, np.uint8)
    centre = 1 + 3 * (1 + 3 * 1)
    flags[centre] = SOLID
    rng = np.random.default_rng(seed)
    f0 = W[:, None] * (1 + 0.2 * (rng.random((Q, n)) - 0.5))
    x = 0 + 3 * (1 + 3 * 1)
    xf = 2 + 3 * (1 + 3 * 1)  # x + c_q, wrapped
    qfrac = np.zeros(n * Q, np.uint8)
    ffmask = np.zeros(n, np.uint32)
    qfrac[x * Q + q] = byte
    if ff_bit:
        ffmask[x] = 1 << q

    def rho_after(qf, ff):
        f = f0.reshape(-1).copy()
        r, _, _, _ = orc.run(f, flags, d, np.full(1, 0.8), np.zeros(1), cs=0.0, periodic_xy=True,
                             qfrac=qf, ffmask=ff)
        return r[x]

    plain = rho_after(np.zeros(n * Q, np.uint8), np.zeros(n, np.uint32))
    got = rho_after(qfrac, ffmask) - plain + f0[OPP[q], x]  # half-way value is f*_opp(q)(x)
    return got, f0[q, x], f0[OPP[q], x], f0[q, xf]


def test_qlibb_special_cases_as_written():
    """The reference's q-LIBB at its special points (the cut fraction is
    b * (1/255) rounded in FP32, as the reference computes it, hence the
    1e-4 relative slack on the interpolation weights)."""
    q = next(i for i in range(1, Q) if (CX[i], CY[i], CZ[i]) == (-1, 0, 0))
    # byte 0: no link, exactly the half-way value
    got, fq, fopp, fff = pulled_population(0, True, q)
    assert abs(got - fopp) < 1e-15
    # q = 1 (byte 255): the midpoint of f*_q(x) and the half-way value
    got, fq, fopp, fff = pulled_population(255, False, q)
    assert abs(got - 0.5 * (fq + fopp)) < 1e-14
    # q just above 1/2 (byte 128): almost exactly f*_q(x) -- the population
    # leaving the wall, not f*_opp(q)(x) as in Bouzidi's scheme
    got, fq, fopp, fff = pulled_population(128, False, q)
    assert abs(got - fq) <= abs(fq - fopp) * (1 - 255 / 256) * (1 + 1e-4) + 1e-15
    assert abs(got - fopp) > 0.9 * abs(fq - fopp)
    # q -> 0 with a two-node mask (byte 1): almost exactly f*_q(x + c_q)
    got, fq, fopp, fff = pulled_population(1, True, q)
    assert abs(got - fff) <= abs(fq - fff) * 2 / 255 * (1 + 1e-4) + 1e-15
    # q < 1/2 without the two-node mask: half-way
    got, fq, fopp, fff = pulled_population(1, False, q)
    assert abs(got - fopp) < 1e-15
    # q just below 1/2 with the mask (byte 127): 2q f*_q(x) + (1 - 2q) f*_q(x + c_q)
    got, fq, fopp, fff = pulled_population(127, True, q)
    assert abs(got - fq) <= abs(fq - fff) * (1 - 254 / 255) * (1 + 1e-4) + 1e-15
