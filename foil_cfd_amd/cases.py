"""Case builders for the FoilCFD workloads (flags, scaling, initial states).

* Taylor-Green vortex box: fully periodic N^3, rho = 1,
  u = (u0 sin kx cos ky cos kz, -u0 cos kx sin ky cos kz, 0), k = 2 pi / N,
  populations at local equilibrium (reference tests/m0_taylor_green.cpp:57-80
  seeds its open-domain variant the same way).
* Wind tunnel around an extruded NACA 2412: buildCleanFoilFlags with the
  default fractional anchors (0.3 nx, 0.5 ny), reference voxelizer.h:57-85.
* The 23.6 M-cell wing: chord 256, 512 x 192 x 240, chord 1.2 m at 30 m/s in
  air (tau clamped at 0.5005, units.h:192-219).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import solver as S


@dataclass
class Case:
    name: str
    dims: tuple
    flags: np.ndarray
    scaling: object
    periodic_xy: bool = False
    cs: float = 0.1733
    ramp: bool = True
    f0: np.ndarray | None = None  # optional canonical initial state (FP64)


def equilibrium_field(rho, u, v, w) -> np.ndarray:
    """Canonical SoA feq (FP64) of per-cell (rho, u, v, w) arrays."""
    n = rho.size
    usq = 1.5 * (u * u + v * v + w * w)
    f = np.empty((S.Q, n), dtype=np.float64)
    for q in range(S.Q):
        cu = 3.0 * (S.CX[q] * u + S.CY[q] * v + S.CZ[q] * w)
        f[q] = S.W[q] * rho * (1.0 + cu + 0.5 * cu * cu - usq)
    return f.reshape(-1)


def tgv_fields(n: int, u0: float = 0.05):
    k = 2.0 * np.pi / n
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x = x.reshape(-1).astype(np.float64)
    y = y.reshape(-1).astype(np.float64)
    z = z.reshape(-1).astype(np.float64)
    u = u0 * np.sin(k * x) * np.cos(k * y) * np.cos(k * z)
    v = -u0 * np.cos(k * x) * np.sin(k * y) * np.cos(k * z)
    w = np.zeros_like(u)
    return np.ones_like(u), u, v, w


def taylor_green(n: int = 64, u0: float = 0.05, nu: float = 0.01) -> Case:
    tau = 3.0 * nu + 0.5
    sc = S.scaling_for_tau(tau, u0, n)
    rho, u, v, w = tgv_fields(n, u0)
    return Case(f"tgv{n}", (n, n, n), np.zeros(n ** 3, dtype=np.uint8), sc, periodic_xy=True,
                cs=0.0, ramp=False, f0=equilibrium_field(rho, u, v, w))


def tgv_kinetic_energy(u, v, w) -> float:
    return 0.5 * float(np.sum(u.astype(np.float64) ** 2 + v.astype(np.float64) ** 2
                              + w.astype(np.float64) ** 2))


def naca_tunnel(nx: int, ny: int, nz: int, chord: int, aoa: float = 8.0, u_lat: float = 0.08,
                cs: float = 0.1, chord_m: float = 1.2, airspeed: float = 30.0,
                nu: float = 1.48e-5, digits: str = "2412", ramp: bool = True,
                tau: float | None = None) -> Case:
    """Wind tunnel around an extruded NACA section. With ``tau`` the physical
    scaling is replaced by one whose lattice viscosity gives exactly that
    relaxation time (a laminar, low-Reynolds variant of the same geometry)."""
    pts, _ = S.generate_naca4(digits, 200)
    layout = S.make_layout(nx, ny, nz, chord)
    flags = S.build_clean_foil_flags(pts, aoa, layout)
    if tau is None:
        sc = S.compute_scaling(chord_m, airspeed, nu, chord, u_lat)
    else:
        sc = S.scaling_for_tau(tau, u_lat, chord)
    return Case(f"naca{digits}_{nx}x{ny}x{nz}_c{chord}_a{aoa:g}", (nx, ny, nz), flags, sc,
                cs=cs, ramp=ramp)


def config2_tunnel(aoa: float = 8.0) -> Case:
    """384 x 128 x 64, chord 128, u_lat 0.08, Cs 0.1."""
    return naca_tunnel(384, 128, 64, 128, aoa=aoa, u_lat=0.08, cs=0.1)


def config3_wing(aoa: float = 8.0) -> Case:
    """512 x 192 x 240 (23.6 M cells), chord 256, TRT + LES (Cs 0.1733)."""
    return naca_tunnel(512, 192, 240, 256, aoa=aoa, u_lat=0.08, cs=0.1733)


def periodic_box(like: Case) -> Case:
    """An all-fluid, fully periodic box with the dimensions, scaling and LES
    constant of another case (no boundaries, no walls): the bulk cost of the
    step on that grid."""
    nx, ny, nz = like.dims
    return Case(f"periodic_box_{nx}x{ny}x{nz}", like.dims, np.zeros(nx * ny * nz, dtype=np.uint8), like.scaling,
                periodic_xy=True, cs=like.cs, ramp=False)


def refinement_patch(case: Case, factor: int = 2, upstream_c: float = 0.35, wake_c: float = 1.0,
                     below_c: float = 0.45, above_c: float = 0.45, digits: str = "2412",
                     aoa: float = 8.0):
    """Patch box and fine flags for a tunnel case the way the reference app
    derives them (derivePatchBox with the chord margins of ui.h:136-140 and
    clearance 3, main.cpp:590-598; buildFinePatchFlags for the fine level)."""
    nx, ny, nz = case.dims
    chord = case.scaling.chord_cells
    cells = lambda c: max(2, int(round(c * chord)))  # noqa: E731
    box = S.derive_patch_box(nx, ny, nz, case.flags, cells(upstream_c), cells(wake_c),
                             cells(below_c), cells(above_c))
    pts, _ = S.generate_naca4(digits, 200)
    layout = S.make_layout(nx, ny, nz, chord)
    fine = S.build_fine_patch_flags(pts, aoa, layout, box, factor)
    return (box.x0, box.y0, box.x1, box.y1), fine


def finer_patch(case: Case, box, factor: int, finer_box, finer_factor: int = 2, digits: str = "2412",
                aoa: float = 8.0):
    """Flags of a nested third level: finer_box (x0, y0, x1, y1) in cells of the
    fine grid of `box` at `factor`, voxelized from the fine level's layout
    (the reference's nested VG box, main.cpp:664-745)."""
    nx, ny, nz = case.dims
    layout = S.make_layout(nx, ny, nz, case.scaling.chord_cells)
    fine_layout = S.make_fine_layout(layout, S.FcPatchBox(*box), factor)
    pts, _ = S.generate_naca4(digits, 200)
    return S.build_fine_patch_flags(pts, aoa, fine_layout, S.FcPatchBox(*finer_box), finer_factor)


def presolve_case(case: Case, chord: int | None = None, aoa: float = 8.0, digits: str = "2412") -> Case:
    """The mesh-sequencing companion of a tunnel case (src/main.cpp:756-790):
    the same section and margins at `chord` cells per chord (default
    max(48, chord / 4)), grid scaled in proportion, same u_lat and physics."""
    nx, ny, nz = case.dims
    c0 = case.scaling.chord_cells
    nc = max(48, c0 // 4) if chord is None else int(chord)
    dims = tuple(max(3, int(round(n * nc / c0))) for n in (nx, ny, nz))
    pts, _ = S.generate_naca4(digits, 200)
    layout = S.make_layout(*dims, nc)
    flags = S.build_clean_foil_flags(pts, aoa, layout)
    sc = case.scaling
    pre = S.compute_scaling(sc.chord_m, sc.airspeed_ms, sc.nu_m2s, nc, sc.u_lat)  # same physical inputs
    return Case(f"{case.name}_pre{nc}", dims, flags, pre, cs=case.cs, ramp=True)


APP_PATCH_MARGINS_C = (0.20, 0.50, 0.10, 0.20)  # upstream, wake, below, above (src/app/ui.h:159-162)
DEFAULT_PARTICLES = 1_000_000                    # src/render/viz.h:28
PARTICLE_SEED = 0xF01DCAFE                       # src/render/viz.cpp:676


def default_patch(case: Case, aoa: float):
    """Patch box of the app's margins derived from the case's (possibly
    vane-merged) flags, and the clean foil flags of the fine level."""
    up, wake, below, above = APP_PATCH_MARGINS_C
    return refinement_patch(case, 2, upstream_c=up, wake_c=wake, below_c=below, above_c=above, aoa=aoa)


def default_tunnel(case: Case, storage: str, aoa: float, wall: str = "newton", particles: int = DEFAULT_PARTICLES,
                   fine_flags=None, coarse_flags=None, box=None):
    """The reference app's default tunnel on top of a wing case: a factor-2
    patch derived from the solid box with the app's margins (main.cpp:590-
    598), the clean foil flags as the surface references of both levels, the
    iMEM wall model on the wall cells of both levels and the tracer
    particles. fine_flags / coarse_flags replace the plain foil flags of a
    level (e.g. foil + vanes; the box then comes from the coarse ones, as
    the app derives it from the vane-merged flags, main.cpp:560-600).
    A given box (coarse cells) needs fine_flags. Returns (solver, box,
    refinement info)."""
    live = case if coarse_flags is None else Case(case.name, case.dims, coarse_flags, case.scaling,
                                                   case.periodic_xy, case.cs, case.ramp, case.f0)
    if box is None:
        box, fine = default_patch(live, aoa)
    else:  # a given box comes with its fine flags
        fine = fine_flags
    s = make_solver(live, storage)
    s.set_surface_reference(case.flags)
    s.init_refinement(box, 2, fine if fine_flags is None else fine_flags)
    s.set_refined_surface_reference(fine)
    s.set_wall_model(wall)
    if particles:
        s.particles_seed(particles, PARTICLE_SEED)
    return s, box, s.refinement_info()


def lowbias32(x):
    """The reference's integer hash (src/render/particles.cu:48-55) on uint32 arrays."""
    x = np.asarray(x, dtype=np.uint32).copy()
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def synthetic_vane_q(qfrac, ffmask, seed=0xC4A11E5, q_min=0.05):
    """Replaces the cut fraction of every vane link (nonzero bytes of the
    dense qfrac[cell*19 + q], ascending) by q ~ U[0, 1) from a fixed-seed
    lowbias32 stream, 24-bit uniform as particles.cu:57-60 draws it, then
    applies the densify rule of vg.cpp:380-392: q < q_min drops the link
    (byte 0, two-node bit cleared), else byte = lround(255 q), 0 -> 1.
    Returns (qfrac, ffmask, kept links)."""
    qf = np.array(qfrac, dtype=np.uint8, copy=True).reshape(-1)
    ff = np.array(ffmask, dtype=np.uint32, copy=True).reshape(-1)
    idx = np.nonzero(qf)[0]
    u = (lowbias32(np.uint32(seed) + np.arange(idx.size, dtype=np.uint32)) & np.uint32(0xFFFFFF)).astype(
        np.float32) * np.float32(1.0 / 16777216.0)
    keep = u >= np.float32(q_min)
    b = np.floor(u.astype(np.float64) * 255.0 + 0.5).astype(np.int64)  # lround of a positive value
    b[b == 0] = 1
    qf[idx] = np.where(keep, b, 0).astype(np.uint8)
    drop = idx[~keep]
    cell, q = drop // 19, drop % 19
    np.bitwise_and.at(ff, cell, ~(np.uint32(1) << q.astype(np.uint32)))
    return qf, ff, int(keep.sum())


def make_solver(case: Case, storage: str = "f32") -> S.LBMSolver:
    s = S.LBMSolver()
    s.set_startup_ramp_enabled(case.ramp)
    s.init(case.dims, case.scaling, case.flags, storage=storage, periodic_xy=case.periodic_xy)
    s.set_smagorinsky_cs(case.cs)
    if case.f0 is not None:
        s.upload_f(case.f0)
    return s
