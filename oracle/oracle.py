"""FoilCFD CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes front end of oracle/liboracle.so (the CPU restatement of the
reference step, lbm_oracle.cpp) and of oracle/_ref/libfcref.so (the
reference's own units.h + NACA generator, built only where /root/reference
exists). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import this module; the product never does.

Every function here is pinned by at least one known answer, closed form or
invariant (exact moments, conservation, planted solutions, the reference's
own test values and the compiled reference's golden vectors); those checks
live in tests/test_oracle.py and the other tests/test_*_oracle.py files.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libfcref.so")

Q = 19
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0], dtype=np.int64)
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1], dtype=np.int64)
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1], dtype=np.int64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)

_lib = None
_ref = None

_vp = C.c_void_p
_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)
_u32p = C.POINTER(C.c_uint32)
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"oracle not built: {ORACLE_LIB} (run `make -C oracle`)")
        L = C.CDLL(ORACLE_LIB)
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_half_to_float.restype = C.c_float
        L.orc_half_to_float.argtypes = [C.c_uint16]
        L.orc_equilibrium_f64.argtypes = [C.c_double] * 4 + [_dp]
        L.orc_equilibrium_f32.argtypes = [C.c_float] * 4 + [_fp]
        L.orc_init_equilibrium_f32.argtypes = [_fp, _u8p, C.c_int, C.c_int, C.c_int, C.c_float]
        L.orc_init_equilibrium_f64.argtypes = [_dp, _u8p, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_spanwise_perturb_f32.argtypes = [_fp, _u8p, C.c_int, C.c_int, C.c_int, C.c_float,
                                               C.c_uint32]
        L.orc_spanwise_perturb_f64.argtypes = [_dp, _u8p, C.c_int, C.c_int, C.c_int, C.c_float,
                                               C.c_uint32]
        L.orc_speck_phase_bits.restype = C.c_uint32
        L.orc_speck_phase_bits.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_smag_prefactor.restype = C.c_float
        L.orc_smag_prefactor.argtypes = [C.c_float]
        common = [_u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_stream_collide_f32.argtypes = [_fp, _fp] + common + [C.c_float] * 4 + [C.c_int] + \
            [_fp] * 4 + [_u16p] * 3 + [_u8p, _u32p]
        L.orc_stream_collide_f64.argtypes = [_dp, _dp] + common + [C.c_double] * 4 + [C.c_int] + \
            [_dp] * 4 + [_u16p] * 3 + [_u8p, _u32p]
        L.orc_run_f32.argtypes = [_fp, _fp] + common + [C.c_int, _fp, _fp, C.c_float, C.c_float] + \
            [_fp] * 4 + [_u16p] * 3 + [_u8p, _u32p]
        L.orc_run_f64.argtypes = [_dp, _dp] + common + [C.c_int, _dp, _dp, C.c_double,
                                                        C.c_double] + \
            [_dp] * 4 + [_u16p] * 3 + [_u8p, _u32p]
        L.orc_run_stored_f64.argtypes = [_dp, _dp] + common + [C.c_int, _dp, _dp, C.c_double,
                                                               C.c_double] + \
            [_dp] * 4 + [C.c_int, C.c_int, C.c_double, C.c_int]
        L.orc_f16c_encode.restype = C.c_uint16
        L.orc_f16c_encode.argtypes = [C.c_float]
        L.orc_f16c_decode.restype = C.c_float
        L.orc_f16c_decode.argtypes = [C.c_uint16]
        L.orc_regularize_fneq_f64.argtypes = [_dp]
        L.orc_set_shell_poison.argtypes = [C.c_int]
        i6 = [C.c_int] * 6
        L.orc_fine_fill_f64.argtypes = [_dp, _dp, _u8p, C.c_int, C.c_int, C.c_int, _u8p, _dp,
                                        C.c_int, C.c_int, C.c_int] + [C.c_int] * 5 + \
            [C.c_double, C.c_double, C.c_int]
        L.orc_fine_fill_f32.argtypes = [_fp, _fp, _u8p, C.c_int, C.c_int, C.c_int, _u8p, _fp,
                                        C.c_int, C.c_int, C.c_int] + [C.c_int] * 5 + \
            [C.c_float, C.c_float, C.c_int]
        L.orc_fine_restrict_f64.argtypes = [_dp, _u8p, C.c_int, C.c_int, C.c_int, _dp, _u8p,
                                            C.c_int, C.c_int, C.c_int] + [C.c_int] * 5 + \
            [C.c_double] + [_dp] * 4
        L.orc_fine_tau_for.restype = C.c_float
        L.orc_fine_tau_for.argtypes = [C.c_float, C.c_int]
        L.orc_run_refined_f64.argtypes = [_dp, _dp, _u8p, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                          _u8p] + [C.c_int] * 6 + [_dp, _dp, C.c_double,
                                                                   C.c_double] + [_dp] * 4
        L.orc_run_refined_slip_f64.argtypes = [_dp, _dp, _u8p, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                               _u8p] + [C.c_int] * 6 + [_dp, _dp, C.c_double,
                                                                        C.c_double] + [_dp] * 4 + [_u16p] * 6 + \
            [_u8p, _u32p, _u8p, _u32p]
        L.orc_run_refined3_f64.argtypes = [_dp, _dp, _u8p, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                           _u8p] + [C.c_int] * 5 + [_dp, _dp, _u8p] + [C.c_int] * 5 + \
            [C.c_int, _dp, _dp, C.c_double, C.c_double] + [_dp] * 4
        L.orc_run_refined_f32.argtypes = [_fp, _fp, _u8p, C.c_int, C.c_int, C.c_int, _fp, _fp,
                                          _u8p] + [C.c_int] * 6 + [_fp, _fp, C.c_float,
                                                                   C.c_float] + [_fp] * 4
        L.orc_force_f32.argtypes = [_fp] + common + [_u16p] * 3 + [_dp]
        L.orc_force_f64.argtypes = [_dp] + common + [_u16p] * 3 + [_dp]
        L.orc_watchdog_f32.restype = C.c_int
        L.orc_watchdog_f32.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_longlong, C.c_int]
        L.orc_reichardt_uplus.restype = C.c_float
        L.orc_reichardt_uplus.argtypes = [C.c_float]
        L.orc_solve_utau.restype = C.c_float
        L.orc_solve_utau.argtypes = [C.c_float] * 4
        L.orc_float_to_half.restype = C.c_uint16
        L.orc_float_to_half.argtypes = [C.c_float]
        _llp = C.POINTER(C.c_longlong)
        L.orc_wall_update_f32.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_int, _llp, _llp,
                                          _fp, _fp, _fp, _fp, _u32p, _fp, C.c_float, C.c_float,
                                          C.c_float, _u16p, _u16p, _u16p, _fp]
        L.orc_particles_seed.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32]
        L.orc_particles_advect.argtypes = [_fp, C.c_int, _fp, _fp, _fp, _u8p, C.c_int, C.c_int,
                                           C.c_int, C.c_float, C.c_float, C.c_float, C.c_uint32]
        L.orc_moments_f32.argtypes = [_fp, C.c_int, C.c_int, C.c_int] + [_dp] * 4
        L.orc_moments_f64.argtypes = [_dp, C.c_int, C.c_int, C.c_int] + [_dp] * 4
        L.orc_solve_utau_iters.restype = C.c_float
        L.orc_solve_utau_iters.argtypes = [C.c_float] * 4 + [C.c_int]
        L.orc_bf16_round.restype = C.c_float
        L.orc_bf16_round.argtypes = [C.c_float]
        _ip = C.POINTER(C.c_int)
        L.orc_mlp_forward.argtypes = [_fp, C.c_int, _fp, C.c_longlong, _dp]
        L.orc_mlp_gradient.restype = C.c_double
        L.orc_mlp_gradient.argtypes = [_fp, C.c_int, _fp, C.c_longlong, _fp, _ip, C.c_int, _dp]
        L.orc_mlp_adam.argtypes = [_fp, _fp, _fp, _dp] + [C.c_double] * 4 + [C.c_int]
        _lib = L
    return _lib


def ref():
    """The reference's own units.h / generateNACA4 (None when not built)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB):
            return None
        R = C.CDLL(REF_LIB)
        sc = [C.c_float, C.c_float, C.c_float, C.c_int, C.c_float]
        R.ref_compute_scaling.argtypes = sc + [_fp]
        R.ref_refined_scaling.argtypes = sc + [C.c_int, _fp]
        R.ref_ramped_tau.restype = C.c_float
        R.ref_ramped_tau.argtypes = sc + [C.c_longlong, C.c_int]
        R.ref_ramped_u.restype = C.c_float
        R.ref_ramped_u.argtypes = sc + [C.c_longlong, C.c_int]
        R.ref_max_effective_re.restype = C.c_float
        R.ref_max_effective_re.argtypes = [C.c_int, C.c_float]
        R.ref_force_to_coefficient.restype = C.c_float
        R.ref_force_to_coefficient.argtypes = sc + [C.c_float, C.c_int]
        R.ref_generate_naca4.restype = C.c_int
        R.ref_generate_naca4.argtypes = [C.c_char_p, C.c_int, _fp, C.c_int, C.POINTER(C.c_int)]
        _ref = R
    return _ref


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _ptr_for(a):
    return _p(a, _dp if a.dtype == np.float64 else _fp)


def equilibrium(rho, ux, uy, uz, f64=True):
    out = np.zeros(Q, dtype=np.float64 if f64 else np.float32)
    if f64:
        lib().orc_equilibrium_f64(rho, ux, uy, uz, _p(out, _dp))
    else:
        lib().orc_equilibrium_f32(rho, ux, uy, uz, _p(out, _fp))
    return out


def init_equilibrium(flags, dims, u_inlet, f64=False):
    nx, ny, nz = dims
    f = np.zeros(Q * nx * ny * nz, dtype=np.float64 if f64 else np.float32)
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    if f64:
        lib().orc_init_equilibrium_f64(_p(f, _dp), _p(fl, _u8p), nx, ny, nz, u_inlet)
    else:
        lib().orc_init_equilibrium_f32(_p(f, _fp), _p(fl, _u8p), nx, ny, nz, u_inlet)
    return f


def spanwise_perturb(f, flags, dims, amplitude, seed=0x5EED):
    nx, ny, nz = dims
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    if f.dtype == np.float64:
        lib().orc_spanwise_perturb_f64(_p(f, _dp), _p(fl, _u8p), nx, ny, nz, amplitude, seed)
    else:
        lib().orc_spanwise_perturb_f32(_p(f, _fp), _p(fl, _u8p), nx, ny, nz, amplitude, seed)


def smag_prefactor(cs):
    return lib().orc_smag_prefactor(cs)


def run(f, flags, dims, taus, uins, lam=3.0 / 16.0, cs=0.1733, periodic_xy=False,
        slip=None, qfrac=None, ffmask=None, macro=True, other=None):
    """Advance ``f`` (canonical SoA, modified in place) by len(taus) steps of the
    reference step sequence; returns (rho, u, v, w) of the last step.
    ``other`` is the initial content of the second ping-pong buffer (default:
    a copy of ``f``); only cells that never get written (solid, slip) keep
    it, and it is read only where an outlet copies such a cell. It is
    updated in place with the final other buffer when given."""
    nx, ny, nz = dims
    n = nx * ny * nz
    f64 = f.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    tmp = f.copy() if other is None else other
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    taus = np.ascontiguousarray(taus, dtype=dt)
    uins = np.ascontiguousarray(uins, dtype=dt)
    m = [np.zeros(n, dtype=dt) for _ in range(4)] if macro else [None] * 4
    uw = [None] * 3 if slip is None else [np.ascontiguousarray(s, dtype=np.uint16) for s in slip]
    pre = float(smag_prefactor(cs))
    fn = lib().orc_run_f64 if f64 else lib().orc_run_f32
    pt = _dp if f64 else _fp
    fn(_p(f, pt), _p(tmp, pt), _p(fl, _u8p), nx, ny, nz, int(periodic_xy), int(periodic_xy),
       len(taus), _p(taus, pt), _p(uins, pt), lam, pre,
       *[_p(a, pt) for a in m], *[_p(a, _u16p) for a in uw],
       _p(None if qfrac is None else np.ascontiguousarray(qfrac, np.uint8), _u8p),
       _p(None if ffmask is None else np.ascontiguousarray(ffmask, np.uint32), _u32p))
    return tuple(m)


STORAGE_FORMATS = {  # (stored mantissa bits, min normal exponent, deviation scale)
    "f32": (23, -126, 1.0),
    "f16": (10, -14, 32768.0),
    "bf16": (7, -126, 1.0),
    "f16c": (-1, 0, 1.0),  # the device's s1e4m11 encode/decode, bit for bit
}


def run_stored(f, flags, dims, taus, uins, fmt, lam=3.0 / 16.0, cs=0.1733, periodic_xy=False,
               other=None):
    """FP64 steps with every population rounded after each step as storage
    format ``fmt`` (a STORAGE_FORMATS key or a (mant, emin, scale) tuple)
    keeps it. ``f`` (FP64 canonical SoA) is advanced in place. ``other`` is
    the second buffer's initial content, as in run() (default: a copy of
    ``f``); it is rounded to the format too and updated in place."""
    nx, ny, nz = dims
    n = nx * ny * nz
    mant, emin, scale = STORAGE_FORMATS[fmt] if isinstance(fmt, str) else fmt
    assert f.dtype == np.float64
    tmp = np.empty_like(f) if other is None else other
    assert tmp.dtype == np.float64 and tmp.size == f.size
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    taus = np.ascontiguousarray(taus, dtype=np.float64)
    uins = np.ascontiguousarray(uins, dtype=np.float64)
    m = [np.zeros(n, dtype=np.float64) for _ in range(4)]
    lib().orc_run_stored_f64(_p(f, _dp), _p(tmp, _dp), _p(fl, _u8p), nx, ny, nz, int(periodic_xy),
                             int(periodic_xy), len(taus), _p(taus, _dp), _p(uins, _dp), lam,
                             float(smag_prefactor(cs)), *[_p(a, _dp) for a in m], int(mant),
                             int(emin), float(scale), int(other is not None))
    return tuple(m)


def force(f, flags, dims, periodic_xy=False, slip=None):
    nx, ny, nz = dims
    out = np.zeros(3, dtype=np.float64)
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    uw = [None] * 3 if slip is None else [np.ascontiguousarray(s, dtype=np.uint16) for s in slip]
    fn = lib().orc_force_f64 if f.dtype == np.float64 else lib().orc_force_f32
    fn(_ptr_for(f), _p(fl, _u8p), nx, ny, nz, int(periodic_xy), int(periodic_xy),
       *[_p(a, _u16p) for a in uw], _p(out, _dp))
    return out


def watchdog(f, dims, offset, flag_base=0):
    nx, ny, nz = dims
    return lib().orc_watchdog_f32(_p(np.ascontiguousarray(f, np.float32), _fp), nx, ny, nz,
                                  offset, flag_base)


def moments(f, dims):
    nx, ny, nz = dims
    n = nx * ny * nz
    out = [np.zeros(n) for _ in range(4)]
    fn = lib().orc_moments_f64 if f.dtype == np.float64 else lib().orc_moments_f32
    fn(_ptr_for(f), nx, ny, nz, *[_p(a, _dp) for a in out])
    return tuple(out)


def set_threads(n):
    lib().orc_set_num_threads(int(n))


def threads():
    return lib().orc_num_threads()


# ---- ramps, restated in FP32 for the step schedules the oracle runs with ----
def ramp_schedule(tau, nu_lat, u_lat, nx, steps, start=0, ramp=True):
    """Per-step (tau, u_in) of the solver's startup ramps (units.h:274-318)
    in FP32, as the solver computes them."""
    f32 = np.float32
    taus = np.empty(steps, dtype=np.float32)
    uins = np.empty(steps, dtype=np.float32)
    ramp_steps = 2 * nx
    vramp = int(f32(0.5) * f32(2.0) * f32(nx))
    for i in range(steps):
        s = start + i
        if not ramp or s >= ramp_steps:
            taus[i] = tau
        else:
            t = f32(s) / f32(ramp_steps)
            nu_now = f32(nu_lat) * (f32(4.0) + t * (f32(1.0) - f32(4.0)))
            taus[i] = f32(3.0) * nu_now + f32(0.5)
        if not ramp or vramp <= 0 or s >= vramp:
            uins[i] = u_lat
        else:
            t = f32(s) / f32(vramp)
            sm = t * t * (f32(3.0) - f32(2.0) * t)
            uins[i] = f32(u_lat) * sm
    return taus, uins


# ---- wall model ---------------------------------------------------------------
def reichardt_uplus(yp):
    return lib().orc_reichardt_uplus(yp)


def solve_utau(ut, ys, nu, seed=0.0):
    return lib().orc_solve_utau(ut, ys, nu, seed)


def solve_utau_iters(ut, ys, nu, seed=0.0, iters=6):
    """Newton on Reichardt's law with a given iteration count (6 = solve_utau)."""
    return lib().orc_solve_utau_iters(ut, ys, nu, seed, iters)


# ---- learned u_tau network (restatement of the GPU network) ------------------
MLP_NPARAMS = 1217
MLP_OFFSETS = {"w1": (0, (32, 3)), "b1": (96, (32,)), "w2": (128, (32, 32)),
               "b2": (1152, (32,)), "w3": (1184, (32,)), "b3": (1216, ())}


def bf16_round(x):
    return lib().orc_bf16_round(x)


def mlp_forward(params, x, emulate=True):
    """Network output for normalised features x (shape (n, 3)); FP64 (with
    the GPU's BF16 operand roundings when emulate)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    assert p.size == MLP_NPARAMS
    xs = np.ascontiguousarray(np.asarray(x, dtype=np.float32).T)
    n = xs.shape[1]
    y = np.zeros(n, dtype=np.float64)
    lib().orc_mlp_forward(_p(p, _fp), int(emulate), _p(xs, _fp), n, _p(y, _dp))
    return y


def mlp_gradient(params, x, t, idx, emulate=True):
    """(gradient of the batch-mean squared error, sum of squared errors)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    xs = np.ascontiguousarray(np.asarray(x, dtype=np.float32).T)
    tt = np.ascontiguousarray(t, dtype=np.float32)
    ii = np.ascontiguousarray(idx, dtype=np.int32)
    n = xs.shape[1]
    assert tt.size == n and ii.min() >= 0 and ii.max() < n
    g = np.zeros(MLP_NPARAMS, dtype=np.float64)
    sse = lib().orc_mlp_gradient(_p(p, _fp), int(emulate), _p(xs, _fp), n, _p(tt, _fp),
                                 ii.ctypes.data_as(C.POINTER(C.c_int)), ii.size, _p(g, _dp))
    return g, sse


def mlp_adam(params, m, v, grad, lr, step, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in place on float32 arrays (params, m, v)."""
    g = np.ascontiguousarray(grad, dtype=np.float64)
    for a in (params, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous and a.size == MLP_NPARAMS
    lib().orc_mlp_adam(_p(params, _fp), _p(m, _fp), _p(v, _fp), _p(g, _dp), lr, beta1, beta2, eps, step)


def float_to_half_bits(x):
    return lib().orc_float_to_half(x)


def regularize_fneq(fneq):
    a = np.ascontiguousarray(fneq, dtype=np.float64).copy()
    lib().orc_regularize_fneq_f64(_p(a, _dp))
    return a


def fine_tau_for(tau_c, m):
    return float(lib().orc_fine_tau_for(tau_c, m))


def fine_dims(box, cdims, m):
    x0, y0, x1, y1 = box
    return (m * (x1 - x0), m * (y1 - y0), m * cdims[2])


def fine_fill(c_t0, c_t1, cflags, cdims, fflags, ff, box, m, time_weight, neq_scale,
              full_volume=False):
    """coarseToFineFill (lbm_refine.cu:123-258) into the fine canonical array ``ff``."""
    f64 = ff.dtype == np.float64
    dt, pt = (np.float64, _dp) if f64 else (np.float32, _fp)
    fn = lib().orc_fine_fill_f64 if f64 else lib().orc_fine_fill_f32
    fd = fine_dims(box, cdims, m)
    a0 = np.ascontiguousarray(c_t0, dtype=dt)
    a1 = np.ascontiguousarray(c_t1, dtype=dt)
    fn(_p(a0, pt), _p(a1, pt), _p(np.ascontiguousarray(cflags, np.uint8), _u8p), *cdims,
       _p(np.ascontiguousarray(fflags, np.uint8), _u8p), _p(ff, pt), *fd, *box, m, time_weight,
       neq_scale, int(full_volume))


def set_shell_poison(on: bool) -> None:
    """Test switch of the shell fill: NaN into every shell population whose
    target cell lies outside the shell's interior (see lbm_oracle.cpp)."""
    lib().orc_set_shell_poison(int(on))


def fine_restrict(ff, fflags, cf, cflags, cdims, box, m, neq_scale_inv, macro=False):
    """fineToCoarseRestrict (lbm_refine.cu:319-429) into the coarse FP64 array ``cf``."""
    n = cdims[0] * cdims[1] * cdims[2]
    m4 = [np.zeros(n) for _ in range(4)] if macro else [None] * 4
    fd = fine_dims(box, cdims, m)
    lib().orc_fine_restrict_f64(_p(np.ascontiguousarray(ff, np.float64), _dp),
                                _p(np.ascontiguousarray(fflags, np.uint8), _u8p), *fd, _p(cf, _dp),
                                _p(np.ascontiguousarray(cflags, np.uint8), _u8p), *cdims, *box, m,
                                neq_scale_inv, *[_p(a, _dp) for a in m4])
    return tuple(m4) if macro else None


def seed_fine(cf, cflags, cdims, fflags, box, m, tau_c, f64=True):
    """The fine level's cold-start state (seedFineFromCoarse,
    lbm_solver.cpp:502-521): rest equilibrium everywhere, then the
    full-volume fill from the current coarse state at weight 0."""
    fd = fine_dims(box, cdims, m)
    ff = init_equilibrium(np.zeros(fd[0] * fd[1] * fd[2], np.uint8), fd, 0.0, f64=f64)
    tau_f = fine_tau_for(tau_c, m)
    neq = np.float32(tau_f) / (np.float32(m) * np.float32(tau_c))
    fine_fill(cf, cf, cflags, cdims, fflags, ff, box, m, 0.0, float(neq), full_volume=True)
    return ff


def run_refined(cf, ff, cflags, fflags, cdims, box, m, taus, uins, lam=3.0 / 16.0, cs=0.1733,
                coarse_slip=None, fine_slip=None, coarse_q=None, fine_q=None):
    """len(taus) refined steps (coarse + m fine sub-steps + restriction,
    lbm_solver.cpp:1144-1290); cf / ff are advanced in place. Returns the
    coarse (rho, u, v, w) of the last step. coarse_slip / fine_slip: optional
    wall-model slip fields (3 half-bit arrays per level); coarse_q / fine_q:
    optional q-LIBB fields (qfrac, ffmask) per level (FP64 runs only)."""
    nx, ny, nz = cdims
    n = nx * ny * nz
    f64 = cf.dtype == np.float64
    if any(a is not None for a in (coarse_slip, fine_slip, coarse_q, fine_q)):
        assert f64 and ff.dtype == np.float64
        ctmp, ftmp = np.empty_like(cf), np.empty_like(ff)
        taus = np.ascontiguousarray(taus, dtype=np.float64)
        uins = np.ascontiguousarray(uins, dtype=np.float64)
        m4 = [np.zeros(n, dtype=np.float64) for _ in range(4)]
        sl = [None if a is None else [np.ascontiguousarray(x, np.uint16) for x in a]
              for a in (coarse_slip, fine_slip)]
        ptrs = []
        for a in sl:
            ptrs += [None] * 3 if a is None else [_p(x, _u16p) for x in a]
        keep = []
        for qa in (coarse_q, fine_q):
            if qa is None:
                ptrs += [None, None]
            else:
                qf = np.ascontiguousarray(qa[0], np.uint8)
                fm = np.ascontiguousarray(qa[1], np.uint32)
                keep += [qf, fm]
                ptrs += [_p(qf, _u8p), _p(fm, _u32p)]
        lib().orc_run_refined_slip_f64(_p(cf, _dp), _p(ctmp, _dp), _p(np.ascontiguousarray(cflags, np.uint8), _u8p),
                                       nx, ny, nz, _p(ff, _dp), _p(ftmp, _dp),
                                       _p(np.ascontiguousarray(fflags, np.uint8), _u8p), *box, m, len(taus),
                                       _p(taus, _dp), _p(uins, _dp), lam, float(smag_prefactor(cs)),
                                       *[_p(a, _dp) for a in m4], *ptrs)
        return tuple(m4)
    dt, pt = (np.float64, _dp) if f64 else (np.float32, _fp)
    assert ff.dtype == cf.dtype
    ctmp, ftmp = np.empty_like(cf), np.empty_like(ff)
    taus = np.ascontiguousarray(taus, dtype=dt)
    uins = np.ascontiguousarray(uins, dtype=dt)
    m4 = [np.zeros(n, dtype=dt) for _ in range(4)]
    fn = lib().orc_run_refined_f64 if f64 else lib().orc_run_refined_f32
    fn(_p(cf, pt), _p(ctmp, pt), _p(np.ascontiguousarray(cflags, np.uint8), _u8p), nx, ny, nz,
       _p(ff, pt), _p(ftmp, pt), _p(np.ascontiguousarray(fflags, np.uint8), _u8p), *box, m,
       len(taus), _p(taus, pt), _p(uins, pt), lam, float(smag_prefactor(cs)),
       *[_p(a, pt) for a in m4])
    return tuple(m4)


def run_refined3(cf, ff, gf, cflags, fflags, gflags, cdims, box, m, gbox, m2, taus, uins,
                 lam=3.0 / 16.0, cs=0.1733):
    """Three-level refined steps (coarse, fine, nested finer patch; FP64),
    lbm_solver.cpp:1144-1290. cf / ff / gf are advanced in place."""
    nx, ny, nz = cdims
    n = nx * ny * nz
    for a in (cf, ff, gf):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    taus = np.ascontiguousarray(taus, dtype=np.float64)
    uins = np.ascontiguousarray(uins, dtype=np.float64)
    m4 = [np.zeros(n) for _ in range(4)]
    u8 = lambda a: _p(np.ascontiguousarray(a, np.uint8), _u8p)
    lib().orc_run_refined3_f64(_p(cf, _dp), _p(np.empty_like(cf), _dp), u8(cflags), nx, ny, nz,
                               _p(ff, _dp), _p(np.empty_like(ff), _dp), u8(fflags), *box, m,
                               _p(gf, _dp), _p(np.empty_like(gf), _dp), u8(gflags), *gbox, m2,
                               len(taus), _p(taus, _dp), _p(uins, _dp), lam, float(smag_prefactor(cs)),
                               *[_p(a, _dp) for a in m4])
    return tuple(m4)


def f16c_encode(x):
    """The device's FP16C (s1e4m11) code of an FP32 deviation."""
    return int(lib().orc_f16c_encode(float(x)))


def f16c_decode(b):
    return float(lib().orc_f16c_decode(int(b)))


# ---- 16-bit store rounding (round to nearest even / stochastic) -------------------
# Restated from the formats' definitions and the stochastic-rounding rule (a
# value rounds up with probability equal to the fraction the kept bits drop):
# the random word's top bits are added below the kept significand before it
# is cut. Pinned by tests/test_oracle_pins.py (round to nearest equals the C
# oracle's FP16C encoder, numpy's float16 conversion and the BF16 bit rule;
# over every value of the random bits the mean decoded value equals the
# input exactly) and compared bit for bit with the device codec
# (fc_probe_store_rounding) in tests/test_gpu_parity.py.
F16C_LIM = np.float32(np.ldexp(np.float32(1) + np.float32(2047 / 2048), -112))  # code 0x7FFF


def store_codes(storage, v, r, stochastic):
    """16-bit codes of FP32 values v (the scaled deviations a step stores)
    with 32-bit random words r: storage 'f16c', 'f16' or 'bf16'."""
    v = np.ascontiguousarray(v, dtype=np.float32)
    r = np.ascontiguousarray(r, dtype=np.uint32)
    if storage == "f16c":
        t = np.fmin(np.fmax(v * np.float32(2.0 ** -112), -F16C_LIM), F16C_LIM)
        b = t.view(np.uint32)
        add = (r >> np.uint32(20)) if stochastic else (np.uint32(0x7FF) + ((b >> np.uint32(12)) & np.uint32(1)))
        u = b + add
        return (((u >> np.uint32(12)) & np.uint32(0x7FFF)) | ((u >> np.uint32(16)) & np.uint32(0x8000))).astype(
            np.uint16)
    b = v.view(np.uint32)
    finite = (b & np.uint32(0x7F800000)) != np.uint32(0x7F800000)
    if storage == "f16":
        if not stochastic:
            return v.astype(np.float16).view(np.uint16)
        t = ((b + (r >> np.uint32(19))) & np.uint32(0xFFFFE000)).view(np.float32)
        return np.where(finite, t, v).astype(np.float16).view(np.uint16)
    if storage == "bf16":
        rn = ((b + np.uint32(0x7FFF) + ((b >> np.uint32(16)) & np.uint32(1))) >> np.uint32(16)).astype(np.uint16)
        if not stochastic:
            return np.where(finite, rn, (b >> np.uint32(16)).astype(np.uint16))
        return np.where(finite, ((b + (r >> np.uint32(16))) >> np.uint32(16)).astype(np.uint16),
                        (b >> np.uint32(16)).astype(np.uint16))
    raise ValueError(storage)


def decode_codes(storage, c):
    """FP64 values of 16-bit codes (the codec's scaled deviations)."""
    c = np.asarray(c, dtype=np.uint16)
    if storage == "f16c":
        sx = c.view(np.int16).astype(np.int32).view(np.uint32)
        u = (sx << np.uint32(12)) & np.uint32(0x87FFF000)
        return u.view(np.float32).astype(np.float64) * 2.0 ** 112
    if storage == "f16":
        return c.view(np.float16).astype(np.float64)
    if storage == "bf16":
        return (c.astype(np.uint32) << np.uint32(16)).view(np.float32).astype(np.float64)
    raise ValueError(storage)


def wall_update(f, dims, wl, utau, nu, cutoff, ema=0.25):
    """One wall-model update (lbm_wallmodel.cu:76-196) on canonical FP32 f.
    wl: dict of SoA arrays (cell, sample, nx, ny, nz, ys, mask); utau updated
    in place. Returns (uwx, uwy, uwz half bits, stats[4])."""
    nx, ny, nz = dims
    n = nx * ny * nz
    uw = [np.zeros(n, np.uint16) for _ in range(3)]
    st = np.zeros(4, np.float32)
    ll = C.POINTER(C.c_longlong)
    f32 = np.ascontiguousarray(f, np.float32)
    lib().orc_wall_update_f32(_p(f32, _fp), nx, ny, nz, len(wl["cell"]),
                              wl["cell"].ctypes.data_as(ll), wl["sample"].ctypes.data_as(ll),
                              _p(wl["nx"], _fp), _p(wl["ny"], _fp), _p(wl["nz"], _fp),
                              _p(wl["ys"], _fp), _p(wl["mask"], _u32p), _p(utau, _fp), nu, cutoff,
                              ema, *[_p(a, _u16p) for a in uw], _p(st, _fp))
    return uw[0], uw[1], uw[2], st


# ---- tracer particles -----------------------------------------------------------
def particles_seed(count, dims, seed):
    pos = np.zeros((count, 4), np.float32)
    lib().orc_particles_seed(_p(pos, _fp), count, *dims, seed)
    return pos


def particles_advect(pos, u, v, w, flags, dims, dt, seed, age_rate=0.002, inlet_fraction=0.3):
    """RK2 advection in place (pos: (n,4) float32)."""
    fl = None if flags is None else np.ascontiguousarray(flags, np.uint8)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
    lib().orc_particles_advect(_p(pos, _fp), len(pos), *[_p(a, _fp) for a in arrs],
                               _p(fl, _u8p), *dims, dt, age_rate, inlet_fraction, seed)


# ---- warm restart (src/sim/snapshot.cpp, lbm_core.cu:162-198, :644-717) ------
# Restated in numpy / pure Python; the layout and the hash chain follow the
# reference's writer and SnapshotKey::hash field by field.
FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3
FCS_MAGIC = 0x31534346
FCS_FORMAT = 1
SOLVER_VERSION = 1
SOLID_FLAG = 1
FLUID_FLAG = 0


def fnv1a64(data: bytes, h: int = FNV_OFFSET) -> int:
    """FNV-1a 64 (snapshot.cpp:33-45)."""
    for b in data:
        h ^= b
        h = (h * FNV_PRIME) & 0xFFFFFFFFFFFFFFFF
    return h


def snapshot_key_hash(geometry_id: str, aoa_bucket: int, u_lat: float, dims, version=SOLVER_VERSION) -> int:
    """SnapshotKey::hash (snapshot.cpp:89-99): id bytes, then the native
    little-endian bytes of aoaBucket, u_lat, nx, ny, nz, solverVersion."""
    import struct
    h = fnv1a64(geometry_id.encode())
    for fmt, v in (("<i", aoa_bucket), ("<f", u_lat), ("<i", dims[0]), ("<i", dims[1]), ("<i", dims[2]),
                   ("<i", version)):
        h = fnv1a64(struct.pack(fmt, v), h)
    return h


def snapshot_key_string(geometry_id: str, aoa_bucket: int, u_lat: float, dims, version=SOLVER_VERSION) -> str:
    """SnapshotKey::toString (snapshot.cpp:101-112)."""
    u32 = float(np.float32(u_lat))
    s = "%s_a%.1f_u%.4f_%dx%dx%d_v%d" % (geometry_id, float(np.float32(aoa_bucket) * np.float32(0.5)), u32,
                                         dims[0], dims[1], dims[2], version)
    return s.replace(":", "-").replace("\\", "-").replace("/", "-")


# FCS1 scaling block, in file order (snapshot.cpp:316-329).
FCS_SCALING = [("chord_m", "<f"), ("airspeed_ms", "<f"), ("nu_m2s", "<f"), ("chord_cells", "<i"),
               ("u_lat", "<f"), ("dx", "<f"), ("dt", "<f"), ("nu_lat_target", "<f"), ("nu_lat", "<f"),
               ("tau", "<f"), ("tau_clamped", "<B"), ("re_target", "<f"), ("re_effective", "<f")]


def fcs1_write(path, geometry_id, aoa_bucket, u_lat, dims, flags_hash, scaling: dict, rho, u, v, w,
               version=SOLVER_VERSION, magic=FCS_MAGIC, fmt=FCS_FORMAT, compression=0, ncells=None):
    """CompactSnapshot::saveToFile (snapshot.cpp:287-347); the keyword
    overrides write deliberately malformed files for the loader's rejections."""
    import struct
    gid = geometry_id.encode()
    out = bytearray()
    out += struct.pack("<III", magic, fmt, len(gid)) + gid
    out += struct.pack("<ifiiii", aoa_bucket, u_lat, dims[0], dims[1], dims[2], version)
    out += struct.pack("<Q", flags_hash)
    for name, f in FCS_SCALING:
        out += struct.pack(f, scaling[name])
    n = len(rho) if ncells is None else ncells
    out += struct.pack("<IQ", compression, n)
    for a in (rho, u, v, w):
        out += np.ascontiguousarray(a, dtype="<f4").tobytes()
    with open(path, "wb") as fh:
        fh.write(bytes(out))


def fcs1_read(path):
    """Parse an FCS1 file into a dict (layout of snapshot.cpp:287-347)."""
    import struct
    b = open(path, "rb").read()
    o = 0

    def take(fmt):
        nonlocal o
        v = struct.unpack_from(fmt, b, o)
        o += struct.calcsize(fmt)
        return v if len(v) > 1 else v[0]

    d = {}
    d["magic"], d["format"], idlen = take("<III")
    d["geometry_id"] = b[o:o + idlen].decode()
    o += idlen
    d["aoa_bucket"], d["u_lat"], nx, ny, nz, d["solver_version"] = take("<ifiiii")
    d["dims"] = (nx, ny, nz)
    d["flags_hash"] = take("<Q")
    d["scaling"] = {name: take(f) for name, f in FCS_SCALING}
    d["compression"], n = take("<IQ")
    d["ncells"] = n
    for name in ("rho", "u", "v", "w"):
        d[name] = np.frombuffer(b, dtype="<f4", count=n, offset=o).copy()
        o += 4 * n
    d["trailing_bytes"] = len(b) - o
    return d


def equilibrium_np(rho, ux, uy, uz):
    """Second-order D3Q19 equilibrium, (Q, n) FP64 (lbm_core.cu:65-73)."""
    rho, ux, uy, uz = (np.asarray(a, np.float64) for a in (rho, ux, uy, uz))
    usq = 1.5 * (ux * ux + uy * uy + uz * uz)
    cu = 3.0 * (CX[:, None] * ux + CY[:, None] * uy + CZ[:, None] * uz)
    return W[:, None] * rho * (1.0 + cu + 0.5 * cu * cu - usq)


def init_from_macro(flags, rho, u, v, w):
    """Compact-snapshot restore (initFromMacroKernel, lbm_core.cu:162-198):
    canonical f (Q, n) = feq(rho, u), solid cells at rest."""
    fl = np.asarray(flags).reshape(-1)
    solid = fl == SOLID_FLAG
    r = np.where(solid, 1.0, np.asarray(rho, np.float64))
    a, b, c = (np.where(solid, 0.0, np.asarray(x, np.float64)) for x in (u, v, w))
    return equilibrium_np(r, a, b, c)


def apply_flag_edits(f, dims, old_flags, new_flags):
    """Flag-edit fixup on canonical f (Q, n), returned as a new array
    (applyEditedFlags + applyFlagEditsKernel, lbm_solver.cpp:1086-1126,
    lbm_core.cu:644-717): solid-state flips only; newly solid -> f = 0;
    newly fluid -> feq at the mean rho and mean u = j / rho over the 18
    neighbours (x / y not wrapped, z wrapped) that are Fluid under the new
    flags and not edited; rest state when none."""
    nx, ny, nz = dims
    f = np.asarray(f, np.float64).reshape(Q, -1)
    old = np.asarray(old_flags).reshape(-1)
    new = np.asarray(new_flags).reshape(-1)
    was, now = old == SOLID_FLAG, new == SOLID_FLAG
    edit = np.zeros(old.size, np.uint8)
    edit[~was & now] = 1
    edit[was & ~now] = 2
    out = f.copy()
    for c in np.nonzero(edit)[0]:
        if edit[c] == 1:
            out[:, c] = 0.0
            continue
        z, rem = divmod(int(c), nx * ny)
        y, x = divmod(rem, nx)
        acc = np.zeros(4)
        found = 0
        for q in range(1, Q):
            xn, yn, zn = x + CX[q], y + CY[q], (z + CZ[q]) % nz
            if xn < 0 or xn >= nx or yn < 0 or yn >= ny:
                continue
            n = xn + nx * (yn + ny * zn)
            if new[n] != FLUID_FLAG or edit[n] != 0:
                continue
            fn = f[:, n]
            r = fn.sum()
            acc += (r, (CX * fn).sum() / r, (CY * fn).sum() / r, (CZ * fn).sum() / r)
            found += 1
        rho, ux, uy, uz = acc / found if found else (1.0, 0.0, 0.0, 0.0)
        out[:, c] = equilibrium_np([rho], [ux], [uy], [uz])[:, 0]
    return out


# ---- boundary-layer readouts and the VG audit (lbm_solver.cpp:700-748,
# :1511-1655; vg_audit.cpp:12-187), restated with numpy on host arrays ---------
def surface_cache(flags, dims):
    """Top solid row per column of the mid-span plane and the solid x extent
    (rebuildSurfaceCache, lbm_solver.cpp:700-728)."""
    nx, ny, nz = dims
    plane = np.asarray(flags).reshape(nz, ny, nx)[nz // 2] == SOLID_FLAG
    top = np.full(nx, -1, np.int64)
    for x in range(nx):
        ys = np.nonzero(plane[:, x])[0]
        if ys.size:
            top[x] = ys.max()
    cols = np.nonzero(top >= 0)[0]
    if cols.size == 0:
        return top, nx, -1, False
    xle, xte = int(cols.min()), int(cols.max())
    return top, xle, xte, xte > xle


def speed32(a, b):
    """FP32 |(a, b)| from the exact FP64 sum of squares (glibc hypotf)."""
    a = np.asarray(a, np.float32).astype(np.float64)
    b = np.asarray(b, np.float32).astype(np.float64)
    return np.sqrt(a * a + b * b).astype(np.float32)


def _f32(x):
    return np.float32(x)


def suction_delta99(u, v, flags, dims, stations, chord_cells, u_lat):
    """extractSuctionDelta99 (lbm_solver.cpp:1511-1584) on macro fields
    (cells each); FP32 arithmetic as the reference's host code."""
    nx, ny, nz = dims
    top, xle, xte, ok = surface_cache(flags, dims)
    out = [{"x_c": float(_f32(s)), "delta99_c": 0.0, "ue_edge": 0.0, "valid": False} for s in stations]
    if not ok:
        return out
    mu = np.asarray(u, np.float32).reshape(nz, ny, nx)[nz // 2]
    mv = np.asarray(v, np.float32).reshape(nz, ny, nx)[nz // 2]
    span = _f32(xte - xle)
    ul = _f32(u_lat)
    for o in out:
        pos = _f32(xle) + _f32(o["x_c"]) * span
        ix = int(min(max(int(np.floor(abs(float(pos)) + 0.5)) * (1 if pos >= 0 else -1), 0), nx - 1))
        ys = int(top[ix])
        if ys < 0 or ys >= ny - 3:
            continue
        ytop = min(ny - 2, ys + chord_cells // 2)
        sp = speed32(mu[ys + 1:ytop + 1, ix], mv[ys + 1:ytop + 1, ix])
        ue = _f32(sp.max()) if sp.size else _f32(0)
        o["ue_edge"] = float(ue)
        if ue < _f32(0.05) * ul:
            continue
        hit = np.nonzero(sp >= _f32(0.99) * ue)[0]
        if hit.size == 0:
            continue
        d99 = _f32(hit[0] + 1) - _f32(0.5)
        cos = _f32(1)
        if ix >= 2 and ix + 2 < nx and top[ix - 2] >= 0 and top[ix + 2] >= 0:
            slope = _f32(top[ix + 2] - top[ix - 2]) / _f32(4)
            cos = _f32(1) / np.sqrt(_f32(1) + slope * slope, dtype=np.float32)
        o["delta99_c"] = float(d99 * cos / _f32(chord_cells))
        o["valid"] = bool(mu[ys + 1, ix] > _f32(-0.02) * ul)
    return out


def separation_onset_xc(u, flags, dims, u_lat):
    """separationOnsetXc (lbm_solver.cpp:1587-1617), sequential scan."""
    nx, ny, nz = dims
    top, xle, xte, ok = surface_cache(flags, dims)
    if not ok:
        return -1.0
    mu = np.asarray(u, np.float32).reshape(nz, ny, nx)[nz // 2]
    span = _f32(xte - xle)
    urev = _f32(-0.02) * _f32(u_lat)
    run = 0
    start = xle + max(2, int(_f32(0.03) * span))
    ix = start
    while ix <= xte and ix < nx:
        ys = int(top[ix])
        if ys < 0 or ys + 2 >= ny:
            run = 0
        elif mu[ys + 2, ix] < urev:
            run += 1
            if run >= 3:
                return float((_f32(ix - run + 1) - _f32(xle)) / span)
        else:
            run = 0
        ix += 1
    return -1.0


def wendt_circulation(beta_rad, ue, h, c, d99):
    """Wendt correlation in lattice units (vg_audit.cpp:12-25), FP32."""
    if ue <= 0 or h <= 0 or c <= 0 or d99 <= 0:
        return 0.0
    f = np.float32
    ar = f(8) * f(h) / (f(3.14159265) * f(c))
    return float(f(1.61) * abs(f(beta_rad)) * f(ue) * f(c) * np.tanh(f(1.41) * f(h) / f(d99), dtype=f)
                 / (f(1) + f(0.48) / ar))


def band_circulation(v, w, ny, nz, y0, y1, z0, z1, f64=False):
    """integrateBandCirculation (vg_audit.cpp:62-101) on a plane indexed
    y + ny*z: omega_x in FP32; sums in FP32 in the reference's z-then-y order,
    or in FP64."""
    v = np.asarray(v, np.float32).reshape(nz, ny)
    w = np.asarray(w, np.float32).reshape(nz, ny)
    out = {"zone": [0.0, 0.0], "ambient": [0.0, 0.0], "zone_cols": 0, "ambient_cols": 0}
    if nz < 3:
        return out
    y0, y1 = max(y0, 0), min(y1, ny - 1)
    z0, z1 = max(z0, 0), min(z1, nz - 1)
    if y1 <= y0 or z1 < z0:
        return out
    acc = np.float64 if f64 else np.float32
    sums = {"zone": [acc(0), acc(0)], "ambient": [acc(0), acc(0)]}
    h = np.float32(0.5)
    for z in range(nz):
        key = "zone" if z0 <= z <= z1 else "ambient"
        out[key + "_cols"] += 1
        zm, zp = (z - 1) % nz, (z + 1) % nz
        for y in range(y0, y1 + 1):
            if y == y0:
                dwdy = w[z, y + 1] - w[z, y]
            elif y == y1:
                dwdy = w[z, y] - w[z, y - 1]
            else:
                dwdy = h * (w[z, y + 1] - w[z, y - 1])
            om = dwdy - h * (v[zp, y] - v[zm, y])
            sums[key][0 if om > 0 else 1] += acc(om)
    out["zone"] = [float(x) for x in sums["zone"]]
    out["ambient"] = [float(x) for x in sums["ambient"]]
    return out


def audit_vg(v, w, flags, dims, vg, chord_cells, ue, d99):
    """auditVGVortexStrength (vg_audit.cpp:103-187) on host fields (cells
    each), band sums in FP64; vg: dict with type, x_c, height_c, length_h,
    beta_deg, pitch_c, gap_h, count."""
    f = np.float32
    out = {"valid": False, "gamma_measured": 0.0, "gamma_predicted": 0.0, "ratio": 0.0, "plane_xc": 0.0}
    if chord_cells <= 0 or ue <= 0 or d99 <= 0 or vg["count"] <= 0:
        return out
    h = f(vg["height_c"]) * f(chord_cells)
    if h <= 0:
        return out
    beta = f(vg["beta_deg"]) * f(3.14159265) / f(180)
    gp = wendt_circulation(beta, ue, h, f(vg["length_h"]) * h, d99)
    out["gamma_predicted"] = gp
    if gp <= 0:
        return out
    pxc = float(min(max(f(vg["x_c"]) + f(vg["height_c"]) * (f(vg["length_h"]) + f(3)), f(0.02)), f(0.97)))
    out["plane_xc"] = pxc
    nx, ny, nz = dims
    top, xle, xte, ok = surface_cache(flags, dims)
    if not ok:
        return out
    pos = f(xle) + f(pxc) * f(xte - xle)
    ix = int(min(max(int(np.floor(abs(float(pos)) + 0.5)) * (1 if pos >= 0 else -1), 0), nx - 1))
    ys = int(top[ix])
    if ys < 0:
        return out
    band = max(2, int(np.ceil(f(4) * h)))
    pitch = f(vg["pitch_c"]) * f(chord_cells)
    half = f(0.5) * (f(vg["count"] - 1) * pitch + f(vg["gap_h"]) * h) + f(3) * h
    mid = nz // 2
    z0, z1 = mid - int(np.ceil(half)), mid + int(np.ceil(half))
    pv = np.asarray(v, np.float32).reshape(nz, ny, nx)[:, :, ix].reshape(-1)
    pw = np.asarray(w, np.float32).reshape(nz, ny, nx)[:, :, ix].reshape(-1)
    b = band_circulation(pv, pw, ny, nz, ys + 1, min(ny - 2, ys + band), z0, z1, f64=True)
    if b["zone_cols"] <= 0:
        return out
    fp = fn = 0.0
    if b["ambient_cols"] >= 4:
        sc = b["zone_cols"] / b["ambient_cols"]
        fp, fn = b["ambient"][0] * sc, -b["ambient"][1] * sc
    pn, nn = max(0.0, b["zone"][0] - fp), max(0.0, -b["zone"][1] - fn)
    n = vg["count"]
    g = (pn + nn) / (2 * n) if vg["type"] in (1, 3) else max(pn, nn) / n
    out.update(valid=True, gamma_measured=g, ratio=g / gp)
    return out


def upsample_macro(fields, sdims, ddims):
    """Trilinear upsample of macro fields (upsampleMacroKernel,
    lbm_refine.cu:437-501) in FP32, the expression order as written; fields
    = (rho, u, v, w), cells of sdims each. Returns the four fields on ddims."""
    snx, sny, snz = sdims
    dnx, dny, dnz = ddims
    f = np.float32
    z, y, x = np.meshgrid(np.arange(dnz), np.arange(dny), np.arange(dnx), indexing="ij")
    x, y, z = (a.reshape(-1) for a in (x, y, z))

    def centre(i, s, d):
        return (i.astype(f) + f(0.5)) * f(s) / f(d) - f(0.5)

    gx, gy, gz = centre(x, snx, dnx), centre(y, sny, dny), centre(z, snz, dnz)
    x0 = np.floor(gx).astype(np.int64)
    y0 = np.floor(gy).astype(np.int64)
    zf = np.floor(gz).astype(np.int64)
    tx = (gx - x0.astype(f)).astype(f)
    ty = (gy - y0.astype(f)).astype(f)
    tz = (gz - zf.astype(f)).astype(f)
    lo = x0 < 0
    x0[lo], tx[lo] = 0, 0
    lo = y0 < 0
    y0[lo], ty[lo] = 0, 0
    x1, y1 = x0 + 1, y0 + 1
    hi = x1 > snx - 1
    x1[hi], tx[hi] = snx - 1, 0
    hi = y1 > sny - 1
    y1[hi], ty[hi] = sny - 1, 0
    z0 = ((zf % snz) + snz) % snz
    z1 = (z0 + 1) % snz
    ux, uy, uz = f(1) - tx, f(1) - ty, f(1) - tz
    wts = [ux * uy * uz, tx * uy * uz, ux * ty * uz, tx * ty * uz, ux * uy * tz, tx * uy * tz, ux * ty * tz,
           tx * ty * tz]
    idx = [(x0, y0, z0), (x1, y0, z0), (x0, y1, z0), (x1, y1, z0), (x0, y0, z1), (x1, y0, z1), (x0, y1, z1),
           (x1, y1, z1)]
    out = []
    for a in fields:
        a = np.asarray(a, f).reshape(-1)
        acc = None
        for wk, (xi, yi, zi) in zip(wts, idx):
            term = (wk * a[xi + snx * (yi + sny * zi)]).astype(f)
            acc = term if acc is None else (acc + term).astype(f)
        out.append(acc)
    return out
