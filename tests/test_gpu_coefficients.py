"""Time-averaged lift and drag coefficients of every storage format against
the FP64 run of the same case: config 2 geometry (NACA 2412, chord 128,
384 x 128 x 64, AoA 8, Cs 0.1, u_lat 0.08), 20k steps, one momentum-exchange
sample every 100 steps, averaged over the second half (reference config 2
protocol).

The relaxation time is 0.52 (chord Reynolds number 1.5e3) instead of the
physical 2.4e6 (tau 0.5005): at tau 0.5005 the reference step itself (the
CPU oracle, which restates it) reaches the 0.30 speed cap within 600 steps on
this grid and on the 128 x 48 x 16 tunnel, so no time average exists there.
FP32 must agree within 0.5 % and FP16C (the 16-bit s1e4m11 storage) within
1 %, both for Cl and for Cd (relative to each coefficient). IEEE FP16 and
BF16 storage are measured and printed (their errors are bounded loosely; the
11-bit mantissa of FP16C is what brings 16-bit storage inside 1 %).
"""
import sys

import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd import cases

pytestmark = pytest.mark.gpu

STEPS, EVERY, AVG_FROM = 20000, 100, 10000


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def averaged_coefficients(case, storage):
    s = cases.make_solver(case, storage)
    span = case.dims[2]
    cl, cd = [], []
    for i in range(STEPS // EVERY):
        s.step_n(EVERY)
        if (i + 1) * EVERY > AVG_FROM:
            F = s.force_now()
            cd.append(fc.force_to_coefficient(case.scaling, float(F[0]), span))
            cl.append(fc.force_to_coefficient(case.scaling, float(F[1]), span))
    assert not s.nan_detected()
    s.close()
    return float(np.mean(cl)), float(np.mean(cd))


def test_time_averaged_coefficients_vs_fp64(gpu):
    case = cases.naca_tunnel(384, 128, 64, 128, aoa=8.0, u_lat=0.08, cs=0.1, tau=0.52)
    say(f"case: config2 {case.dims} {STEPS} steps, averaging from step {AVG_FROM}")
    res = {}
    for st in ("f64", "f32", "f16c", "f16", "bf16"):
        res[st] = averaged_coefficients(case, st)
        say(f"{st}: <Cl> {res[st][0]:.5f}  <Cd> {res[st][1]:.5f}")
    cl64, cd64 = res["f64"]
    errs = {}
    for st in ("f32", "f16c", "f16", "bf16"):
        el = abs(res[st][0] - cl64) / abs(cl64)
        ed = abs(res[st][1] - cd64) / abs(cd64)
        errs[st] = (el, ed)
        say(f"{st} vs f64: Cl err {100 * el:.3f} %, Cd err {100 * ed:.3f} %")
    assert max(errs["f32"]) < 5e-3
    assert max(errs["f16c"]) < 1e-2
    assert max(errs["f16"]) < 0.1
    assert max(errs["bf16"]) < 1.0
