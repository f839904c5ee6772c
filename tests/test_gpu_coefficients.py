"""Time-averaged lift and drag coefficients of every storage format against
the FP64 run of the same case, one momentum-exchange sample every 100 steps
(the reference config 2 protocol).

The relaxation time is 0.52 instead of the physical 2.4e6-Reynolds value
(tau 0.5005): at tau 0.5005 the reference step itself (the CPU oracle, which
restates it) reaches the 0.30 speed cap within 600 steps on the config 2 grid
and on the 128 x 48 x 16 tunnel, so no time average exists there.

* 128 x 48 x 16 tunnel (chord 40, Re = 480, steady after ~2000 steps), 6000
  steps averaged over the second half: FP32 within 0.5 %, FP16C (the 16-bit
  s1e4m11 storage) within 1 %, both for Cl and Cd. IEEE FP16 (~4 % in Cl) and
  BF16 (~12 %) are measured and bounded loosely: their 10- and 7-bit
  mantissas stall the slow viscous relaxation of the steady state.
* Config 2 geometry (384 x 128 x 64, chord 128, Re = 1.5e3), 20k steps
  averaged over the second half: FP32 within 0.5 %; the 16-bit formats are
  printed with the sample spread of the averaged signal (this flow is not
  steady, so a 16-bit run is a different realisation of it).
"""
import sys

import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def averaged_coefficients(case, storage, steps, avg_from, every=100):
    s = cases.make_solver(case, storage)
    span = case.dims[2]
    cl, cd = [], []
    for i in range(steps // every):
        s.step_n(every)
        if (i + 1) * every > avg_from:
            F = s.force_now()
            cd.append(fc.force_to_coefficient(case.scaling, float(F[0]), span))
            cl.append(fc.force_to_coefficient(case.scaling, float(F[1]), span))
    assert not s.nan_detected()
    s.close()
    return np.asarray(cl), np.asarray(cd)


def compare(case, steps, avg_from, storages):
    res = {}
    for st in storages:
        cl, cd = averaged_coefficients(case, st, steps, avg_from)
        res[st] = (cl.mean(), cd.mean(), cl.std(), cd.std())
        say(f"{st}: <Cl> {res[st][0]:.5f} (sd {res[st][2]:.2e})  <Cd> {res[st][1]:.5f} "
            f"(sd {res[st][3]:.2e})")
    cl64, cd64 = res["f64"][:2]
    errs = {}
    for st in storages[1:]:
        el = abs(res[st][0] - cl64) / abs(cl64)
        ed = abs(res[st][1] - cd64) / abs(cd64)
        errs[st] = (el, ed)
        say(f"{st} vs f64: Cl err {100 * el:.3f} %, Cd err {100 * ed:.3f} %")
    return errs


def test_time_averaged_coefficients_steady_tunnel(gpu):
    case = cases.naca_tunnel(128, 48, 16, 40, aoa=8.0, u_lat=0.08, cs=0.1, tau=0.52)
    say(f"case: tunnel {case.dims} tau 0.52, 6000 steps, averaging from step 3000")
    errs = compare(case, 6000, 3000, ("f64", "f32", "f16c", "f16", "bf16"))
    assert max(errs["f32"]) < 5e-3
    assert max(errs["f16c"]) < 1e-2
    assert max(errs["f16"]) < 0.1
    assert max(errs["bf16"]) < 0.25


def test_time_averaged_coefficients_config2_geometry(gpu):
    case = cases.naca_tunnel(384, 128, 64, 128, aoa=8.0, u_lat=0.08, cs=0.1, tau=0.52)
    say(f"case: config2 geometry {case.dims} tau 0.52, 20000 steps, averaging from step 10000")
    errs = compare(case, 20000, 10000, ("f64", "f32", "f16c", "f16", "bf16"))
    assert max(errs["f32"]) < 5e-3
    assert max(errs["f16c"]) < 5e-2
