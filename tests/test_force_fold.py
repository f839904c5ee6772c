"""The force fold (EMA + trailing-average statistics) against closed forms,
through the C ABI on caller-supplied samples (host code: no GPU). The solver
runs the same fold on its own momentum-exchange samples.

Reference: EMA fold lbm_solver.cpp:633-654 (alpha = 1 - exp(-dSteps /
windowSteps), first sample seeds), readout lbm_solver.cpp:1414-1509 (gate,
samples before one flow-through dropped, mean / min / max / median),
256-sample ring lbm_solver.cpp:120, coefficient C = F / (1/2 u^2 Nc span)
units.h:168-173."""
import math

import numpy as np

import foil_cfd_amd as fc
from foil_cfd_amd.solver import ForceFold

SCALING = fc.compute_scaling(1.2, 30.0, 1.48e-5, 256, 0.08)
SPAN = 240
QREF = 0.5 * 0.08 * 0.08 * 256 * SPAN  # force of a unit coefficient


def coef(force):
    return force / QREF


def test_ema_weight_is_one_minus_exp_of_the_elapsed_window_fraction():
    fold = ForceFold()
    w = 6400.0
    for d in (1.0, 100.0, 640.0, 6400.0, 20000.0):
        fold.clear()
        fold.add(0, 0.0, w, 0.0, 0.0)           # seeds the EMA at 0
        fold.add(int(d), d, w, 0.0, QREF)       # a unit step in Cl
        cl = fold.readout(SCALING, SPAN, 10**9, False).cl
        assert abs(cl - (1 - math.exp(-d / w))) < 2e-6 * max(1.0, cl), (d, cl)
    # equally spaced samples of a unit step: 1 - (1 - alpha)^k = 1 - exp(-k d / W)
    fold.clear()
    fold.add(0, 0.0, w, 0.0, 0.0)
    for k in range(1, 65):
        fold.add(100 * k, 100.0, w, -QREF, 2 * QREF)
    r = fold.readout(SCALING, SPAN, 10**9, False)
    want = 1 - math.exp(-6400.0 / w)
    assert abs(r.cl - 2 * want) < 1e-5 and abs(r.cd + want) < 1e-5
    assert abs(r.lift_to_drag - r.cl / r.cd) < 1e-6
    fold.close()


def test_first_sample_seeds_and_window_has_a_floor_of_one_step():
    fold = ForceFold()
    fold.add(500, 500.0, 6400.0, 0.3 * QREF, 0.7 * QREF)  # seeds, whatever the weight
    r = fold.readout(SCALING, SPAN, 0, False)
    assert abs(r.cl - 0.7) < 1e-6 and abs(r.cd - 0.3) < 1e-6
    fold.clear()
    fold.add(0, 0.0, 0.0, 0.0, 0.0)
    fold.add(1, 1.0, 0.0, 0.0, QREF)  # window 0 -> 1 step: alpha = 1 - 1/e
    assert abs(fold.readout(SCALING, SPAN, 0, False).cl - (1 - math.exp(-1))) < 1e-6
    fold.close()


def test_gate_and_startup_transient():
    fold = ForceFold()
    one_ft = 6400
    # start-up transient (before one flow-through): wild values that must not count
    for k in range(1, 64):
        fold.add(100 * k, 100.0, 6400.0, 50 * QREF, -80 * QREF)
    vals = [0.8, 1.2, 1.0, 0.9, 1.1]
    for i, v in enumerate(vals):
        fold.add(one_ft + 100 * i, 100.0, 6400.0, 0.1 * v * QREF, v * QREF)
    closed = fold.readout(SCALING, SPAN, one_ft, False)
    assert not closed.valid and closed.cl_avg == 0.0 and closed.cl_median == 0.0
    r = fold.readout(SCALING, SPAN, one_ft, True)
    assert r.valid
    assert abs(r.cl_avg - np.mean(vals)) < 1e-6
    assert abs(r.cd_avg - 0.1 * np.mean(vals)) < 1e-7
    assert abs(r.cl_min - 0.8) < 1e-6 and abs(r.cl_max - 1.2) < 1e-6
    assert abs(r.cl_median - 1.0) < 1e-7                 # odd count: the middle sample
    assert abs(r.ld_avg - 10.0) < 1e-5 and abs(r.ld_median - 10.0) < 1e-5
    fold.add(one_ft + 500, 100.0, 6400.0, 0.1 * 1.3 * QREF, 1.3 * QREF)
    r = fold.readout(SCALING, SPAN, one_ft, True)
    assert abs(r.cl_median - 0.5 * (1.0 + 1.1)) < 1e-7  # even count: mean of the middle two
    fold.close()


def test_ring_keeps_the_last_256_samples():
    fold = ForceFold()
    n = 300
    for k in range(n):
        fold.add(10000 + k, 1.0, 6400.0, QREF, float(k) * QREF)
    r = fold.readout(SCALING, SPAN, 0, True)
    kept = np.arange(n - 256, n, dtype=np.float64)
    assert abs(r.cl_avg - kept.mean()) < 1e-4 * kept.mean()
    assert abs(r.cl_min - (n - 256)) < 1e-4 and abs(r.cl_max - (n - 1)) < 1e-4
    assert abs(r.cl_median - 0.5 * (kept[127] + kept[128])) < 1e-4
    fold.close()


def test_zero_drag_guard():
    fold = ForceFold()
    fold.add(7000, 0.0, 6400.0, 1e-9 * QREF, QREF)  # |Cd| < 1e-6
    r = fold.readout(SCALING, SPAN, 0, True)
    assert r.lift_to_drag == 0.0 and r.ld_avg == 0.0
    fold.close()
