"""The allocation-free force entry point (fc_force_now_noalloc) against
fc_force_now, through the C ABI: two calls in a row on different solvers,
the second with a smaller force list, then a refined solver whose force runs
on the fine grid. Same kernels and summation order, so the results agree bit
for bit, and the scratch a call leaves behind does not leak into the next."""
import sys

import numpy as np
import pytest

from foil_cfd_amd import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_force_now_noalloc_matches_force_now(gpu):
    results = []
    for dims, chord in (((256, 96, 32), 64), ((128, 48, 16), 40)):
        say(f"case: no-alloc force, tunnel {dims} chord {chord}, f32, 60 steps")
        case = cases.naca_tunnel(*dims, chord, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56)
        s = cases.make_solver(case, "f32")
        s.step_n(60)
        s.synchronize()
        a = s.force_now_noalloc()
        b = s.force_now()
        say(f"  noalloc {a} force_now {b}")
        assert np.array_equal(a, b) and np.all(np.isfinite(a)) and abs(a[0]) > 0
        results.append(a)
        s.close()
    assert not np.array_equal(results[0], results[1])
    say("case: no-alloc force, refined tunnel 192x72x16 (forces on the fine grid), f32")
    case = cases.naca_tunnel(192, 72, 16, 48, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56)
    box, fine = cases.refinement_patch(case, aoa=8.0)
    s = cases.make_solver(case, "f32")
    s.init_refinement(box, 2, fine)
    assert s.refinement_info()["forces_from_fine"]
    s.step_n(20)
    s.synchronize()
    a, b = s.force_now_noalloc(), s.force_now()
    say(f"  noalloc {a} force_now {b}")
    assert np.array_equal(a, b)
    s.close()
