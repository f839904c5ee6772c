"""Batched stepping of independent solvers (fc_group_*): a group of uneven
grids stepped with one launch per step (HIP graph replay) matches the same
solvers stepped one by one bit for bit, for every launch shape the
autotuner can pick; ineligible members are refused."""
import sys

import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd import cases, group

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_polar(storage, wall=False):
    """Three uneven NACA 2412 tunnels (chords 40 / 48 / 56) at -4, 6, 16 deg."""
    cs = group.polar_sweep(3, -4.0, 16.0, chords=[40, 48, 56], u_lat=0.06, cs=0.1)
    solvers = []
    for k, c in enumerate(cs):
        s = cases.make_solver(c, storage)
        if wall and k == 1:
            s.set_wall_model("newton")
        solvers.append(s)
    return cs, solvers


def state(s):
    s.synchronize()
    s.forces()  # folds the batch's force sample
    return s.download_f(f64=True), s.read_macro(), s.force_raw(), s.step_count()


def same(a, b, forces=True):
    """Bit-identical populations, macro fields and step count (and the last
    force sample when both runs sampled at the same batches)."""
    fa, ma, ra, na = a
    fb, mb, rb, nb = b
    nan = np.isnan(fa) & np.isnan(fb)
    ok = not np.count_nonzero((fa != fb) & ~nan)
    ok = ok and all(np.array_equal(x, y) for x, y in zip(ma, mb)) and na == nb
    if forces:
        ok = ok and np.array_equal(np.asarray(ra[0]), np.asarray(rb[0])) and ra[1] == rb[1]
    return ok


@pytest.mark.parametrize("storage", ["f16c", "f32"])
def test_group_matches_individual_steps(gpu, storage):
    say(f"case: group of 3 uneven tunnels vs individual stepping, {storage}, 37 + 300 steps")
    # member 1 carries the wall model and member 0 starts 3 steps ahead (odd
    # parity): both the uniform and the mixed-member kernels run
    _, ind = small_polar(storage, wall=True)
    ind[0].step_n(3)
    for s in ind:
        s.step_n(37)
        state(s)  # each batch's force sample is taken (and folded) in both runs
        s.step_n(300)
    want = [state(s) for s in ind]
    for s in ind:
        s.close()
    _, mem = small_polar(storage, wall=True)
    mem[0].step_n(3)
    g = group.SolverGroup(mem)
    g.step(37)
    for s in mem:
        state(s)
    g.step(300)  # crosses the 256-step table chunk
    got = [state(s) for s in mem]
    for k, (a, b) in enumerate(zip(got, want)):
        assert same(a, b), f"member {k} differs"
    info = g.info()
    assert info["members"] == 3 and info["cells"] == sum(s.ncells for s in mem)
    g.close()
    for s in mem:
        s.close()


def test_group_launch_shapes_identical(gpu):
    say("case: group launch shapes and graph on/off give identical states (f16c)")
    ref = None
    for wpb, slab, graph, budget in ((4, True, True, 0), (1, False, True, 1), (8, True, False, 0),
                                     (2, False, False, 6 << 20), (4, True, True, 12 << 20)):
        _, mem = small_polar("f16c")
        g = group.SolverGroup(mem)
        g.set_launch(wpb, slab, graph)
        g.set_cache_budget(budget)
        g.step(51)
        st = [state(s) for s in mem]
        if ref is None:
            ref = st
        else:
            assert all(same(a, b) for a, b in zip(st, ref)), (wpb, slab, graph, budget)
        say(f"  shape {wpb} {slab} {graph} budget {budget}: {g.info()['sub_batches']} sub-batches")
        g.close()
        for s in mem:
            s.close()


def test_group_autotune(gpu):
    say("case: group autotune over 5 budgets and 8 launch shapes, then parity with individual stepping")
    _, mem = small_polar("f16c")
    g = group.SolverGroup(mem)
    best = g.autotune(steps_per_trial=5)  # 13 trials x 2 batches x 5 steps
    say(f"  best shape: {best}")
    assert best["waves_per_block"] in (1, 2, 4, 8) and best["last_step_ms"] > 0 and best["glups"] > 0
    g.step(21)
    ms = g.last_step_ms()
    assert ms > 0
    got = [state(s) for s in mem]
    _, ind = small_polar("f16c")
    for s in ind:
        for _ in range(26):
            s.step_n(5)
        s.step_n(21)
    want = [state(s) for s in ind]
    assert all(same(a, b, forces=False) for a, b in zip(got, want))
    g.close()
    for s in mem + ind:
        s.close()


def test_group_refuses_ineligible_members(gpu):
    say("case: group refusals")
    c = cases.naca_tunnel(64, 32, 8, 16, aoa=4.0, cs=0.1)
    a = cases.make_solver(c, "f16c")
    b = cases.make_solver(c, "f32")
    with pytest.raises(RuntimeError, match="storage"):
        group.SolverGroup([a, b])
    p = cases.make_solver(cases.taylor_green(16), "f16c")
    with pytest.raises(RuntimeError, match="periodic"):
        group.SolverGroup([a, p])
    d = cases.make_solver(c, "f64")
    with pytest.raises(RuntimeError, match="16/32-bit"):
        group.SolverGroup([d])
    with pytest.raises(RuntimeError, match="twice"):
        group.SolverGroup([a, a])
    for s in (a, b, p, d):
        s.close()
