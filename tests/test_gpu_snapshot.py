"""Warm restart on the GPU (src/sim/snapshot.h/.cpp; applyEditedFlags and
restoreFromMacroscopic, lbm_solver.cpp:1086-1126 and :1773-1803; fixup kernel
lbm_core.cu:644-717):

* an exact (device-resident) snapshot restores the populations bit for bit:
  stepping on from a restore reproduces the uninterrupted run exactly, on the
  same solver and on a fresh one of the same grid, in either layout parity and
  with a fine patch;
* a compact snapshot restores f = feq(rho, u) of the captured macro fields,
  checked against the oracle's FP64 restatement, through the FCS1 file;
* the flag-edit fixup changes exactly the edited cells, as the oracle's
  restatement of the reference fixup does from the same pre-edit state.

Tolerances: the fixup and the equilibrium are computed in FP32 for the 16/32
bit formats (FP64 for FP64 storage) and stored in the run's format; the
bounds sit at a few storage ulps of |f - w| <= 0.05, far below the O(1e-3)
that a wrong neighbour, sign or layout slot produces."""
import sys
import time

import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd import cases
from foil_cfd_amd import snapshot as S
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
Q = fc.Q
TOL = {"f64": 1e-13, "f32": 2e-7, "f16c": 1e-6}
# FP16C keeps 12 significant bits of f - w: half an ulp is 2^-12 of |f - w|
# (the stored f = 0 of a newly solid rest direction, f - w = -1/3, rounds by
# up to 6e-5).
REL = {"f64": 0.0, "f32": 0.0, "f16c": 2.0 ** -12}


def excess(got, want, storage):
    """max of |got - want| beyond the storage bound (<= 0 passes)."""
    dev = np.abs(want - orc.W[:, None])
    return float((np.abs(got - want) - (TOL[storage] + REL[storage] * dev)).max())


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tunnel(nx=96, ny=48, nz=8, chord=32):
    return cases.naca_tunnel(nx, ny, nz, chord, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56, ramp=False)


def key_of(case):
    return S.SnapshotKey("naca:2412", S.SnapshotKey.bucket_for_aoa(8.0), case.scaling.u_lat, case.dims)


def state(s, fine=False):
    s.synchronize()
    out = [s.download_f(f64=True), *s.read_macro()]
    if fine:
        out.append(s.download_fine_f(f64=True))
    return out


def same(a, b):
    for x, y in zip(a, b):
        x, y = np.asarray(x), np.asarray(y)
        if x.shape != y.shape:
            return False
        nan = np.isnan(x) & np.isnan(y)
        if np.count_nonzero((x != y) & ~nan):
            return False
    return True


@pytest.mark.parametrize("storage,n0", [("f16c", 37), ("f32", 40), ("f64", 13)])
def test_exact_snapshot_restore_is_bit_exact(gpu, storage, n0):
    say(f"case: exact snapshot, tunnel 96x48x8 {storage}, capture after {n0} steps, 50 more")
    case = tunnel()
    key = key_of(case)
    a = cases.make_solver(case, storage)
    a.step_n(n0)
    snap = S.VramSnapshot()
    assert not snap.has_data()
    snap.capture(a, key)
    assert snap.has_data() and snap.key().equals(key)
    assert snap.flags_hash() == orc.fnv1a64(case.flags.tobytes())
    a.step_n(50)
    want = state(a)
    assert snap.restore(a)
    assert not a.forces().valid  # the force gate closes for a short re-settle
    a.step_n(50)
    assert same(state(a), want), "restore on the same solver"
    b = cases.make_solver(case, storage)
    b.step_n(3)
    assert snap.restore(b)
    b.step_n(50)
    assert same(state(b), want), "restore on a fresh solver"
    # refused: another grid, another storage format of the same width
    other = cases.make_solver(tunnel(nx=100), storage)
    assert not snap.restore(other)
    alt = {"f16c": "bf16", "f32": "f16c", "f64": "f32"}[storage]
    c = cases.make_solver(case, alt)
    assert not snap.restore(c)
    snap.release()
    assert not snap.has_data() and not snap.restore(a)
    for s in (a, b, other, c):
        s.close()
    snap.close()


def test_exact_snapshot_with_fine_patch(gpu):
    say("case: exact snapshot with a factor-2 fine patch, tunnel 160x64x8 f32, capture after 21 steps")
    case = cases.naca_tunnel(160, 64, 8, 32, aoa=8.0, cs=0.1, u_lat=0.05, tau=0.56, ramp=False)
    box, fine = cases.refinement_patch(case)
    key = key_of(case)
    a = cases.make_solver(case, "f32")
    a.init_refinement(box, 2, fine)
    a.step_n(21)
    snap = S.VramSnapshot()
    snap.capture(a, key)
    a.step_n(30)
    want = state(a, fine=True)
    assert snap.restore(a)
    a.step_n(30)
    assert same(state(a, fine=True), want)
    b = cases.make_solver(case, "f32")
    b.init_refinement(box, 2, fine)
    assert snap.restore(b)
    b.step_n(30)
    assert same(state(b, fine=True), want)
    plain = cases.make_solver(case, "f32")  # no patch: level shapes differ
    assert not snap.restore(plain)
    for s in (a, b, plain):
        s.close()
    snap.close()


@pytest.mark.parametrize("storage", ["f32", "f16c", "f64"])
def test_compact_snapshot_restore_matches_oracle(gpu, storage, tmp_path):
    say(f"case: compact snapshot through FCS1, tunnel 96x48x8 {storage}, 60 steps")
    case = tunnel()
    key = key_of(case)
    a = cases.make_solver(case, storage)
    a.step_n(60)
    a.synchronize()
    rho, u, v, w = a.read_macro()
    snap = S.CompactSnapshot()
    snap.capture(a, key)
    for x, y in zip(snap.fields(), (rho, u, v, w)):
        assert np.array_equal(x, y)
    assert snap.flags_hash() == orc.fnv1a64(case.flags.tobytes())
    assert snap.scaling().tau == case.scaling.tau
    p = tmp_path / key.file_name()
    snap.save_to_file(p)
    d = orc.fcs1_read(p)
    assert d["dims"] == case.dims and d["geometry_id"] == "naca:2412" and d["aoa_bucket"] == 16
    assert all(np.array_equal(d[n], x) for n, x in zip(("rho", "u", "v", "w"), (rho, u, v, w)))
    back = S.CompactSnapshot()
    back.load_from_file(p)
    b = cases.make_solver(case, storage)
    b.step_n(7)
    assert back.restore(b)
    got = b.download_f(f64=True).reshape(Q, -1)
    want = orc.init_from_macro(case.flags, rho, u, v, w)
    err = float(np.abs(got - want).max())
    say(f"  compact restore max |f - feq| vs oracle: {err:.2e}")
    assert excess(got, want, storage) <= 0
    rb, ub, vb, wb = b.read_macro()  # the restored fields are the solver's macro arrays
    assert np.array_equal(rb, rho) and np.array_equal(ub, u)
    assert not b.forces().valid  # settling transient: gate closed
    b.step_n(200)
    b.synchronize()
    assert not b.nan_detected()
    # refused on another grid
    other = cases.make_solver(tunnel(nz=10), storage)
    assert not back.restore(other)
    for s in (a, b, other):
        s.close()


def edited_flags(case):
    """A 3x3-cell vane of fluid cells just above the foil becomes solid, the
    foil's last three solid columns open up, and one buried solid cell opens
    with every neighbour still solid (the rest-state branch)."""
    nx, ny, nz = case.dims
    fl = case.flags.reshape(nz, ny, nx)
    new = fl.copy()
    solid = fl[0] == fc.SOLID
    xs = np.nonzero(solid.any(axis=0))[0]
    x0, x1 = int(xs.min()), int(xs.max())
    xm = (x0 + x1) // 2
    ytop = int(np.nonzero(solid[:, xm])[0].max())
    new[:, ytop + 2:ytop + 5, xm:xm + 3] = fc.SOLID
    for x in range(x1 - 2, x1 + 1):
        new[:, solid[:, x], x] = fc.FLUID
    # buried cell: all 18 neighbours solid in the new flags
    ys, xs2 = np.nonzero((new[0] == fc.SOLID))
    buried = None
    for y, x in zip(ys, xs2):
        if 1 <= y < ny - 1 and 1 <= x < nx - 1 and (new[:, y - 1:y + 2, x - 1:x + 2] == fc.SOLID).all():
            buried = (y, x)
            break
    assert buried is not None
    new[nz // 2, buried[0], buried[1]] = fc.FLUID
    return new.reshape(-1)


@pytest.mark.parametrize("storage,n0", [("f32", 40), ("f32", 41), ("f16c", 41), ("f64", 40)])
def test_flag_edit_fixup_matches_oracle(gpu, storage, n0):
    layout = "L" if n0 % 2 == 0 else "S"
    say(f"case: flag-edit fixup, tunnel 96x48x8 {storage}, {n0} steps (layout {layout})")
    case = tunnel()
    s = cases.make_solver(case, storage)
    s.step_n(n0)
    s.synchronize()
    f0 = s.download_f(f64=True).reshape(Q, -1)
    new = edited_flags(case)
    count = s.apply_edited_flags(new)
    f1 = s.download_f(f64=True).reshape(Q, -1)
    old_solid = case.flags == fc.SOLID
    new_solid = new == fc.SOLID
    edited = old_solid != new_solid
    assert count == int(edited.sum()) and count > 0
    say(f"  {int((~old_solid & new_solid).sum())} newly solid, {int((old_solid & ~new_solid).sum())} newly fluid")
    keep = ~edited
    nan = np.isnan(f0[:, keep]) & np.isnan(f1[:, keep])
    assert not np.count_nonzero((f0[:, keep] != f1[:, keep]) & ~nan), "an unedited population changed"
    want = orc.apply_flag_edits(f0, case.dims, case.flags, new)
    g, h = f1[:, edited], want[:, edited]
    assert np.isfinite(g).all() and np.isfinite(h).all()
    err = float(np.abs(g - h).max())
    say(f"  edited cells max |f - oracle|: {err:.2e}")
    assert excess(g, h, storage) <= 0
    assert np.array_equal(s.read_flags(), new)
    assert not s.forces().valid  # the force gate reopens after half a flow-through
    s.step_n(200)
    s.synchronize()
    assert not s.nan_detected()
    s.close()


def test_vg_edit_flow(gpu):
    """The reference's VG-edit flow: capture the clean field, edit, step,
    restore the clean snapshot, then apply the next edit on top of it."""
    say("case: VG edit flow, tunnel 96x48x8 f32")
    case = tunnel()
    key = key_of(case)
    s = cases.make_solver(case, "f32")
    s.step_n(80)
    s.synchronize()
    clean = s.download_f(f64=True)
    snap = S.VramSnapshot()
    snap.capture(s, key)
    vg = edited_flags(case)
    s.apply_edited_flags(vg)
    s.step_n(40)
    assert snap.restore(s)
    restored = s.download_f(f64=True)
    assert same([restored], [clean])
    # back to the clean flags on the restored field: the vane cells reopen
    # (equilibrium from their neighbours), the foil cells close again
    s.apply_edited_flags(case.flags)
    got = s.download_f(f64=True).reshape(Q, -1)
    want = orc.apply_flag_edits(clean.reshape(Q, -1), case.dims, vg, case.flags)
    ed = (vg == fc.SOLID) != (case.flags == fc.SOLID)
    assert excess(got[:, ed], want[:, ed], "f32") <= 0
    s.step_n(100)
    s.synchronize()
    assert not s.nan_detected()
    s.close()
    snap.close()


def test_exact_snapshot_copy_rate(gpu):
    """Capture / restore of a 3.1 M-cell FP16C lattice: device copies at HBM
    rate (reported; asserted only to complete and to restore exactly)."""
    say("case: exact snapshot copy rate, tunnel 256x128x96 f16c")
    case = tunnel(256, 128, 96, 96)
    s = cases.make_solver(case, "f16c")
    s.step_n(11)
    snap = S.VramSnapshot()
    snap.capture(s, key_of(case))
    s.synchronize()
    ref = s.download_f(f64=False)
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        assert snap.restore(s)
    dt = (time.perf_counter() - t0) / reps
    gbs = 2 * snap.nbytes() / dt / 1e9
    say(f"  snapshot {snap.nbytes() / 1e6:.1f} MB, restore {dt * 1e3:.3f} ms incl. sync -> {gbs:.0f} GB/s "
        f"(read + write)")
    assert same([s.download_f(f64=False)], [ref])
    s.close()
    snap.close()
