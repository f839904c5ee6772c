"""Warm-start snapshots on the CPU (no device calls): FNV-1a against the
published Fowler-Noll-Vo test vectors, the snapshot key's hash / printable
form / AoA bucket against the oracle restatement, the FCS1 file format
written by the library and parsed by an independent reader (and the
reverse), every loader rejection, and the LRU disk cache
(src/sim/snapshot.h, src/sim/snapshot.cpp). The oracle's warm-restart
restatements (equilibrium from macros, flag-edit fixup) are pinned here by
their invariants."""
import os
import time

import numpy as np
import pytest

from foil_cfd_amd import snapshot as S
from foil_cfd_amd._abi import FcScaling
from oracle import oracle as O

# Published FNV-1a 64 test vectors (Fowler / Noll / Vo reference test suite).
FNV_VECTORS = [(b"", 0xCBF29CE484222325), (b"a", 0xAF63DC4C8601EC8C), (b"foobar", 0x85944171F73967E8)]


@pytest.mark.parametrize("data,want", FNV_VECTORS)
def test_fnv1a_published_vectors(data, want):
    assert O.fnv1a64(data) == want
    assert S.hash_flags(np.frombuffer(data, dtype=np.uint8)) == want


def test_hash_flags_matches_oracle_on_a_flag_field():
    rng = np.random.default_rng(3)
    flags = rng.integers(0, 9, size=4097, dtype=np.uint8)
    assert S.hash_flags(flags) == O.fnv1a64(flags.tobytes())
    flags2 = flags.copy()
    flags2[2000] ^= 1
    assert S.hash_flags(flags2) != S.hash_flags(flags)


KEYS = [("naca:2412", 16, 0.08, (768, 320, 96)), ("dat:ls413.dat", -8, 0.05, (512, 192, 64)),
        ("stl:0123456789abcdef", 0, 0.1, (64, 32, 8)), ("", 3, 0.0625, (1, 1, 1)),
        ("a/b\\c:d", 31, 0.07, (100, 50, 20))]


@pytest.mark.parametrize("gid,bucket,ulat,dims", KEYS)
def test_snapshot_key_hash_and_string(gid, bucket, ulat, dims):
    k = S.SnapshotKey(gid, bucket, ulat, dims)
    assert k.hash() == O.snapshot_key_hash(gid, bucket, ulat, dims)
    s = k.to_string()
    assert s == O.snapshot_key_string(gid, bucket, ulat, dims)
    assert ":" not in s and "/" not in s and "\\" not in s
    assert k.file_name() == "%016x.fcs" % k.hash()


def test_snapshot_key_fields_all_enter_the_hash_and_equality():
    base = S.SnapshotKey("naca:2412", 16, 0.08, (768, 320, 96))
    variants = [S.SnapshotKey("naca:0012", 16, 0.08, (768, 320, 96)),
                S.SnapshotKey("naca:2412", 17, 0.08, (768, 320, 96)),
                S.SnapshotKey("naca:2412", 16, 0.0801, (768, 320, 96)),
                S.SnapshotKey("naca:2412", 16, 0.08, (769, 320, 96)),
                S.SnapshotKey("naca:2412", 16, 0.08, (768, 321, 96)),
                S.SnapshotKey("naca:2412", 16, 0.08, (768, 320, 97)),
                S.SnapshotKey("naca:2412", 16, 0.08, (768, 320, 96), 2)]
    hashes = {base.hash()} | {v.hash() for v in variants}
    assert len(hashes) == len(variants) + 1
    assert base.equals(S.SnapshotKey("naca:2412", 16, 0.08, (768, 320, 96)))
    assert not any(base.equals(v) for v in variants)


def test_aoa_bucket_rounds_half_away_from_zero():
    # bucket = lround(2 aoa): halves round away from zero
    cases = {0.0: 0, 0.24: 0, 0.25: 1, -0.25: -1, 1.25: 3, -1.25: -3, 8.0: 16, 7.74: 15, -4.0: -8}
    for aoa, want in cases.items():
        assert S.SnapshotKey.bucket_for_aoa(aoa) == want, aoa


def _scaling():
    s = FcScaling()
    s.chord_m, s.airspeed_ms, s.nu_m2s, s.chord_cells = 1.2, 30.0, 1.48e-5, 256
    s.u_lat, s.dx, s.dt = 0.08, 1.2 / 256, 0.08 * (1.2 / 256) / 30.0
    s.nu_lat_target, s.nu_lat, s.tau, s.tau_clamped = 1.1e-5, 8.3e-4, 0.5025, 1
    s.re_target, s.re_effective = 2.43e6, 24672.0
    return s


def _sc_dict(s):
    return {name: getattr(s, name) for name, _ in O.FCS_SCALING}


def _fields(dims, seed=0):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    return [rng.standard_normal(n).astype(np.float32) for _ in range(4)]


def test_fcs1_written_by_library_parses_with_independent_reader(tmp_path):
    dims = (12, 7, 5)
    key = S.SnapshotKey("naca:2412", 11, 0.08, dims)
    rho, u, v, w = _fields(dims)
    snap = S.CompactSnapshot()
    sc = _scaling()
    snap.set(key, 0x0123456789ABCDEF, sc, rho, u, v, w)
    p = tmp_path / "a.fcs"
    snap.save_to_file(p)
    d = O.fcs1_read(p)
    assert d["magic"] == 0x31534346 and d["format"] == 1 and d["compression"] == 0
    assert open(p, "rb").read(4) == b"FCS1"
    assert d["geometry_id"] == "naca:2412" and d["aoa_bucket"] == 11 and d["dims"] == dims
    assert d["u_lat"] == np.float32(0.08) and d["solver_version"] == 1
    assert d["flags_hash"] == 0x0123456789ABCDEF
    want = _sc_dict(sc)
    for name, _ in O.FCS_SCALING:
        assert np.float32(d["scaling"][name]) == np.float32(want[name]), name
    assert d["ncells"] == rho.size and d["trailing_bytes"] == 0
    for name, a in zip(("rho", "u", "v", "w"), (rho, u, v, w)):
        assert np.array_equal(d[name], a)
    # header size: 12 + id + 24 + 8 + 49 + 12
    assert os.path.getsize(p) == 12 + len("naca:2412") + 24 + 8 + 49 + 12 + 16 * rho.size


def test_fcs1_written_by_independent_writer_loads(tmp_path):
    dims = (9, 4, 3)
    rho, u, v, w = _fields(dims, 1)
    sc = _sc_dict(_scaling())
    p = tmp_path / "b.fcs"
    O.fcs1_write(p, "stl:00ff00ff00ff00ff", -5, 0.05, dims, 0xFEEDFACECAFEBEEF, sc, rho, u, v, w)
    snap = S.CompactSnapshot()
    snap.load_from_file(p)
    assert snap.has_data()
    k = snap.key()
    assert (k.geometry_id, k.aoa_bucket, k.dims, k.solver_version) == ("stl:00ff00ff00ff00ff", -5, dims, 1)
    assert np.float32(k.u_lat) == np.float32(0.05)
    assert snap.flags_hash() == 0xFEEDFACECAFEBEEF
    got = snap.scaling()
    for name, _ in O.FCS_SCALING:
        assert np.float32(getattr(got, name)) == np.float32(sc[name]), name
    for a, b in zip(snap.fields(), (rho, u, v, w)):
        assert np.array_equal(a, b)
    # and a save of what was loaded is byte-identical
    p2 = tmp_path / "b2.fcs"
    snap.save_to_file(p2)
    assert open(p, "rb").read() == open(p2, "rb").read()


@pytest.mark.parametrize("override,reason", [
    ({"magic": 0x32534346}, "bad magic"),
    ({"fmt": 2}, "unsupported format version"),
    ({"version": 2}, "solver version mismatch"),
    ({"compression": 1}, "unsupported payload compression"),
    ({"ncells": 0}, "payload size mismatch"),
    ({"ncells": 9 * 4 * 3 + 1}, "payload size mismatch"),
])
def test_fcs1_loader_rejections(tmp_path, override, reason):
    dims = (9, 4, 3)
    rho, u, v, w = _fields(dims, 2)
    p = tmp_path / "bad.fcs"
    O.fcs1_write(p, "naca:0012", 0, 0.08, dims, 1, _sc_dict(_scaling()), rho, u, v, w, **override)
    snap = S.CompactSnapshot()
    with pytest.raises(RuntimeError, match=reason):
        snap.load_from_file(p)
    assert not snap.has_data()


def test_fcs1_loader_rejects_truncation_and_long_ids(tmp_path):
    dims = (9, 4, 3)
    rho, u, v, w = _fields(dims, 4)
    p = tmp_path / "t.fcs"
    O.fcs1_write(p, "naca:0012", 0, 0.08, dims, 1, _sc_dict(_scaling()), rho, u, v, w)
    full = open(p, "rb").read()
    hdr = 12 + len("naca:0012")
    for cut, reason in ((hdr + 10, "truncated header"), (hdr + 24 + 8 + 20, "truncated scaling block"),
                        (hdr + 24 + 8 + 49 + 6, "truncated payload header"), (len(full) - 3,
                                                                              "truncated field payload")):
        q = tmp_path / f"cut{cut}.fcs"
        open(q, "wb").write(full[:cut])
        with pytest.raises(RuntimeError, match=reason):
            S.CompactSnapshot().load_from_file(q)
    q = tmp_path / "longid.fcs"
    O.fcs1_write(q, "x" * 4097, 0, 0.08, dims, 1, _sc_dict(_scaling()), rho, u, v, w)
    with pytest.raises(RuntimeError, match="corrupt geometry id"):
        S.CompactSnapshot().load_from_file(q)
    q = tmp_path / "id4096.fcs"
    O.fcs1_write(q, "y" * 4096, 0, 0.08, dims, 1, _sc_dict(_scaling()), rho, u, v, w)
    s = S.CompactSnapshot()
    s.load_from_file(q)
    assert s.key().geometry_id == "y" * 4096
    with pytest.raises(RuntimeError, match="cannot open"):
        S.CompactSnapshot().load_from_file(tmp_path / "missing.fcs")


def test_compact_snapshot_set_rejects_size_mismatch():
    snap = S.CompactSnapshot()
    rho, u, v, w = _fields((4, 4, 4))
    with pytest.raises(RuntimeError):
        snap.set(S.SnapshotKey("g", 0, 0.08, (4, 4, 5)), 0, _scaling(), rho, u, v, w)
    assert not snap.has_data()
    with pytest.raises(RuntimeError):
        snap.save_to_file("/nonexistent-dir/x.fcs")


def _snap(gid, dims=(8, 4, 4), seed=0):
    s = S.CompactSnapshot()
    s.set(S.SnapshotKey(gid, 0, 0.08, dims), 7, _scaling(), *_fields(dims, seed))
    return s


def test_disk_cache_lru_eviction_and_hits(tmp_path):
    d = tmp_path / "cache"
    one = None
    c = S.DiskSnapshotCache(d, budget_bytes=10 << 30)
    a = _snap("geom:a")
    c.store(a)
    one = c.used_bytes()
    assert one == os.path.getsize(d / a.key().file_name())
    c.close()
    # budget for two entries; timestamps spaced so the order is unambiguous
    c = S.DiskSnapshotCache(d, budget_bytes=2 * one + one // 2)
    assert c.used_bytes() == one  # startup scan
    b, e, f = _snap("geom:b", seed=1), _snap("geom:c", seed=2), _snap("geom:d", seed=3)
    past = time.time() - 100
    os.utime(d / a.key().file_name(), (past, past))
    c.store(b)
    os.utime(d / b.key().file_name(), (past + 10, past + 10))
    assert c.find(a.key()) is not None  # hit refreshes a: b is now the oldest
    c.store(e)
    assert c.find(b.key()) is None and c.find(a.key()) is not None and c.find(e.key()) is not None
    assert c.used_bytes() == 2 * one
    # re-storing an existing key replaces it without growing the total
    c.store(e)
    assert c.used_bytes() == 2 * one
    # asynchronous store from the worker thread
    c.store_async(f)
    f.close()  # the queue holds its own reference to the data
    assert c.flush() == 0
    assert c.find(S.SnapshotKey("geom:d", 0, 0.08, (8, 4, 4))) is not None
    assert c.used_bytes() == 2 * one
    hit = c.find(S.SnapshotKey("geom:d", 0, 0.08, (8, 4, 4)))
    back = S.CompactSnapshot()
    back.load_from_file(hit)
    assert np.array_equal(back.fields()[0], _fields((8, 4, 4), 3)[0])
    # shrinking the budget evicts immediately, nothing protected
    c.set_budget_bytes(one)
    assert c.used_bytes() == one
    c.set_budget_bytes(0)
    assert c.used_bytes() == 0 and not [p for p in os.listdir(d) if p.endswith(".fcs")]
    c.close()


def test_disk_cache_store_keeps_a_single_oversized_entry(tmp_path):
    c = S.DiskSnapshotCache(tmp_path / "c2", budget_bytes=10)
    s = _snap("big")
    c.store(s)  # never evicts the entry just stored
    assert c.find(s.key()) is not None and c.used_bytes() > 10
    c.close()


# ---- oracle restatements pinned by invariants -------------------------------------
def test_oracle_init_from_macro_moments():
    rng = np.random.default_rng(7)
    n = 200
    rho = 1 + 0.05 * rng.standard_normal(n)
    u, v, w = (0.05 * rng.standard_normal(n) for _ in range(3))
    flags = np.zeros(n, np.uint8)
    flags[::7] = O.SOLID_FLAG
    f = O.init_from_macro(flags, rho, u, v, w)
    fl = flags == O.SOLID_FLAG
    r = np.where(fl, 1.0, rho)
    U = [np.where(fl, 0.0, a) for a in (u, v, w)]
    cs = (O.CX, O.CY, O.CZ)
    assert np.allclose(f.sum(0), r, rtol=0, atol=1e-14)
    for a in range(3):
        assert np.allclose((cs[a][:, None] * f).sum(0), r * U[a], atol=1e-14)
        for b in range(3):
            pi = (cs[a][:, None] * cs[b][:, None] * f).sum(0)
            assert np.allclose(pi, r * (U[a] * U[b] + (a == b) / 3.0), atol=1e-14)
    assert np.array_equal(f[:, fl], np.tile(O.W[:, None], (1, fl.sum())))


def _linear_field(dims, g=(1e-3, -2e-3, 5e-4), u0=(0.05, 0.01, -0.02)):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = (a.reshape(-1).astype(np.float64) for a in (x, y, z))
    rho = 1 + g[0] * x + g[1] * y + g[2] * z
    u = u0[0] + 1e-3 * x - 2e-3 * z
    v = u0[1] + 3e-4 * y
    w = u0[2] + 5e-4 * x - 1e-4 * y
    return rho, u, v, w


def test_oracle_flag_edit_linear_field_recovers_centre_state():
    # Full symmetric neighbourhood of a linear field: the mean of rho and of
    # u over the 18 links equals the centre value (sum of c_q = 0), so the
    # fixed cell gets the equilibrium of the field at its own position. A
    # wrong neighbour offset or sign breaks the symmetry.
    dims = (9, 9, 9)
    n = 729
    rho, u, v, w = _linear_field(dims)
    f = O.equilibrium_np(rho, u, v, w)
    old = np.zeros(n, np.uint8)
    c = 4 + 9 * (4 + 9 * 4)
    old[c] = O.SOLID_FLAG
    new = np.zeros(n, np.uint8)
    out = O.apply_flag_edits(f, dims, old, new)
    want = O.equilibrium_np([rho[c]], [u[c]], [v[c]], [w[c]])[:, 0]
    assert np.allclose(out[:, c], want, rtol=0, atol=1e-15)
    mask = np.ones(n, bool)
    mask[c] = False
    assert np.array_equal(out[:, mask], f[:, mask])


def test_oracle_flag_edit_solid_enclosed_wrap_and_exclusions():
    dims = (6, 6, 4)
    nx, ny, nz = dims
    n = nx * ny * nz
    idx = lambda x, y, z: x + nx * (y + ny * z)  # noqa: E731
    rng = np.random.default_rng(11)
    rho = 1 + 0.01 * rng.standard_normal(n)
    u, v, w = (0.02 * rng.standard_normal(n) for _ in range(3))
    f = O.equilibrium_np(rho, u, v, w)
    old = np.zeros(n, np.uint8)
    new = np.zeros(n, np.uint8)
    # newly solid -> zeros
    a = idx(1, 1, 1)
    new[a] = O.SOLID_FLAG
    # enclosed: solid block 3x3x4 (all z) around (4,4,*); centre column opens
    for x in range(3, 6):
        for y in range(3, 6):
            for z in range(nz):
                old[idx(x, y, z)] = new[idx(x, y, z)] = O.SOLID_FLAG
    b = idx(4, 4, 2)
    new[b] = 0
    # z wrap: cell at z=0 whose only fluid neighbour is (0, 4, nz-1) across the wrap
    for x in range(0, 3):
        for y in range(3, 6):
            for z in range(nz):
                old[idx(x, y, z)] = new[idx(x, y, z)] = O.SOLID_FLAG
    c = idx(1, 4, 0)
    new[c] = 0
    new[idx(1, 4, nz - 1)] = old[idx(1, 4, nz - 1)] = 0
    out = O.apply_flag_edits(f, dims, old, new)
    assert np.array_equal(out[:, a], np.zeros(19))
    assert np.allclose(out[:, b], O.W)
    nbr = idx(1, 4, nz - 1)
    want = O.equilibrium_np([rho[nbr]], [u[nbr]], [v[nbr]], [w[nbr]])[:, 0]
    assert np.allclose(out[:, c], want, atol=1e-15)
    # an edited neighbour is excluded: open (1,4,nz-1)'s partner too
    new2 = new.copy()
    old2 = old.copy()
    old2[nbr] = O.SOLID_FLAG  # now nbr is itself newly fluid -> excluded for c
    out2 = O.apply_flag_edits(f, dims, old2, new2)
    assert np.allclose(out2[:, c], O.W)
    # flag changes that keep the solid state are not edits
    new3 = new.copy()
    new3[idx(2, 1, 1)] = 3
    out3 = O.apply_flag_edits(f, dims, old, new3)
    assert np.array_equal(out3[:, idx(2, 1, 1)], f[:, idx(2, 1, 1)])
