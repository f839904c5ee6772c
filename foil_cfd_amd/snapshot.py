"""Warm-start snapshots (fc_snapshot_* / fc_vram_* / fc_compact_* C ABI).

Mirrors src/sim/snapshot.h: ``SnapshotKey`` (exact cache key, FNV-1a hash,
printable form), ``hash_flags``, ``VramSnapshot`` (device-resident exact copy
of the populations), ``CompactSnapshot`` (host rho/u/v/w with the FCS1 file
format) and ``DiskSnapshotCache`` (LRU directory of FCS1 files).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from ._abi import FC_SOLVER_VERSION, FcDims, FcScaling, FcSnapshotKey, check, load


@dataclass
class SnapshotKey:
    """SnapshotKey, snapshot.h:33-64."""
    geometry_id: str
    aoa_bucket: int
    u_lat: float
    dims: tuple
    solver_version: int = FC_SOLVER_VERSION

    @staticmethod
    def bucket_for_aoa(aoa_deg: float) -> int:
        return int(load().fc_snapshot_aoa_bucket(float(aoa_deg)))

    def _c(self):
        # keeps the id bytes alive with the struct
        self._id = self.geometry_id.encode()
        return FcSnapshotKey(self._id, int(self.aoa_bucket), float(self.u_lat), FcDims(*self.dims),
                             int(self.solver_version))

    def hash(self) -> int:
        k = self._c()
        return int(load().fc_snapshot_key_hash(C.byref(k)))

    def to_string(self) -> str:
        k = self._c()
        buf = C.create_string_buffer(4400)
        check(load().fc_snapshot_key_string(C.byref(k), buf, 4400), "fc_snapshot_key_string")
        return buf.value.decode()

    def file_name(self) -> str:
        k = self._c()
        buf = C.create_string_buffer(64)
        check(load().fc_snapshot_key_file_name(C.byref(k), buf, 64), "fc_snapshot_key_file_name")
        return buf.value.decode()

    def equals(self, other: "SnapshotKey") -> bool:
        a, b = self._c(), other._c()
        return bool(load().fc_snapshot_key_equal(C.byref(a), C.byref(b)))

    @staticmethod
    def _from_c(k: FcSnapshotKey) -> "SnapshotKey":
        return SnapshotKey(k.geometry_id.decode(), k.aoa_bucket, k.u_lat, (k.dims.nx, k.dims.ny, k.dims.nz),
                           k.solver_version)


def hash_flags(flags: np.ndarray) -> int:
    """hashFlags: FNV-1a 64 of the flag bytes."""
    f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
    return int(load().fc_hash_flags(f.ctypes.data_as(C.POINTER(C.c_uint8)), f.size))


def _err():
    return C.create_string_buffer(1024)


class VramSnapshot:
    """VramSnapshot, snapshot.h:77-122 (one object = one device slot)."""

    def __init__(self):
        self._lib = load()
        self._h = self._lib.fc_vram_snapshot_create()
        if not self._h:
            raise MemoryError("fc_vram_snapshot_create failed")

    def capture(self, solver, key: SnapshotKey) -> None:
        k, e = key._c(), _err()
        check(self._lib.fc_vram_snapshot_capture(self._h, solver._h, C.byref(k), e, 1024),
              "fc_vram_snapshot_capture", e.value.decode())

    def restore(self, solver) -> bool:
        return self._lib.fc_vram_snapshot_restore(self._h, solver._h) == 0

    def has_data(self) -> bool:
        return bool(self._lib.fc_vram_snapshot_has_data(self._h))

    def key(self) -> SnapshotKey:
        k, buf = FcSnapshotKey(), C.create_string_buffer(4097)
        check(self._lib.fc_vram_snapshot_get_key(self._h, C.byref(k), buf, 4097), "fc_vram_snapshot_get_key")
        return SnapshotKey._from_c(k)

    def flags_hash(self) -> int:
        return int(self._lib.fc_vram_snapshot_flags_hash(self._h))

    def nbytes(self) -> int:
        return int(self._lib.fc_vram_snapshot_bytes(self._h))

    def release(self) -> None:
        self._lib.fc_vram_snapshot_release(self._h)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fc_vram_snapshot_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class CompactSnapshot:
    """CompactSnapshot, snapshot.h:124-180."""

    def __init__(self):
        self._lib = load()
        self._h = self._lib.fc_compact_snapshot_create()
        if not self._h:
            raise MemoryError("fc_compact_snapshot_create failed")

    def capture(self, solver, key: SnapshotKey) -> None:
        k, e = key._c(), _err()
        check(self._lib.fc_compact_snapshot_capture(self._h, solver._h, C.byref(k), e, 1024),
              "fc_compact_snapshot_capture", e.value.decode())

    def set(self, key: SnapshotKey, flags_hash: int, scaling: FcScaling, rho, u, v, w) -> None:
        """Fill from host arrays (no solver needed)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in (rho, u, v, w)]
        k = key._c()
        fp = C.POINTER(C.c_float)
        check(self._lib.fc_compact_snapshot_set(self._h, C.byref(k), int(flags_hash), C.byref(scaling),
                                                *[a.ctypes.data_as(fp) for a in arrs], arrs[0].size),
              "fc_compact_snapshot_set")

    def restore(self, solver) -> bool:
        return self._lib.fc_compact_snapshot_restore(self._h, solver._h) == 0

    def save_to_file(self, path) -> None:
        e = _err()
        check(self._lib.fc_compact_snapshot_save(self._h, os.fsencode(path), e, 1024), "fc_compact_snapshot_save",
              e.value.decode())

    def load_from_file(self, path) -> None:
        e = _err()
        check(self._lib.fc_compact_snapshot_load(self._h, os.fsencode(path), e, 1024), "fc_compact_snapshot_load",
              e.value.decode())

    def has_data(self) -> bool:
        return bool(self._lib.fc_compact_snapshot_has_data(self._h))

    def key(self) -> SnapshotKey:
        k, buf = FcSnapshotKey(), C.create_string_buffer(4097)
        check(self._lib.fc_compact_snapshot_get_key(self._h, C.byref(k), buf, 4097),
              "fc_compact_snapshot_get_key")
        return SnapshotKey._from_c(k)

    def flags_hash(self) -> int:
        return int(self._lib.fc_compact_snapshot_flags_hash(self._h))

    def scaling(self) -> FcScaling:
        s = FcScaling()
        check(self._lib.fc_compact_snapshot_scaling(self._h, C.byref(s)), "fc_compact_snapshot_scaling")
        return s

    def fields(self):
        n = int(np.prod(self.key().dims))
        out = [np.empty(n, dtype=np.float32) for _ in range(4)]
        fp = C.POINTER(C.c_float)
        check(self._lib.fc_compact_snapshot_fields(self._h, *[a.ctypes.data_as(fp) for a in out]),
              "fc_compact_snapshot_fields")
        return tuple(out)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fc_compact_snapshot_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class DiskSnapshotCache:
    """DiskSnapshotCache, snapshot.h:182-219."""

    def __init__(self, directory, budget_bytes: int = 10 << 30):
        self._lib = load()
        self._h = self._lib.fc_snapshot_cache_open(os.fsencode(directory), int(budget_bytes))
        if not self._h:
            raise OSError(f"cannot open snapshot cache directory {directory}")

    def find(self, key: SnapshotKey):
        k, buf = key._c(), C.create_string_buffer(4096)
        r = self._lib.fc_snapshot_cache_find(self._h, C.byref(k), buf, 4096)
        if r < 0:
            raise RuntimeError("fc_snapshot_cache_find failed")
        return buf.value.decode() if r == 1 else None

    def store(self, snap: CompactSnapshot) -> None:
        e = _err()
        check(self._lib.fc_snapshot_cache_store(self._h, snap._h, e, 1024), "fc_snapshot_cache_store",
              e.value.decode())

    def store_async(self, snap: CompactSnapshot) -> None:
        check(self._lib.fc_snapshot_cache_store_async(self._h, snap._h), "fc_snapshot_cache_store_async")

    def flush(self) -> int:
        return int(self._lib.fc_snapshot_cache_flush(self._h))

    def set_budget_bytes(self, budget: int) -> None:
        check(self._lib.fc_snapshot_cache_set_budget(self._h, int(budget)), "fc_snapshot_cache_set_budget")

    def used_bytes(self) -> int:
        return int(self._lib.fc_snapshot_cache_used_bytes(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fc_snapshot_cache_close(self._h)
            self._h = None

    def __del__(self):
        self.close()


def aoa_from_bucket(bucket: int) -> float:
    return 0.5 * bucket


__all__ = ["SnapshotKey", "hash_flags", "VramSnapshot", "CompactSnapshot", "DiskSnapshotCache",
           "aoa_from_bucket"]
