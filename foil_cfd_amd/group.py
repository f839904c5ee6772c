"""Batched stepping of independent solvers (fc_group_* C ABI).

No reference counterpart: the reference steps one solver at a time. A group
steps several initialized ``LBMSolver`` objects with one batched step launch
per step (captured into a HIP graph and replayed); ``step(n)`` is equivalent
to ``s.step_n(n)`` on every member. ``polar_sweep`` builds the cases of an
angle-of-attack polar with uneven grids.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import cases as _cases
from ._abi import FcGroupInfo, check, load


def _info_dict(i: FcGroupInfo) -> dict:
    return {"members": i.members, "cells": int(i.cells), "waves_per_block": i.waves_per_block,
            "slab_order": i.slab_order, "use_graph": i.use_graph,
            "cache_budget": int(i.cache_budget), "sub_batches": i.sub_batches,
            "last_step_ms": float(i.last_step_ms), "glups": float(i.glups)}


class SolverGroup:
    def __init__(self, solvers):
        self._lib = load()
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s._h for s in self.solvers])
        err = C.create_string_buffer(256)
        h = self._lib.fc_group_create(arr, len(self.solvers), err, 256)
        if not h:
            raise RuntimeError(f"fc_group_create failed: {err.value.decode(errors='replace')}")
        self._h = h

    def _check(self, rc, what):
        if rc != 0:
            buf = C.create_string_buffer(256)
            self._lib.fc_group_last_error(self._h, buf, 256)
            check(rc, what, buf.value.decode(errors="replace"))

    def step(self, n: int):
        self._check(self._lib.fc_group_step(self._h, n), "fc_group_step")

    def set_launch(self, waves_per_block: int = 4, slab_order: bool = True, use_graph: bool = True):
        self._check(self._lib.fc_group_set_launch(self._h, waves_per_block, int(slab_order), int(use_graph)),
                    "fc_group_set_launch")

    def set_cache_budget(self, nbytes: int):
        self._check(self._lib.fc_group_set_cache_budget(self._h, int(nbytes)), "fc_group_set_cache_budget")

    def last_step_ms(self) -> float:
        ms = C.c_double(0.0)
        self._check(self._lib.fc_group_last_step_ms(self._h, C.byref(ms)), "fc_group_last_step_ms")
        return float(ms.value)

    def info(self) -> dict:
        i = FcGroupInfo()
        self._check(self._lib.fc_group_get_info(self._h, C.byref(i)), "fc_group_get_info")
        return _info_dict(i)

    def autotune(self, steps_per_trial: int = 20) -> dict:
        i = FcGroupInfo()
        self._check(self._lib.fc_group_autotune(self._h, steps_per_trial, C.byref(i)), "fc_group_autotune")
        return _info_dict(i)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fc_group_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def lowbias32(x: int) -> int:
    """The lowbias32 integer hash (32-bit)."""
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def polar_chords(count: int = 16, lo: int = 96, hi: int = 192, seed: int = 0xF01DCAFE):
    """Chords (cells) of the polar cases: a fixed-seed lowbias32 draw in [lo, hi]."""
    return [lo + lowbias32(seed + 0x9E3779B9 * (i + 1)) % (hi - lo + 1) for i in range(count)]


def polar_sweep(count: int = 16, aoa_lo: float = -4.0, aoa_hi: float = 16.0, chords=None, u_lat: float = 0.08,
                cs: float = 0.1733, ramp: bool = True, tau: float | None = None):
    """The cases of an angle-of-attack polar: NACA 2412, AoA evenly spaced in
    [aoa_lo, aoa_hi], grid (2c, 0.75c, 0.9c) for chord c (uneven grids)."""
    chords = polar_chords(count) if chords is None else list(chords)
    aoas = np.linspace(aoa_lo, aoa_hi, count) if count > 1 else np.array([aoa_lo])
    out = []
    for c, a in zip(chords, aoas):
        nx, ny, nz = 2 * c, int(round(0.75 * c)), int(round(0.9 * c))
        out.append(_cases.naca_tunnel(nx, ny, nz, c, aoa=float(a), u_lat=u_lat, cs=cs, ramp=ramp, tau=tau))
    return out
