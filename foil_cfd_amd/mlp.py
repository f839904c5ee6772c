"""Learned u_tau network: GPU training and batched inference (fc_mlp_* C ABI).

No reference counterpart: the reference solves Reichardt's law for u_tau by
Newton in every wall cell (src/sim/lbm_wallmodel.cu:55-73). Here the solver
harvests (y_s, |u_t|, nu, Newton u_tau) samples from its own wall cells
(``Solver.wall_harvest``), ``MlpTrainer`` fits the 3 -> 32 -> 32 -> 1 network
on the GPU with the library's forward / backward / Adam kernels, and
``export()`` produces the parameter block that ``Solver.set_wall_mlp`` loads
for the ``"mlp"`` wall-model mode (Newton stays the fallback outside the
training envelope).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import FC_MLP_NPARAMS, FcMlpParams, check, load

NPARAMS = FC_MLP_NPARAMS


def _f32(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MlpTrainer:
    """Harvested samples on the GPU plus FP32 master parameters and Adam state."""

    def __init__(self, samples, seed: int = 0x5EED):
        self._lib = load()
        smp = np.ascontiguousarray(samples, dtype=np.float32).reshape(-1, 4)
        err = C.create_string_buffer(256)
        h = self._lib.fc_mlp_trainer_create(_f32(smp), smp.shape[0], seed & 0xFFFFFFFF, err, 256)
        if not h:
            raise RuntimeError(f"fc_mlp_trainer_create failed: {err.value.decode(errors='replace')}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fc_mlp_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def samples(self) -> int:
        return int(self._lib.fc_mlp_trainer_samples(self._h))

    def train(self, steps: int, batch: int = 1024, lr: float = 3e-3) -> np.ndarray:
        """Runs `steps` Adam steps; returns each step's batch mean squared error."""
        losses = np.zeros(max(steps, 1), dtype=np.float32)
        check(self._lib.fc_mlp_trainer_train(self._h, steps, batch, lr, _f32(losses)),
              "fc_mlp_trainer_train")
        return losses[:steps]

    def gradient(self, idx, params=None):
        """(grad[1217], sum of squared errors) over the samples idx (multiple of 32)."""
        ii = np.ascontiguousarray(idx, dtype=np.int32)
        g = np.zeros(NPARAMS, dtype=np.float32)
        sse = C.c_float(0.0)
        p = None
        if params is not None:
            p = np.ascontiguousarray(params, dtype=np.float32)
            assert p.size == NPARAMS
        check(self._lib.fc_mlp_trainer_gradient(
            self._h, None if p is None else _f32(p), ii.ctypes.data_as(C.POINTER(C.c_int)), ii.size,
            _f32(g), C.byref(sse)), "fc_mlp_trainer_gradient")
        return g, float(sse.value)

    def params(self) -> np.ndarray:
        p = np.zeros(NPARAMS, dtype=np.float32)
        check(self._lib.fc_mlp_trainer_get_params(self._h, _f32(p)), "fc_mlp_trainer_get_params")
        return p

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == NPARAMS
        check(self._lib.fc_mlp_trainer_set_params(self._h, _f32(p)), "fc_mlp_trainer_set_params")

    def data(self):
        """(normalised inputs (n, 3), normalised targets (n,)) as held on the GPU."""
        n = self.samples
        x = np.zeros(3 * n, dtype=np.float32)
        t = np.zeros(n, dtype=np.float32)
        check(self._lib.fc_mlp_trainer_data(self._h, _f32(x), _f32(t)), "fc_mlp_trainer_data")
        return x.reshape(3, n).T.copy(), t

    def export(self) -> FcMlpParams:
        out = FcMlpParams()
        check(self._lib.fc_mlp_trainer_export(self._h, C.byref(out)), "fc_mlp_trainer_export")
        return out

    def fit(self, steps: int = 4000, batch: int = 1024, lr: float = 3e-3, final_lr: float = 1e-4):
        """Adam with a geometric learning-rate decay over four phases."""
        losses = []
        phases = 4
        for k in range(phases):
            rate = lr * (final_lr / lr) ** (k / (phases - 1))
            losses.append(self.train(steps // phases, batch, rate))
        return np.concatenate(losses)


def infer(params: FcMlpParams, samples) -> np.ndarray:
    """Network u_tau of (y_s, |u_t|, nu[, -]) rows on the GPU (-1 outside the envelope)."""
    lib = load()
    smp = np.asarray(samples, dtype=np.float32)
    if smp.shape[-1] == 3:
        smp = np.concatenate([smp, np.zeros((smp.shape[0], 1), np.float32)], axis=1)
    smp = np.ascontiguousarray(smp.reshape(-1, 4))
    out = np.zeros(smp.shape[0], dtype=np.float32)
    check(lib.fc_mlp_infer(C.byref(params), _f32(smp), smp.shape[0], _f32(out)), "fc_mlp_infer")
    return out


def params_to_dict(p: FcMlpParams) -> dict:
    """Readable copy of an exported parameter block (BF16 weights widened to float)."""
    def bf(a, n):
        bits = np.ctypeslib.as_array(a, shape=(n,)).astype(np.uint32) << 16
        return bits.view(np.float32).copy()
    return {
        "in_mean": np.array(p.in_mean[:3], np.float32), "in_inv_std": np.array(p.in_inv_std[:3], np.float32),
        "env_lo": np.array(p.env_lo[:], np.float32), "env_hi": np.array(p.env_hi[:], np.float32),
        "out_mean": float(p.out_mean), "out_std": float(p.out_std),
        "w1": bf(p.w1, 96).reshape(32, 3), "b1": np.array(p.b1[:], np.float32),
        "w2": bf(p.w2, 1024).reshape(32, 32), "b2": np.array(p.b2[:], np.float32),
        "w3": bf(p.w3, 32), "b3": float(p.b3),
    }
