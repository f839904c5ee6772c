"""Learned u_tau network on the GPU (fc_mlp_* through the C ABI).

* gradients of the library's forward / backward kernels (BF16 matrix-core
  tiles, FP32 accumulate) against the CPU oracle's restatement with the same
  BF16 operand roundings, at the initial and at a trained parameter point;
* Adam's first step (|dP| = lr wherever the gradient is not tiny);
* the fused wave-cooperative inference against the oracle forward pass;
* training on samples the solver harvests from its own wall cells, and the
  trained network serving the wall model (Newton as fallback and arbiter);
* what the network gains over the Newton solve without it, measured against
  the exact root of Reichardt's law (one Newton iteration saved from the
  reference's analytic start, which already converges in two to three).
"""
import sys

import numpy as np
import pytest

from foil_cfd_amd import cases, mlp
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KAPPA, REICH_C = 0.41, 7.8  # Reichardt's law, lbm_wallmodel.cuh:27-29


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def exact_utau(ut, ys, nu):
    """Root of u_tau * u+(y_s u_tau / nu) = u_t by FP64 bisection (u+ is
    Reichardt's law, monotone in u_tau)."""
    ut, ys, nu = (np.asarray(a, np.float64) for a in (ut, ys, nu))
    lo = np.full(ut.shape, 1e-14)
    hi = np.maximum(ut, 10.0 * np.sqrt(nu * ut / ys))
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        yp = ys * mid / nu
        up = np.log1p(KAPPA * yp) / KAPPA + REICH_C * (1 - np.exp(-yp / 11) - yp / 11 * np.exp(-yp / 3))
        big = mid * up > ut
        hi = np.where(big, mid, hi)
        lo = np.where(big, lo, mid)
    return 0.5 * (lo + hi)


def synthetic_samples(n, rng):
    """Wall samples spanning the lattice range of the reference cases: y_s
    0.5..8 cells, |u_t| 1e-3..0.1, nu 10^-5.5..10^-2.5 (Re_y 0.2..1.6e6)."""
    ys = 10 ** rng.uniform(np.log10(0.5), np.log10(8.0), n)
    ut = 10 ** rng.uniform(-3, -1, n)
    nu = 10 ** rng.uniform(-5.5, -2.5, n)
    return np.stack([ys, ut, nu, exact_utau(ut, ys, nu)], axis=1).astype(np.float32)


def blocks(g):
    return {"w1": g[0:96], "b1": g[96:128], "w2": g[128:1152], "b2": g[1152:1184],
            "w3": g[1184:1216], "b3": g[1216:1217]}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def check_gradient(tr, x, t, idx, params=None, tol=3e-3, vs_exact=None):
    g, sse = tr.gradient(idx, params)
    p = tr.params() if params is None else params
    go, sseo = orc.mlp_gradient(p, x, t, idx, emulate=True)
    gx, _ = orc.mlp_gradient(p, x, t, idx, emulate=False)
    r = rel(g, go)
    say(f"  batch {idx.size}: rel L2 vs oracle(bf16) {r:.2e}, vs exact FP64 {rel(g, gx):.2e}, "
        f"sse {sse:.6g} vs {sseo:.6g}")
    assert r < tol
    for name, gb in blocks(g).items():
        ob = blocks(go)[name]
        assert rel(gb, ob) < 10 * tol, (name, rel(gb, ob))
    assert sse == pytest.approx(sseo, rel=1e-4)
    # Against the all-FP64 network: away from a fit the BF16 operand rounding
    # moves the gradient by ~1e-2; near a fit the residuals are as small as the
    # rounding's effect on the output, so the two gradients part (not checked).
    if vs_exact is not None:
        assert rel(g, gx) < vs_exact
    return g


def test_gradient_matches_oracle(gpu):
    say("case: MLP gradient vs oracle, 4096 synthetic samples")
    rng = np.random.default_rng(11)
    tr = mlp.MlpTrainer(synthetic_samples(4096, rng), seed=1234)
    x, t = tr.data()
    assert x.shape == (4096, 3) and np.all(np.isfinite(x)) and np.all(np.isfinite(t))
    # initial (Glorot) point, one tile and many tiles
    check_gradient(tr, x, t, rng.integers(0, 4096, 32).astype(np.int32), vs_exact=0.05)
    check_gradient(tr, x, t, rng.integers(0, 4096, 1024).astype(np.int32), vs_exact=0.05)
    # a trained point
    losses = tr.train(300, 512, 3e-3)
    assert losses[-20:].mean() < 0.5 * losses[:20].mean()
    check_gradient(tr, x, t, rng.integers(0, 4096, 512).astype(np.int32))
    # explicit parameters (not the trainer's own)
    p = (rng.standard_normal(mlp.NPARAMS) * 0.3).astype(np.float32)
    check_gradient(tr, x, t, rng.integers(0, 4096, 256).astype(np.int32), params=p)
    tr.close()


def test_adam_first_step(gpu):
    say("case: MLP Adam first step")
    rng = np.random.default_rng(12)
    tr = mlp.MlpTrainer(synthetic_samples(2048, rng), seed=7)
    p0 = tr.params()
    tr.set_params(p0)  # restarts Adam
    lr = 1e-3
    tr.train(1, 1024, lr)
    d = np.abs(tr.params().astype(np.float64) - p0)
    # first step = lr * g / (|g| + 1e-8): |dP| = lr up to FP32 rounding of P
    # and the eps term (parameters whose batch gradient is below ~1e-5)
    ok = np.abs(d - lr) <= lr * 2e-3 + 2.0 ** -23 * np.abs(p0)
    say(f"  {ok.mean():.4f} of parameters moved by lr")
    assert ok.mean() > 0.98
    assert np.all(d <= lr * (1 + 2e-3) + 2.0 ** -23 * np.abs(p0))
    tr.close()


def test_inference_matches_oracle_forward(gpu):
    say("case: MLP inference vs oracle forward")
    rng = np.random.default_rng(13)
    tr = mlp.MlpTrainer(synthetic_samples(8192, rng), seed=99)
    tr.fit(1200, 512, 3e-3, 3e-4)
    prm = tr.export()
    P = tr.params()
    d = mlp.params_to_dict(prm)
    # exported BF16 weights are round-to-nearest-even of the master weights
    for name, sl in (("w1", slice(0, 96)), ("w2", slice(128, 1152)), ("w3", slice(1184, 1216))):
        want = np.array([orc.bf16_round(float(v)) for v in P[sl]], np.float32)
        assert np.array_equal(d[name].reshape(-1), want)
    # 2000 samples: a partial last wavefront (2000 = 31 * 64 + 16)
    held = synthetic_samples(2000, np.random.default_rng(14))
    u = mlp.infer(prm, held)
    inside = u > 0
    assert inside.mean() > 0.97
    x = ((np.log(held[:, :3].astype(np.float64)) - d["in_mean"]) * d["in_inv_std"]).astype(np.float32)
    y = orc.mlp_forward(P, x, emulate=True)
    uo = held[:, 1] * np.exp(y * d["out_std"] + d["out_mean"])
    e = np.abs(u[inside] - uo[inside]) / uo[inside]
    say(f"  GPU vs oracle u_tau: median {np.median(e):.2e} max {e.max():.2e}")
    assert np.median(e) < 2e-5 and e.max() < 3e-3
    # outside the envelope the network declines
    far = held[:64].copy()
    far[:, 0] *= 1e3
    assert np.all(mlp.infer(prm, far) == -1.0)
    tr.close()


def test_trained_network_gain_over_newton(gpu):
    """The network's estimate against the exact root, and Newton from the
    network's estimate against Newton from the reference's cold start."""
    say("case: MLP gain over Newton, 65536 training / 4096 held-out samples")
    rng = np.random.default_rng(21)
    tr = mlp.MlpTrainer(synthetic_samples(65536, rng), seed=5)
    losses = tr.fit(8000, 1024, 3e-3, 5e-5)
    say(f"  loss {losses[:50].mean():.3e} -> {losses[-200:].mean():.3e}")
    prm = tr.export()
    held = synthetic_samples(4096, np.random.default_rng(22))
    exact = held[:, 3].astype(np.float64)
    net = mlp.infer(prm, held)
    inside = net > 0
    assert inside.mean() > 0.97
    h, ex, nt = held[inside], exact[inside], net[inside].astype(np.float64)
    e_net = np.abs(nt - ex) / ex

    def newton(seed, k):
        return np.array([orc.solve_utau_iters(float(a[1]), float(a[0]), float(a[2]), float(s), k)
                         for a, s in zip(h, seed)], np.float64)

    zero = np.zeros(len(h))
    rows = {}
    for k in (1, 2, 3, 4, 6):
        rows[f"cold+{k}"] = np.abs(newton(zero, k) - ex) / ex
    for k in (1, 2):
        rows[f"net+{k}"] = np.abs(newton(nt, k) - ex) / ex
    say(f"  network alone: median {np.median(e_net):.2e}  p99 {np.quantile(e_net, 0.99):.2e}  "
        f"max {e_net.max():.2e}")
    for name, e in rows.items():
        say(f"  Newton {name:7s}: median {np.median(e):.2e}  p99 {np.quantile(e, 0.99):.2e}  max {e.max():.2e}")
    assert np.median(e_net) < 3e-3 and np.quantile(e_net, 0.99) < 2e-2
    # What the network gains over the Newton solve without it. Alone it is
    # more accurate than one Newton step from the reference's analytic start
    # (median); as a seed, one step from it matches two from the analytic
    # start (median) and is 100x tighter than one (p99), and two steps reach
    # the FP32 floor that takes three without it.
    assert np.median(e_net) < np.median(rows["cold+1"])
    assert np.median(rows["net+1"]) <= np.median(rows["cold+2"])
    assert np.quantile(rows["net+1"], 0.99) < np.quantile(rows["cold+1"], 0.99) / 100
    assert rows["net+2"].max() <= 2 * rows["cold+3"].max() + 1e-7
    tr.close()


def test_train_on_harvest_and_serve_in_solver(gpu):
    say("case: harvest -> GPU training -> MLP wall model, tunnel 192x64x16")
    case = cases.naca_tunnel(192, 64, 16, 48, aoa=8.0, cs=0.1, ramp=False)
    s = cases.make_solver(case, "f32")
    s.step_n(200)
    s.set_wall_model("newton")
    s.set_wall_harvest(True)
    got = []
    for _ in range(30):
        s.step_n(10)
        s.synchronize()
        hv = s.wall_harvest()
        got.append(hv[hv[:, 3] > 0])
    smp = np.concatenate(got)
    say(f"  harvested {len(smp)} samples from {s.wall_cell_count()} wall cells")
    assert len(smp) > 1000
    tr = mlp.MlpTrainer(smp, seed=3)
    assert tr.samples == len(smp)
    tr.fit(3000, 512, 3e-3, 1e-4)
    prm = tr.export()
    s.set_wall_mlp(prm)
    s.set_wall_model("mlp")
    s.step_n(10)
    s.synchronize()
    r = s.wall_model_readout()
    say(f"  wall cells {r.cells}, served by the network {r.mlp_cells}")
    assert r.mode == 2 and r.mlp_cells > 0.8 * r.cells
    newton = s.wall_harvest()  # Newton solution seeded by the network
    live = newton[:, 3] > 0
    net = mlp.infer(prm, newton[live])
    ok = net > 0
    e = np.abs(net[ok] - newton[live][ok, 3]) / newton[live][ok, 3]
    say(f"  network vs Newton: median {np.median(e):.2e} max {e.max():.2e}")
    assert np.median(e) < 5e-3 and e.max() < 5e-2
    s.step_n(200)
    s.synchronize()
    assert not s.nan_detected()
    f = s.forces()
    assert np.isfinite(f.cl) and np.isfinite(f.cd)
    tr.close()
    s.close()
