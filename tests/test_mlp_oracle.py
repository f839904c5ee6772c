"""Pins of the oracle's learned-u_tau network restatement (CPU only).

The network has no reference counterpart, so the oracle is pinned by the
mathematics: its reverse-mode gradient against central finite differences of
its own loss (FP64, no BF16 emulation), the forward pass against a closed
form with a single active path, BF16 rounding against torch's conversion,
and Adam against the constant-gradient property of bias correction (every
step moves each parameter by lr * sign(g)).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc

NP = orc.MLP_NPARAMS


def random_params(rng, scale=0.5):
    p = np.zeros(NP, dtype=np.float32)
    p[0:96] = rng.uniform(-1.2, 1.2, 96)
    p[96:128] = rng.uniform(-0.3, 0.3, 32)
    p[128:1152] = rng.uniform(-scale, scale, 1024)
    p[1152:1184] = rng.uniform(-0.3, 0.3, 32)
    p[1184:1216] = rng.uniform(-0.6, 0.6, 32)
    p[1216] = 0.1
    return p


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 4000),
                        # exact ties: bit 15 set, low bits clear, both parities of bit 16
                        (np.array([0x3F808000, 0x3F818000, 0xBF808000, 0x40490000 | 0x8000],
                                  dtype=np.uint32)).view(np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    got = np.array([orc.bf16_round(float(v)) for v in x], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_forward_single_path_closed_form():
    # W1[0] = (a, 0, 0), W2[0][0] = b, w3[0] = c, b3 = d, all else zero:
    # y = d + c tanh(b tanh(a x0)).
    a, b, c, d = 0.75, -1.5, 0.625, 0.25
    p = np.zeros(NP, dtype=np.float32)
    p[0] = a
    p[128] = b
    p[1184] = c
    p[1216] = d
    x = np.array([[-2.0, 0.3, 5.0], [0.0, 1.0, 1.0], [1.7, -4.0, 0.0]], dtype=np.float32)
    want = d + c * np.tanh(b * np.tanh(a * x[:, 0].astype(np.float64)))
    np.testing.assert_allclose(orc.mlp_forward(p, x, emulate=False), want, rtol=0, atol=1e-15)
    # Every other unit contributes through its bias only: b2[j] = 1 and w3[j]
    # = 0.5 for j > 0 adds 31 * 0.5 * tanh(1).
    p[1152 + 1:1184] = 1.0
    p[1184 + 1:1216] = 0.5
    np.testing.assert_allclose(orc.mlp_forward(p, x, emulate=False), want + 31 * 0.5 * np.tanh(1.0),
                               rtol=0, atol=1e-13)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    p = random_params(rng)
    n = 96
    x = rng.standard_normal((n, 3)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    idx = rng.integers(0, n, 64).astype(np.int32)
    g, sse = orc.mlp_gradient(p, x, t, idx, emulate=False)
    assert sse > 0

    def loss(q):
        return orc.mlp_gradient(q, x, t, idx, emulate=False)[1] / idx.size

    # every parameter class, including b3
    picks = np.concatenate([rng.choice(np.arange(0, 96), 12, replace=False),
                            rng.choice(np.arange(96, 128), 6, replace=False),
                            rng.choice(np.arange(128, 1152), 24, replace=False),
                            rng.choice(np.arange(1152, 1184), 6, replace=False),
                            rng.choice(np.arange(1184, 1216), 6, replace=False), [1216]])
    gn = np.linalg.norm(g)
    for k in picks:
        qp, qm = p.copy(), p.copy()
        qp[k] = np.float32(p[k] + np.float32(2e-3))
        qm[k] = np.float32(p[k] - np.float32(2e-3))
        h = float(qp[k]) - float(qm[k])  # exact step of the float32 parameters
        fd = (loss(qp) - loss(qm)) / h
        assert abs(fd - g[k]) <= 2e-6 * gn + 1e-5 * abs(g[k]), (k, fd, g[k])


def test_bf16_emulation_changes_results_by_bf16_amounts():
    rng = np.random.default_rng(3)
    p = random_params(rng)
    x = rng.standard_normal((64, 3)).astype(np.float32)
    t = rng.standard_normal(64).astype(np.float32)
    idx = np.arange(64, dtype=np.int32)
    ye, yx = orc.mlp_forward(p, x, True), orc.mlp_forward(p, x, False)
    assert 0 < np.max(np.abs(ye - yx)) < 0.05
    ge, _ = orc.mlp_gradient(p, x, t, idx, True)
    gx, _ = orc.mlp_gradient(p, x, t, idx, False)
    rel = np.linalg.norm(ge - gx) / np.linalg.norm(gx)
    assert 1e-5 < rel < 0.03
    # With BF16-exact weights the emulated and exact forward passes differ
    # only by the rounding of the 32 first-layer activations: bound it by
    # sum_j |w3_j| * sum_k |W2_jk| * 2^-9 * |h1_k| (tanh' <= 1).
    pb = np.array([orc.bf16_round(float(v)) for v in p], dtype=np.float32)
    ye, yx = orc.mlp_forward(pb, x, True), orc.mlp_forward(pb, x, False)
    bound = np.sum(np.abs(pb[1184:1216])[:, None] * np.abs(pb[128:1152].reshape(32, 32))) * 2.0 ** -9
    assert np.all(np.abs(ye - yx) <= bound)


def test_adam_constant_gradient_moves_by_lr_sign():
    rng = np.random.default_rng(5)
    p = rng.standard_normal(NP).astype(np.float32)
    m = np.zeros(NP, np.float32)
    v = np.zeros(NP, np.float32)
    # |g| >= 1e-3 >> eps = 1e-8, so eps shifts a step by at most 1e-5 relative
    g = rng.choice([-1.0, 1.0], NP) * 10.0 ** rng.uniform(-3, 1, NP)
    lr = 1e-3
    start = p.astype(np.float64).copy()
    for step in range(1, 6):
        orc.mlp_adam(p, m, v, g, lr, step)
        np.testing.assert_allclose(p - start, -lr * step * np.sign(g), rtol=0, atol=2e-6 * step + 1e-6)


@pytest.mark.parametrize("iters", [1, 2, 4, 6])
def test_newton_iteration_count(iters):
    # the 6-iteration solve is solve_utau; fewer iterations leave a larger
    # Reichardt residual from the same cold start
    ut, ys, nu = 0.05, 2.0, 1e-4
    full = orc.solve_utau(ut, ys, nu)
    got = orc.solve_utau_iters(ut, ys, nu, 0.0, iters)
    if iters == 6:
        assert got == full
    res = abs(got * orc.reichardt_uplus(ys * got / nu) - ut) / ut
    res6 = abs(full * orc.reichardt_uplus(ys * full / nu) - ut) / ut
    assert res6 < 1e-5
    assert res >= res6 * 0.999
