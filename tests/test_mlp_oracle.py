"""The CPU f64 MLP that checks the device network, pinned without the device.

* its backpropagation against central finite differences of its own loss;
* its bf16 rounding against hand-worked bit patterns;
* the dimensionless form of the return it learns (y + beta y^(1/n) = 1) against
  the oracle's Herschel-Bulkley solver (constitutive.rs:280-362);
* the encoding constants against the formulas they come from.
"""
import math

import numpy as np

import mlp as M
import oracle as O


def test_bf16_rounding_bit_patterns():
    one = 1.0
    ulp = 2.0 ** -7  # bf16 spacing at 1
    vals = np.array([one, one + ulp, one + ulp / 2, one + 3 * ulp / 2, -2.5, 3.0e-3, 65504.0])
    got = M.bf16(vals)
    # ties go to even: 1 + ulp/2 -> 1, 1 + 3 ulp/2 -> 1 + 2 ulp
    # 3e-3 = 1.536 * 2^-9 and 0.536 * 128 = 68.6 rounds to 69; 65504 rounds up to 2^16
    want = np.array([one, one + ulp, one, one + 2 * ulp, -2.5, (1 + 69 / 128) * 2.0 ** -9, 65536.0])
    assert np.array_equal(got, want), got


def test_f64_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init(1) + rng.normal(0, 0.05, M.N_PARAMS) * (np.arange(M.N_PARAMS) % 7 == 0)
    x = rng.normal(0, 1, (5, 12))
    t = rng.normal(0, 1, (5, 7))
    w = np.array([1, 2, 1, 1, 3, 1, 10.0])
    loss, g = M.loss_grad(p, x, t, w, emulate=False)
    idx = rng.choice(M.N_PARAMS, 60, replace=False)
    # Every block gets checked, including the last biases.
    idx = np.concatenate([idx, [0, 12 * 64, 12 * 64 + 64, M.N_PARAMS - 1, M.N_PARAMS - 8]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, x, t, w, emulate=False)[0] - M.loss_grad(pm, x, t, w, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    assert loss > 0


def test_emulated_network_is_close_to_f64():
    # bf16 rounding of inputs, weights and activations moves the outputs by about 1e-2.
    rng = np.random.default_rng(2)
    p = M.init(3)
    x = rng.normal(0, 1, (64, 12))
    y64 = M.forward(p, x, emulate=False)[0]
    y16 = M.forward(p, x, emulate=True)[0]
    rel = np.abs(y16 - y64).max() / np.abs(y64).max()
    assert 1e-4 < rel < 5e-2, rel


def test_dimensionless_return_matches_oracle_solver():
    # q = sigma_y + (q_tr - sigma_y) y with y + beta y^(1/n) = 1,
    # beta = (3 G dt / a) (a / K)^(1/n), a = q_tr - sigma_y.
    rng = np.random.default_rng(4)
    G, dt = 5e4, 6.7e-5
    for _ in range(200):
        ys = rng.uniform(17.0, 520.0)
        qtr = ys * (1.0 + 10 ** rng.uniform(-4, 1.5))
        K = rng.uniform(0.5, 20.0) * 3 ** ((1 + 0.5) / 2)
        n = rng.uniform(0.3, 1.0)
        rc, q, inc, it = O.hb_return(qtr, ys, K, n, G, dt)
        assert rc == 0
        a = qtr - ys
        beta = (3 * G * dt / a) * (a / K) ** (1.0 / n)
        y = M.hb_overstress_fraction(np.array([beta]), np.array([n]))[0]
        assert abs(ys + a * y - q) <= 1e-9 * qtr, (qtr, ys, K, n)
        assert abs(inc - a * (1 - y) / (3 * G)) <= 1e-9 * qtr / (3 * G)


def test_input_encoding_determines_the_return():
    # From inputs 6 (r), 10 (K / 3G dt) and 11 (n) alone: beta = (1/r) (r / K')^(1/n).
    rng = np.random.default_rng(5)
    for _ in range(50):
        s3 = rng.uniform(3.0, 15.0)
        a = 10 ** rng.uniform(-3, 2.5)
        K = rng.uniform(0.5, 40.0)
        n = rng.uniform(0.3, 1.0)
        f6 = math.log10(a / s3) / 4
        f10 = math.log10(K / s3) / 4
        r, kp = 10 ** (4 * f6), 10 ** (4 * f10)
        beta_inputs = (1 / r) * (r / kp) ** (1 / n)
        beta = (s3 / a) * (a / K) ** (1 / n)
        assert abs(beta_inputs - beta) <= 1e-9 * beta
