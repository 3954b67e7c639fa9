"""The return-map surrogate on the device: MFMA forward and backward against the
CPU f64 network, training on logged exact returns, and deployment in the
substep with the exact return as fallback, against the f64 oracle.

Tolerances: forward outputs within 1e-3 of the emulated f64 network (normwise),
gradients within 5e-3 per parameter block; the substep with the surrogate
within the engine's contract (positions and velocities 1e-4, stress 1e-3).
"""
import os

import numpy as np
import pytest

import cases
import mlp as M
from test_gpu_parity import build_oracle, compare_states, engine_from_oracle

pytestmark = pytest.mark.gpu

BLOCKS = ["W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4"]


def blocks(v):
    out, o = {}, 0
    for (a, b), (wn, bn) in zip(M.SIZES, zip(BLOCKS[::2], BLOCKS[1::2])):
        out[wn] = v[o:o + a * b]
        o += a * b
        out[bn] = v[o:o + b]
        o += b
    return out


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = M.init(seed)
    # Non-zero biases so that every block of the gradient is exercised.
    return (p + (np.abs(p) == 0) * rng.normal(0, 0.1, p.size)).astype(np.float32)


def test_forward_matches_cpu_network(engine):
    p = random_params(11)
    net = engine.Network(p)
    rng = np.random.default_rng(12)
    x = rng.normal(0, 1, (1000, 12)).astype(np.float32)  # not a multiple of 16
    y = net.forward(x)
    ref = M.forward(p.astype(np.float64), x.astype(np.float64), emulate=True)[0]
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print("forward err", err, flush=True)
    assert err < 1e-3
    assert np.array_equal(net.params(), p)


def test_gradient_matches_cpu_f64_network(engine):
    p = random_params(21)
    net = engine.Network(p)
    rng = np.random.default_rng(22)
    x = rng.normal(0, 1, (777, 12)).astype(np.float32)
    t = rng.normal(0, 1, (777, 7)).astype(np.float32)
    w = np.array([1, 1, 1, 1, 1, 1, 20], np.float32)
    loss, g = net.loss_grad(x, t, w)
    rloss, rg = M.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), w, emulate=True)
    assert abs(loss - rloss) <= 1e-4 * rloss
    gb, rb = blocks(g), blocks(rg)
    for name in BLOCKS:
        e = np.linalg.norm(gb[name] - rb[name]) / np.linalg.norm(rb[name])
        print(f"grad {name}: rel err {e:.2e}", flush=True)
        assert e < 5e-3, name


def small_dataset(engine, samples=300_000):
    import litter_physics_amd.surrogate as S
    return S.collect(samples, seed=7, instances=8, windows=16, gap=40, log=lambda m: None)


@pytest.fixture(scope="module")
def trained(engine):
    """A network trained here on logged exact returns (kept small so the test is quick)."""
    import litter_physics_amd.surrogate as S
    x, t, aux = small_dataset(engine)
    assert len(x) > 100_000
    rng = np.random.default_rng(0)
    order = rng.permutation(len(x))
    hold, tr = order[:20_000], order[20_000:]
    net = engine.Network(seed=5)
    losses = net.train(x[tr], t[tr], S.LOSS_WEIGHTS, epochs=6, batch=4096, lr=3e-3, seed=1)
    return net, losses, (x[hold], t[hold], aux[hold])


def test_training_on_the_device_reduces_the_loss(engine, trained):
    net, losses, _ = trained
    print("loss per epoch", losses, flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.05 * losses[0]


def test_network_gains_over_the_plain_starting_point(engine, trained):
    # With one Newton correction after the prediction, far more returns meet the
    # residual gate than one correction from the exact solver's own start.
    import litter_physics_amd.surrogate as S
    net, _, (x, t, aux) = trained
    rep = S.evaluate(net, x, t, aux)
    print(rep, flush=True)
    with_net = rep["network+1 Newton"]["fallback_rate"]
    without = rep["no network, 1 Newton from q_tr/2 overstress"]["fallback_rate"]
    assert with_net < 0.5 * without
    assert rep["network"]["q_rel_err_p99"] < 0.05


def weights_for_deployment(engine, trained):
    import litter_physics_amd.surrogate as S
    if os.path.exists(S.WEIGHTS):
        return S.load_network(), True
    return trained[0], False


def test_substep_with_surrogate_matches_oracle(engine, trained):
    net, committed = weights_for_deployment(engine, trained)
    spec = cases.paste_column_32()
    osim = build_oracle(spec)
    gsim = engine_from_oracle(engine, osim)
    gsim.set_surrogate(net, tolerance=1e-4, polish=2)
    dt = min(osim.stability_limit(), spec.config.max_dt)
    for k in range(100):
        assert osim.try_step(dt) == 0
        assert gsim.try_step(dt).accepted == 1, k
    considered, accepted = gsim.surrogate_stats()
    fallback = 1.0 - accepted / max(considered, 1)
    print(f"surrogate: {considered} yielded returns, fallback rate {fallback:.4f} (committed weights: {committed})",
          flush=True)
    assert considered > 0
    compare_states(gsim, osim)
    lo, lg = osim.ledger(), gsim.ledger()
    assert abs(lg.plastic_dissipation - lo.plastic_dissipation) / lo.plastic_dissipation < 1e-3
    if committed:
        assert fallback < 0.01


def test_exact_path_unchanged_without_surrogate(engine, trained):
    # Switching the surrogate off restores the exact return (no offers counted).
    net, _ = weights_for_deployment(engine, trained)
    spec = cases.paste_column_32()
    osim = build_oracle(spec)
    gsim = engine_from_oracle(engine, osim)
    gsim.set_surrogate(net)
    gsim.set_surrogate(None)
    dt = min(osim.stability_limit(), spec.config.max_dt)
    for _ in range(20):
        assert osim.try_step(dt) == 0
        assert gsim.try_step(dt).accepted == 1
    assert gsim.surrogate_stats() == (0, 0)
    compare_states(gsim, osim)
