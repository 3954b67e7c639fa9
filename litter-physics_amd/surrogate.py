"""Training data, training and evaluation of the Herschel-Bulkley return-map surrogate.

Data are exact returns logged by the engine (mpm_log_returns) from batches of
paste slumps with per-instance rheology (shear yield stress U[10, 300] Pa,
consistency U[0.5, 20] Pa s^n, flow index U[0.3, 1.0]) in two shapes: like the
small-slump configuration (32^3 grids, dx 2.5 or 1.5 mm, column heights
20-80 mm), and like the batched calibration sweep (grids log-uniform 16^3 to
64^3 cells over a 100 mm cube, so dx 1.6-6.3 mm, 40 x 40 mm columns 20-80 mm
tall).  The network trains on the GPU through mpm_mlp_train.

    python litter-physics_amd/surrogate.py train --samples 20000000 --epochs 6

writes the f32 master parameters to surrogate_weights.npy beside this file and a
JSON report of the held-out evaluation.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "surrogate_weights.npy")
# Loss weight per output: six stress components, then the overstress fraction
# (the output the deployed return uses).
LOSS_WEIGHTS = np.array([1, 1, 1, 1, 1, 1, 20], np.float32)


def engine():
    """The engine package (works whether or not it was imported under its module name)."""
    mod = sys.modules.get("litter_physics_amd")
    if mod is not None:
        return mod
    import importlib.util
    spec = importlib.util.spec_from_file_location("litter_physics_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["litter_physics_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def slump_specs(E, count, seed, cells=32, yield_range=(10.0, 300.0), k_range=(0.5, 20.0), n_range=(0.3, 1.0),
                height_range=(0.02, 0.08), spacings=(0.0025, 0.0015)):
    """Paste slumps with random rheology (shear units) and column shape from `seed`."""
    rng = np.random.default_rng(seed)
    bulk, shear = 2e5, 5e4
    young = 9.0 * bulk * shear / (3.0 * bulk + shear)
    poisson = (3.0 * bulk - 2.0 * shear) / (2.0 * (3.0 * bulk + shear))
    face = [0.0, 0.0, 0.0, 0.0, math.inf, 0.0]
    specs = []
    for i in range(count):
        dx = float(spacings[i % len(spacings)])
        domain = [cells * dx] * 3
        mat = E.paste(1200.0, young, poisson, rng.uniform(*yield_range), rng.uniform(*k_range),
                      rng.uniform(*n_range), shear_units=True)
        cfg = E.config(wall_friction=0.0, cfl=0.4, max_dt=1e-4, min_dt=1e-12, face_friction=face)
        height = min(rng.uniform(*height_range), 0.9 * domain[2])
        width = min(rng.uniform(0.02, 0.05), 0.7 * domain[0])
        st = E.place_particles("slump", mat, dx, domain, [width, width, height], seed=int(rng.integers(1 << 31)),
                               jitter=0.5)
        specs.append({"material": mat, "cfg": cfg, "domain": domain, "spacing": dx, "state": st})
    return specs


def sweep_like_specs(E, count, seed, yield_range=(10.0, 300.0), k_range=(0.5, 20.0), n_range=(0.3, 1.0)):
    """Paste slumps shaped like the batched calibration sweep: grids log-uniform
    16^3..64^3 cells over a 100 mm cube (dx = 100 mm / cells), 40 x 40 mm
    columns 20-80 mm tall, random rheology (shear units) from `seed`."""
    rng = np.random.default_rng(seed)
    bulk, shear = 2e5, 5e4
    young = 9.0 * bulk * shear / (3.0 * bulk + shear)
    poisson = (3.0 * bulk - 2.0 * shear) / (2.0 * (3.0 * bulk + shear))
    face = [0.0, 0.0, 0.0, 0.0, math.inf, 0.0]
    specs = []
    for _ in range(count):
        cells = int(round(math.exp(rng.uniform(math.log(16), math.log(64)))))
        dx = 0.1 / cells
        domain = [cells * dx] * 3
        mat = E.paste(1200.0, young, poisson, rng.uniform(*yield_range), rng.uniform(*k_range),
                      rng.uniform(*n_range), shear_units=True)
        cfg = E.config(wall_friction=0.0, cfl=0.4, max_dt=1e-4, min_dt=1e-12, face_friction=face)
        st = E.place_particles("slump", mat, dx, domain, [0.04, 0.04, rng.uniform(0.02, 0.08)],
                               seed=int(rng.integers(1 << 31)), jitter=0.5)
        specs.append({"material": mat, "cfg": cfg, "domain": domain, "spacing": dx, "state": st})
    return specs


def collect(samples, seed=3, instances=64, sweep_instances=0, windows=40, gap=48, log=print):
    """Log exact returns: `windows` single logged steps, the first ones after 1, 2,
    4, ... steps (the loading transient) and then `gap` steps apart, randomly
    subsampled to `samples` in total.  Returns (inputs, targets, aux)."""
    E = engine()
    specs = slump_specs(E, instances, seed) + sweep_like_specs(E, sweep_instances, seed + 1000)
    total = sum(len(s["state"]) for s in specs)
    batch = E.Batch(specs)
    rng = np.random.default_rng(seed + 1)
    per = max(samples // windows, 1)
    xs, ts, auxs = [], [], []
    got = 0
    for w in range(windows):
        if got >= samples:
            break
        batch.advance(1e9, max_steps=min(1 << w, gap))
        batch.log_returns(total)
        batch.advance(1e9, max_steps=1)
        x, t, a = batch.take_returns(total)
        batch.log_returns(0)
        want = min(per if w < windows - 1 else samples - got, samples - got)
        if len(x) > want:
            keep = np.sort(rng.choice(len(x), want, replace=False))
            x, t, a = x[keep], t[keep], a[keep]
        xs.append(x)
        ts.append(t)
        auxs.append(a)
        got += len(x)
        log(f"window {w}: {len(x)} samples, {got} total")
    return np.concatenate(xs), np.concatenate(ts), np.concatenate(auxs)


def residual(aux, y, corrections=0):
    """|f(q)| / max(q_tr, 1) of the return equation after `corrections` Newton
    steps from the overstress fraction y (f64, the device's safeguard)."""
    qtr, ys, s3, K, n = (aux[:, i].astype(np.float64) for i in range(5))
    avail = qtr - ys
    d = avail * np.clip(np.asarray(y, np.float64), 0.0, 1.0)
    inv_n = 1.0 / n
    for _ in range(corrections):
        over = d / K
        f = d - avail + s3 * over ** inv_n
        with np.errstate(divide="ignore", invalid="ignore"):
            deriv = np.where(over > 0, 1.0 + s3 * inv_n * over ** (inv_n - 1.0) / K, np.inf)
            nd = d - f / deriv
        d = np.where(np.isfinite(nd) & (nd > 0) & (nd < avail), nd, d)
    f = d - avail + s3 * (d / K) ** inv_n
    return np.abs(f) / np.maximum(qtr, 1.0), ys + d


def evaluate(net, x, t, aux, tol=1e-4):
    """Held-out accuracy of the network, with and without one Newton correction,
    against the exact returns; the baseline is the exact solver's own starting
    point (half the overstress) given the same single correction."""
    y = net.forward(x)
    q_exact = aux[:, 5].astype(np.float64)
    out = {"samples": int(len(x))}
    half = np.full(len(x), 0.5)
    for name, yy, k in (("network", y[:, 6], 0), ("network+1 Newton", y[:, 6], 1), ("network+2 Newton", y[:, 6], 2),
                        ("no network, 1 Newton from q_tr/2 overstress", half, 1),
                        ("no network, 2 Newton from q_tr/2 overstress", half, 2)):
        res, q = residual(aux, yy, k)
        err = np.abs(q - q_exact) / q_exact
        out[name] = {"fallback_rate": float(np.mean(res > tol)), "residual_median": float(np.median(res)),
                     "q_rel_err_p99": float(np.quantile(err, 0.99)), "q_rel_err_max": float(err.max())}
    out["stress_rmse_outputs_0_5"] = float(np.sqrt(np.mean((y[:, :6] - t[:, :6]) ** 2)))
    return out


def load_network(path=WEIGHTS):
    E = engine()
    return E.Network(np.load(path).astype(np.float32))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--samples", type=int, default=20_000_000)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--sweep-instances", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    E = engine()

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    x, t, aux = collect(args.samples, seed=args.seed, instances=args.instances,
                        sweep_instances=args.sweep_instances, log=log)
    log(f"collected {len(x)} samples in {time.perf_counter() - t0:.1f} s")
    rng = np.random.default_rng(args.seed + 2)
    order = rng.permutation(len(x))
    hold = order[: max(len(x) // 50, 1)]
    train = order[len(hold):]
    net = E.Network(seed=args.seed)
    t1 = time.perf_counter()
    losses = []
    for ep in range(args.epochs):
        # One epoch per call keeps a progress line per epoch; the learning-rate
        # schedule restarts, decaying over the remaining epochs.
        lr = args.lr * 0.5 * (1.0 + math.cos(math.pi * ep / args.epochs))
        losses += list(net.train(x[train], t[train], LOSS_WEIGHTS, epochs=1, batch=args.batch, lr=max(lr, 1e-5),
                                 seed=args.seed + ep))
        log(f"epoch {ep}: loss {losses[-1]:.3e} ({time.perf_counter() - t1:.1f} s)")
    report = {"train_samples": int(len(train)), "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
              "loss_per_epoch": losses, "train_seconds": time.perf_counter() - t1,
              "heldout": evaluate(net, x[hold], t[hold], aux[hold])}
    np.save(args.out, net.params())
    text = json.dumps(report, indent=1)
    print(text, flush=True)
    if args.report:
        with open(args.report, "w") as f:
            f.write(text)


if __name__ == "__main__":
    main()
