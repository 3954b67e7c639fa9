"""CPU f64 reference of the return-map MLP (test infrastructure only).

The network is 12 -> 64 -> 64 -> 64 -> 7 with SiLU, parameters in the master
layout W1[12][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][7] b4[7]
(W[k][n] maps input k to output n).  The loss is sum_i sum_u w_u (y_iu - t_iu)^2 / n.

`emulate=True` rounds to bf16 exactly where the device kernels do: inputs,
weights, SiLU outputs and the output/hidden deltas before they enter a matrix
product; everything else (pre-activations, biases, accumulation, bias
gradients) stays f64.  `emulate=False` is the plain f64 network, whose gradient
the tests pin by central finite differences.

The network has no counterpart in the reference; the return map it stands in
for is solve_herschel_bulkley_return (core/src/mpm/constitutive.rs:280-362).
"""
import numpy as np

SIZES = [(12, 64), (64, 64), (64, 64), (64, 7)]
N_PARAMS = sum(a * b + b for a, b in SIZES)


def bf16(x):
    """Round to the nearest bf16 (ties to even), returned as f64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = [], 0
    for a, b in SIZES:
        w = p[o:o + a * b].reshape(a, b)
        o += a * b
        out.append((w, p[o:o + b]))
        o += b
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in layers])


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, emulate=True):
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(p)
    h = r(x)
    zs, hs = [], [h]
    for i, (w, b) in enumerate(layers):
        z = h @ r(w) + b
        if i == 3:
            return z, zs, hs
        zs.append(z)
        h = r(silu(z))
        hs.append(h)


def loss_grad(p, x, t, weights=None, emulate=True):
    """Loss and gradient in the master layout."""
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    w_u = np.ones(7) if weights is None else np.asarray(weights, dtype=np.float64)
    n = x.shape[0]
    layers = unpack(p)
    y, zs, hs = forward(p, x, emulate)
    err = y - np.asarray(t, dtype=np.float64)
    loss = float((w_u * err * err).sum() / n)
    d = 2.0 * w_u * err / n
    grads = [None] * 4
    for i in range(3, -1, -1):
        dr = r(d)
        grads[i] = (hs[i].T @ dr, d.sum(0))
        if i > 0:
            d = (dr @ r(layers[i][0]).T) * silu_grad(zs[i - 1])
    return loss, pack(grads)


def init(seed=0):
    """Xavier-uniform weights from numpy's generator, zero biases (tests only)."""
    rng = np.random.default_rng(seed)
    layers = []
    for a, b in SIZES:
        lim = np.sqrt(6.0 / (a + b))
        layers.append((rng.uniform(-lim, lim, (a, b)), np.zeros(b)))
    return pack(layers)


def hb_overstress_fraction(beta, n_exp, iters=200):
    """y in (0, 1) with y + beta y^(1/n) = 1 (the dimensionless return), by bisection in f64."""
    beta = np.asarray(beta, dtype=np.float64)
    inv = 1.0 / np.asarray(n_exp, dtype=np.float64)
    lo = np.zeros_like(beta)
    hi = np.ones_like(beta)
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        f = mid + beta * mid ** inv - 1.0
        hi = np.where(f > 0, mid, hi)
        lo = np.where(f > 0, lo, mid)
    return 0.5 * (lo + hi)
