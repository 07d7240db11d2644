"""CPU oracle of the learned Davidson preconditioner (float64 numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the engine's MLP kernels (forward, loss,
gradient, training-pair recording), never by the product package.  The reference has no
learned preconditioner; this restates the engine's own definition, and it is pinned by central
finite differences of its loss (tests/test_oracle_mlp.py), by the closed form of a network
whose last layer is zero (the diagonal preconditioner exactly), and by planted mistakes.

Network (parameters: one float32 vector of NPARAM entries, used rounded to bfloat16):
    z1 = b1 + sum_{p in alpha} W1[p] + sum_{p in beta} W1[64 + p] + xh W1[128] + xr W1[129]
    a1 = relu(z1);  z2 = b2 + a1 W2;  a2 = relu(z2);  m = b3 + a2 . w3
Preconditioned correction of residual component r_i at Ritz value theta:
    t_i = r_i / (theta - H_ii) * (1 + m_i),  xh = asinh(H_ii - theta),  xr = asinh(|r_i| / rms(r))
Training pair of determinant i (u, v already carry the importance weight and the snapshot
normalisation): loss = (1 / n) sum_s (u_s (1 + m_s) - v_s)^2.
"""
from __future__ import annotations

import numpy as np

H = 64
NIN = 130
OFF_W1 = 0
OFF_B1 = OFF_W1 + NIN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1

SAMPLE_DTYPE = np.dtype([("alpha", "<u8"), ("beta", "<u8"), ("xh", "<f4"), ("xr", "<f4"),
                         ("u", "<f4"), ("v", "<f4")])


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float64."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    lsb = (b >> 16) & 1
    b = ((b + 0x7FFF + lsb) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = bf16_round(params)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(NIN, H), b1=p[OFF_B1:OFF_W2],
                W2=p[OFF_W2:OFF_B2].reshape(H, H), b2=p[OFF_B2:OFF_W3],
                w3=p[OFF_W3:OFF_B3], b3=p[OFF_B3])


def inputs(samples):
    """Dense input matrix (n, 130): alpha bits, beta bits, xh, xr (features rounded to
    bfloat16, like the parameters)."""
    n = samples.size
    X = np.zeros((n, NIN))
    for p in range(64):
        X[:, p] = (samples["alpha"] >> np.uint64(p)) & np.uint64(1)
        X[:, 64 + p] = (samples["beta"] >> np.uint64(p)) & np.uint64(1)
    X[:, 128] = bf16_round(samples["xh"])
    X[:, 129] = bf16_round(samples["xr"])
    return X


def forward(params, samples):
    w = unpack(params)
    X = inputs(samples)
    z1 = w["b1"] + X @ w["W1"]
    a1 = np.maximum(z1, 0.0)
    z2 = w["b2"] + a1 @ w["W2"]
    a2 = np.maximum(z2, 0.0)
    m = w["b3"] + a2 @ w["w3"]
    return m, (X, z1, a1, z2, a2)


def loss_grad(params, samples):
    """Loss and its gradient with respect to the (bfloat16-rounded) parameters."""
    w = unpack(params)
    m, (X, z1, a1, z2, a2) = forward(params, samples)
    u = samples["u"].astype(np.float64)
    v = samples["v"].astype(np.float64)
    n = max(1, samples.size)
    e = u * (1.0 + m) - v
    loss = float(np.sum(e * e) / n)
    dm = 2.0 * e * u / n
    g = np.zeros(NPARAM)
    g[OFF_B3] = dm.sum()
    g[OFF_W3:OFF_B3] = a2.T @ dm
    dz2 = np.outer(dm, w["w3"]) * (z2 > 0)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (a1.T @ dz2).ravel()
    dz1 = (dz2 @ w["W2"].T) * (z1 > 0)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (X.T @ dz1).ravel()
    return loss, g


def loss_only(params64, samples):
    """Loss with parameters used as given (no bfloat16 rounding): for finite differences."""
    p = np.asarray(params64, np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(NIN, H)
    X = inputs(samples)
    a1 = np.maximum(p[OFF_B1:OFF_W2] + X @ W1, 0.0)
    a2 = np.maximum(p[OFF_B2:OFF_W3] + a1 @ p[OFF_W2:OFF_B2].reshape(H, H), 0.0)
    m = p[OFF_B3] + a2 @ p[OFF_W3:OFF_B3]
    e = samples["u"] * (1.0 + m) - samples["v"]
    return float(np.sum(e * e) / max(1, samples.size))


def init_params(seed=0, zero_last=True):
    """He-style initial weights; a zero last layer makes the network the diagonal preconditioner."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    p[OFF_W1:OFF_B1] = rng.standard_normal(NIN * H) * np.sqrt(2.0 / 24.0)
    p[OFF_W2:OFF_B2] = rng.standard_normal(H * H) * np.sqrt(2.0 / H)
    if not zero_last:
        p[OFF_W3:OFF_B3] = rng.standard_normal(H) * np.sqrt(1.0 / H)
    return p


def precondition(params, alpha, beta, hii, theta, r):
    """t = r / (theta - H_ii) * (1 + m) over a whole vector, det = ia * n_beta + ib (in either
    sign convention of the determinant basis: the features use |r_i|, so t flips with r); the
    regulariser of the diagonal preconditioner is the reference's (chemistry/tpb/davidson.h:
    855-862)."""
    n = hii.size
    nb = beta.size
    ia = np.arange(n) // nb
    ib = np.arange(n) % nb
    rms = np.sqrt(np.dot(r, r) / n)
    s = np.zeros(n, dtype=SAMPLE_DTYPE)
    s["alpha"] = alpha[ia]
    s["beta"] = beta[ib]
    s["xh"] = np.arcsinh(hii - theta).astype(np.float32)
    s["xr"] = np.arcsinh(np.abs(r) / rms if rms > 0 else 0.0 * r).astype(np.float32)
    m, _ = forward(params, s)
    den = theta - hii
    den = np.where(np.abs(den) > 1e-12, den, den - 1e-12)
    return r / den * (1.0 + m)


def hash01(i, seed):
    """Counter-based uniform [0, 1) of determinant index i: the splitmix64 finaliser (Steele,
    Lea, Flood 2014) of seed ^ ((i + 1) * 0x9E3779B97F4A7C15), top 53 bits."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) ^ ((np.asarray(i, np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def record_terms(sigma, hii, x, xs):
    """Per-determinant quantities of a training snapshot (caller's basis, float64): Rayleigh
    quotient theta, residual r of the normalised iterate, exact correction d = x*^ - (x^.x*^) x^
    (x* sign-aligned with x), diagonal correction t, and u, v, q = u^2 + v^2 as recorded."""
    xx = x @ x
    hx = sigma(x)
    theta = (x @ hx) / xx
    sx = 1.0 / np.sqrt(xx)
    xxs = x @ xs
    sxs = (-1.0 if xxs < 0 else 1.0) / np.sqrt(xs @ xs)
    proj = abs(xxs) / np.sqrt(xx * (xs @ xs))
    xn = x * sx
    r = hx * sx - theta * xn
    d = xs * sxs - proj * xn
    den = theta - hii
    den = np.where(np.abs(den) > 1e-12, den, den - 1e-12)
    t = r / den
    c = (t @ d) / (d @ d)
    vnorm = abs(c) * np.sqrt(d @ d)
    u = t / vnorm
    v = c * d / vnorm
    return dict(theta=theta, r=r, d=d, t=t, u=u, v=v, q=u * u + v * v,
                rms=np.sqrt((r @ r) / x.size))


def record(sigma, hii, alpha, beta, x, xs, target, seed):
    """The samples kept by the engine's recorder, sorted by determinant index."""
    T = record_terms(sigma, hii, x, xs)
    n = x.size
    p = np.minimum(1.0, target * T["q"] / T["q"].sum())
    keep = np.nonzero(hash01(np.arange(n), seed) < p)[0]
    nb = beta.size
    s = np.zeros(keep.size, dtype=SAMPLE_DTYPE)
    s["alpha"] = alpha[keep // nb]
    s["beta"] = beta[keep % nb]
    s["xh"] = np.arcsinh(hii[keep] - T["theta"]).astype(np.float32)
    s["xr"] = np.arcsinh(np.abs(T["r"][keep]) / T["rms"]).astype(np.float32)
    w = 1.0 / np.sqrt(p[keep])
    s["u"] = (T["u"][keep] * w).astype(np.float32)
    s["v"] = (T["v"][keep] * w).astype(np.float32)
    return s, keep, T["theta"], p
