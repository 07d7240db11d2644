"""The CPU oracle of the learned preconditioner (oracle/mlp.py), pinned by mathematics.

* its analytic gradient equals central finite differences of the loss (parameters chosen
  exactly representable in bfloat16, so the rounding is the identity);
* planted mistakes in the backward pass (a dropped ReLU mask, a transposed W2, a missing
  factor 2) fail that comparison;
* a network with a zero last layer is the diagonal preconditioner exactly;
* the recorder's importance sampling is unbiased: the weighted sample sums of u^2, u v, v^2
  estimate the full-vector sums, and a huge target keeps every determinant with weight 1.
"""
import numpy as np
import pytest

import mlp as M
import oracle as O
from sbd_amd import synth


def samples(n, seed=0, L=12):
    rng = np.random.default_rng(seed)
    s = np.zeros(n, dtype=M.SAMPLE_DTYPE)
    for f in ("alpha", "beta"):
        bits = np.zeros(n, np.uint64)
        for i in range(n):
            for p in rng.choice(L, 4, replace=False):
                bits[i] |= np.uint64(1) << np.uint64(p)
        s[f] = bits
    s["xh"] = rng.uniform(0, 4, n)
    s["xr"] = rng.uniform(0, 3, n)
    s["u"] = rng.standard_normal(n)
    s["v"] = s["u"] * rng.uniform(0.5, 1.5, n)
    return s


def bf16_params(seed):
    p = M.init_params(seed, zero_last=False)
    p[M.OFF_B1:M.OFF_W2] = np.random.default_rng(seed + 1).uniform(-0.2, 0.2, M.H)
    return M.bf16_round(p).astype(np.float32)


def fd_check(grad_fn, p, s, idx, h=1e-6):
    _, g = grad_fn(p, s)
    p64 = p.astype(np.float64)
    err = []
    for q in idx:
        a = p64.copy(); a[q] += h
        b = p64.copy(); b[q] -= h
        fd = (M.loss_only(a, s) - M.loss_only(b, s)) / (2 * h)
        err.append(abs(fd - g[q]) / max(1e-6, abs(fd) + abs(g[q])))
    return max(err)


def check_indices(seed=3):
    rng = np.random.default_rng(seed)
    groups = [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
              (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NPARAM)]
    idx = []
    for a, b in groups:
        idx += list(rng.choice(np.arange(a, b), min(24, b - a), replace=False))
    # the xh and xr rows and bits that are set in some sample
    idx += [M.OFF_W1 + 128 * M.H + 5, M.OFF_W1 + 129 * M.H + 9, M.OFF_W1 + 3 * M.H + 7]
    return idx


def test_gradient_matches_finite_differences():
    s = samples(96)
    p = bf16_params(7)
    assert np.array_equal(M.bf16_round(p), p.astype(np.float64))
    assert fd_check(M.loss_grad, p, s, check_indices()) < 1e-5


def planted(kind):
    def grad(params, s):
        w = M.unpack(params)
        m, (X, z1, a1, z2, a2) = M.forward(params, s)
        u = s["u"].astype(np.float64)
        e = u * (1 + m) - s["v"]
        n = s.size
        dm = (1.0 if kind == "factor" else 2.0) * e * u / n
        g = np.zeros(M.NPARAM)
        g[M.OFF_B3] = dm.sum()
        g[M.OFF_W3:M.OFF_B3] = a2.T @ dm
        dz2 = np.outer(dm, w["w3"]) * (z2 > 0)
        g[M.OFF_B2:M.OFF_W3] = dz2.sum(0)
        g[M.OFF_W2:M.OFF_B2] = (a1.T @ dz2).ravel()
        W2 = w["W2"] if kind == "transpose" else w["W2"].T
        dz1 = dz2 @ W2
        if kind != "mask":
            dz1 = dz1 * (z1 > 0)
        g[M.OFF_B1:M.OFF_W2] = dz1.sum(0)
        g[M.OFF_W1:M.OFF_B1] = (X.T @ dz1).ravel()
        return float(np.sum(e * e) / n), g
    return grad


@pytest.mark.parametrize("kind", ["mask", "transpose", "factor"])
def test_planted_backward_mistakes_are_caught(kind):
    s = samples(96)
    p = bf16_params(7)
    assert fd_check(planted(kind), p, s, check_indices()) > 1e-3


def test_zero_last_layer_is_the_diagonal_preconditioner():
    inst = synth.make_custom(8, 3, 2, 6, None, None, seed_ints=5)
    so = O.System.from_instance(inst)
    hii = so.hdiag()
    r = np.random.default_rng(1).standard_normal(so.ndet)
    theta = hii.min() - 0.3
    p = M.init_params(0, zero_last=True)
    t = M.precondition(p, inst.alpha, inst.beta, hii, theta, r)
    assert np.array_equal(t, r / (theta - hii))


def test_recorder_weights_are_unbiased():
    inst = synth.make_custom(8, 3, 2, 6, None, None, seed_ints=5)
    so = O.System.from_instance(inst)
    hii = so.hdiag()
    ref = so.davidson(max_it=100, num_block=10, eps=1e-10)
    xs = ref["W"]
    x = xs + 2e-3 * np.random.default_rng(2).standard_normal(so.ndet)
    sigma = lambda v: so.mult(v, hii)
    T = M.record_terms(sigma, hii, x, xs)
    assert abs(np.linalg.norm(T["v"]) - 1.0) < 1e-12
    assert T["theta"] < hii.min()
    # every determinant kept with weight 1 when the target is large enough
    s, keep, theta, p = M.record(sigma, hii, inst.alpha, inst.beta, x, xs, 1e12, 0)
    assert keep.size == so.ndet and np.all(p == 1.0)
    # sampling: the mean over seeds of the weighted sums matches the full sums
    full = np.array([T["u"] @ T["u"], T["u"] @ T["v"], T["v"] @ T["v"]])
    est = []
    for seed in range(40):
        s, keep, theta, p = M.record(sigma, hii, inst.alpha, inst.beta, x, xs, 200, seed)
        u = s["u"].astype(np.float64)
        v = s["v"].astype(np.float64)
        est.append([u @ u, u @ v, v @ v])
    est = np.mean(est, 0)
    assert np.max(np.abs(est - full) / np.abs(full)) < 0.1, (est, full)
