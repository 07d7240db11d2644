"""The learned preconditioner's GPU kernels against the CPU oracle (oracle/mlp.py), and what
the trained network does inside the solver.

Tolerances: the engine evaluates the network in FP32 on the matrix cores, the oracle in
float64 with the same bfloat16-rounded parameters and features, so outputs agree to ~1e-6
relative; gradients are compared block by block in relative L2 norm (1e-4).
"""
import numpy as np
import pytest

import mlp as M
import oracle as O
from sbd_amd import mlp_train, synth
from sbd_amd._lib import Engine, SbdxError, mlp_forward, mlp_loss_grad, mlp_train as train

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def random_samples(n, L=36, ne=8, seed=0):
    rng = np.random.default_rng(seed)
    s = np.zeros(n, dtype=M.SAMPLE_DTYPE)
    for f in ("alpha", "beta"):
        occ = np.argsort(rng.random((n, L)), axis=1)[:, :ne]
        s[f] = np.bitwise_or.reduce(np.left_shift(np.uint64(1), occ.astype(np.uint64)), axis=1)
    s["xh"] = rng.uniform(0.0, 4.0, n)
    s["xr"] = rng.uniform(0.0, 3.0, n)
    s["u"] = rng.standard_normal(n)
    s["v"] = s["u"] * rng.uniform(0.3, 1.7, n)
    return s


def params(seed):
    p = M.init_params(seed, zero_last=False)
    p[M.OFF_B1:M.OFF_W2] = np.random.default_rng(seed).uniform(-0.3, 0.3, M.H)
    p[M.OFF_B2:M.OFF_W3] = np.random.default_rng(seed + 1).uniform(-0.3, 0.3, M.H)
    p[M.OFF_B3] = 0.05
    return p


GROUPS = {"W1": (M.OFF_W1, M.OFF_B1), "b1": (M.OFF_B1, M.OFF_W2), "W2": (M.OFF_W2, M.OFF_B2),
          "b2": (M.OFF_B2, M.OFF_W3), "w3": (M.OFF_W3, M.OFF_B3), "b3": (M.OFF_B3, M.NPARAM)}


@pytest.mark.parametrize("n", [1, 64, 1000])
def test_forward_matches_oracle(n):
    print(f"[mlp forward n={n}]", flush=True)
    s = random_samples(n, seed=n)
    p = params(1)
    got = mlp_forward(s, p).astype(np.float64)
    ref, _ = M.forward(p, s)
    assert np.max(np.abs(got - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [64, 777, 5000])
def test_gradient_matches_oracle(n):
    print(f"[mlp gradient n={n}]", flush=True)
    s = random_samples(n, seed=100 + n)
    p = params(2)
    loss, g = mlp_loss_grad(s, p)
    rl, rg = M.loss_grad(p, s)
    assert abs(loss - rl) <= 1e-5 * abs(rl)
    for name, (a, b) in GROUPS.items():
        err = np.linalg.norm(g[a:b] - rg[a:b]) / max(1e-30, np.linalg.norm(rg[a:b]))
        assert err < 1e-4, (name, err)


def small_case():
    inst = synth.make_custom(10, 3, 3, 8, 60, 50, seed_ints=5)
    return inst, O.System.from_instance(inst), Engine.from_instance(inst)


def test_apply_matches_oracle():
    inst, so, eng = small_case()
    hii = so.hdiag()
    r = np.random.default_rng(4).standard_normal(so.ndet)
    theta = hii.min() - 0.4
    p = params(3)
    eng.mlp_set(p)
    got = eng.mlp_apply(theta, r)
    ref = M.precondition(p, inst.alpha, inst.beta, hii, theta, r)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    eng.mlp_set(None)
    with pytest.raises(SbdxError):
        eng.mlp_apply(theta, r)


def test_record_matches_oracle():
    inst, so, eng = small_case()
    hii = so.hdiag()
    xs = so.davidson(max_it=200, num_block=10, eps=1e-10)["W"]
    x = xs + 3e-3 * np.random.default_rng(6).standard_normal(so.ndet)
    dx = torch.tensor(x, device="cuda")
    dxs = torch.tensor(xs, device="cuda")
    target = 400
    got, cnt, theta = eng.mlp_record(dx.data_ptr(), dxs.data_ptr(), target, seed=9)
    ref, keep, rtheta, p = M.record(lambda v: so.mult(v, hii), hii, inst.alpha, inst.beta, x, xs,
                                    target, 9)
    assert abs(theta - rtheta) < 1e-12 * abs(rtheta)
    assert cnt == got.size
    key = lambda s: s["alpha"].astype(object) * (1 << 64) + s["beta"].astype(object)
    gk = {k: i for i, k in enumerate(key(got))}
    rk = {k: i for i, k in enumerate(key(ref))}
    common = sorted(set(gk) & set(rk))
    # membership can differ only where p_i sits within rounding of the hash value
    assert len(common) >= 0.99 * max(len(gk), len(rk))
    g = got[[gk[k] for k in common]]
    r = ref[[rk[k] for k in common]]
    for f in ("xh", "xr", "u", "v"):
        assert np.max(np.abs(g[f] - r[f])) <= 1e-5 * max(1.0, np.max(np.abs(r[f]))), f


def test_zero_last_layer_is_the_diagonal_preconditioner():
    """A network with w3 = b3 = 0 outputs m = 0: preconditioner 1 returns r / (theta - H_ii)
    to the last bit or two, and the solver follows the diagonal preconditioner's iteration (to
    the rounding of those last bits, which the iteration amplifies) to the same energy."""
    inst, so, eng = small_case()
    hii = eng.hdiag()  # the engine's own diagonal (the oracle's agrees to 1e-12, not bitwise)
    r = np.random.default_rng(8).standard_normal(so.ndet)
    theta = hii.min() - 0.25
    eng.mlp_set(M.init_params(5, zero_last=True))
    ref = r / (theta - hii)
    assert np.max(np.abs(eng.mlp_apply(theta, r) - ref) / np.abs(ref)) <= 4e-16
    a = eng.davidson(method=1, precision=1, num_block=6, max_iteration=60, eps=1e-9, preconditioner=0)
    b = eng.davidson(method=1, precision=1, num_block=6, max_iteration=60, eps=1e-9, preconditioner=1)
    assert a["converged"] and b["converged"] and b["precond_fallback"] == 0
    assert abs(a["energy"] - b["energy"]) < 1e-9
    k = min(a["trace"].size, b["trace"].size)
    assert np.max(np.abs(a["trace"][:k] - b["trace"][:k])) < 1e-6
    assert abs(a["steps"] - b["steps"]) <= 6


def c5_case():
    """The learned preconditioner's workload family at test size: 48 orbitals, 12 + 12
    electrons, rank-96 integrals, 200 x 220 sampled strings (44,000 determinants)."""
    inst = synth.make_custom(48, 12, 12, 96, 200, 220, seed_ints=3)
    return inst, O.System.from_instance(inst), Engine.from_instance(inst)


def test_c5_family_record_train_apply_solve():
    """Every GPU part of the learned preconditioner on the 48-orbital family against the
    oracle: the recorder's pairs (features, diagonal and exact corrections, weights), the loss
    and gradient on those pairs, training on the GPU, the trained network's correction, and
    the mixed-precision solver with it reaching the FP64 energy."""
    print("[mlp c5 family 200x220]", flush=True)
    inst, so, eng = c5_case()
    hii = so.hdiag()
    xs = so.davidson(max_it=300, num_block=10, eps=1e-10)["W"]
    # an iterate 0.96 along x*: on this family the weights concentrate on few determinants
    # near x*, so the recorder keeps ~1,000 of 44,000 at this target
    x = xs + 0.3 * np.random.default_rng(11).standard_normal(so.ndet) / np.sqrt(so.ndet)
    x /= np.linalg.norm(x)
    dx = torch.tensor(x, device="cuda")
    dxs = torch.tensor(xs, device="cuda")
    got, cnt, theta = eng.mlp_record(dx.data_ptr(), dxs.data_ptr(), 30000, seed=21)
    ref, _, rtheta, _ = M.record(lambda v: so.mult(v, hii), hii, inst.alpha, inst.beta, x, xs, 30000, 21)
    assert abs(theta - rtheta) < 1e-12 * abs(rtheta) and cnt == got.size and got.size > 300
    key = lambda s: s["alpha"].astype(object) * (1 << 64) + s["beta"].astype(object)
    gk = {k: i for i, k in enumerate(key(got))}
    rk = {k: i for i, k in enumerate(key(ref))}
    common = sorted(set(gk) & set(rk))
    assert len(common) >= 0.99 * max(len(gk), len(rk))
    g, r = got[[gk[k] for k in common]], ref[[rk[k] for k in common]]
    for f in ("xh", "xr", "u", "v"):
        assert np.max(np.abs(g[f] - r[f])) <= 1e-5 * max(1.0, np.max(np.abs(r[f]))), f
    # loss and gradient of a non-trivial network on the recorded pairs
    p = params(7)
    loss, grad = mlp_loss_grad(got, p)
    rl, rg = M.loss_grad(p, got)
    assert abs(loss - rl) <= 1e-5 * abs(rl)
    for name, (a, b) in GROUPS.items():
        assert np.linalg.norm(grad[a:b] - rg[a:b]) <= 1e-4 * max(1e-30, np.linalg.norm(rg[a:b])), name
    # training from the diagonal preconditioner (zero last layer) lowers the loss
    p0 = mlp_train.init_params(0)
    l0, _ = mlp_loss_grad(got, p0)
    pt, trace = train(got, p0, epochs=40, batch=256, lr=2e-3, seed=3)
    l1, _ = mlp_loss_grad(got, pt)
    assert np.all(np.isfinite(trace)) and l1 < l0, (l0, l1)
    # the trained network's correction, then the solver with it
    eng.mlp_set(pt)
    rr = np.random.default_rng(12).standard_normal(so.ndet)
    th = hii.min() - 0.3
    assert np.max(np.abs(eng.mlp_apply(th, rr) - M.precondition(pt, inst.alpha, inst.beta, hii, th, rr))) \
        <= 1e-5 * np.max(np.abs(rr / (th - hii)))
    # from the HF vector the one-vector FP64 solver settles on an excited state of this instance;
    # the FP64 reference is the two-vector block solver, which reaches the lowest one
    e64 = eng.davidson(method=1, precision=0, block=2, num_block=8, max_iteration=400, eps=1e-9)["energy"]
    res = eng.davidson(method=1, precision=1, block=2, num_block=8, max_iteration=400, eps=1e-9,
                       preconditioner=1)
    assert res["converged"] and abs(res["energy"] - e64) < 1e-9, (res["energy"], e64)


def test_training_lowers_the_loss_and_the_solver_stays_exact():
    """Pairs recorded from two 36-orbital instances train the network on the GPU; the loss
    drops below the untrained (diagonal) value, and on a held-out instance the learned
    preconditioner converges to the same FP64 energy (1e-9 Ha)."""
    print("[mlp training]", flush=True)
    pairs = []
    for sd in (100, 101):
        eng = Engine.from_instance(mlp_train.family_instance("c2", 150, sd))
        s, info = mlp_train.collect(eng, 20000, sd)
        assert info["converged"] and s.size > 0
        pairs.append(s)
    samples = np.concatenate(pairs)
    p0 = mlp_train.init_params(0)
    l0, _ = mlp_loss_grad(samples, p0)
    p, trace = train(samples, p0, epochs=15, batch=4096, lr=2e-3, seed=1)
    l1, _ = mlp_loss_grad(samples, p)
    assert np.all(np.isfinite(trace)) and l1 < l0, (l0, l1)
    eng = Engine.from_instance(mlp_train.family_instance("c2", 150, 200))
    x = torch.zeros(eng.ndet, dtype=torch.float64, device="cuda")
    ref = mlp_train.solve(eng, x.data_ptr(), eps=1e-9)
    eng.mlp_set(p)
    r = mlp_train.solve(eng, x.data_ptr(), preconditioner=1, eps=1e-9)
    assert r["converged"] and abs(r["energy"] - ref["energy"]) < 1e-9


@pytest.mark.xfail(strict=False, reason=(
    "expected negative: on these rank-72 random-integral instances a per-determinant scaling "
    "of the diagonal correction is a weak lever (the least-squares scalings that best match the "
    "exact corrections are near zero for most determinant classes), so the network trained on "
    "two small instances does not reliably take fewer steps than the diagonal preconditioner; "
    "the measured gains and losses at scale are recorded with the workload's profiles"))
def test_learned_preconditioner_takes_fewer_steps_on_a_held_out_instance():
    """Steps to 1e-6 Ha on a held-out 36-orbital instance with the diagonal and with the trained
    learned preconditioner (network trained on pairs from two other seeds); the gain claimed is
    fewer steps."""
    print("[mlp gain]", flush=True)
    pairs = [mlp_train.collect(Engine.from_instance(mlp_train.family_instance("c2", 150, sd)),
                               20000, sd)[0] for sd in (100, 101)]
    p, _ = train(np.concatenate(pairs), mlp_train.init_params(0), epochs=15, batch=4096, lr=2e-3,
                 seed=1)
    eng = Engine.from_instance(mlp_train.family_instance("c2", 150, 200))
    x = torch.zeros(eng.ndet, dtype=torch.float64, device="cuda")
    ref = mlp_train.solve(eng, x.data_ptr(), eps=1e-9)
    ev = mlp_train.evaluate(eng, p, ref["energy"])
    print(f"steps to 1e-6 Ha: diagonal {ev[0]['steps_to_tol']}, learned {ev[1]['steps_to_tol']} "
          f"(fallback at {ev[1]['fallback']})", flush=True)
    assert 0 < ev[1]["steps_to_tol"] < ev[0]["steps_to_tol"]
