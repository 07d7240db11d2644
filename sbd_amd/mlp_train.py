"""Training pipeline of the learned Davidson preconditioner (GPU; the kernels are in libsbdx).

    python -m sbd_amd.mlp_train --family c3 --strings 2000 --train 100-115 --eval 200-203 \\
        --eval-strings 10000 --out sbd_amd/weights/c3_family.npy

For every training seed of a workload family (integrals, alpha and beta strings all drawn from
that seed): solve to convergence (block Davidson, mixed precision), then rerun from the HF
vector one thick-restart cycle at a time with the diagonal preconditioner and, at every
restart iterate x, record importance-sampled (features, diagonal correction, exact correction)
pairs against the converged x* (sbdx_mlp_record).  The pairs of all seeds train the network
with Adam on the GPU (sbdx_mlp_train).  Evaluation on held-out seeds: steps and wall time to
1e-6 Ha of the converged energy with the diagonal and with the learned preconditioner.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import synth
from ._lib import MLP_NPARAM, MLP_SAMPLE, Engine, mlp_train

FAMILIES = {
    # name: (norb, electrons per spin, integral rank)
    "c2": (36, 8, 72),
    "c3": (48, 12, 96),
}

SOLVE = dict(method=1, precision=1, block=1, num_block=10)


def family_instance(family, strings, seed):
    L, ne, rank = FAMILIES[family]
    return synth.make_custom(L, ne, ne, rank, strings, strings, seed_ints=seed,
                             seed_alpha=10_000 + seed, seed_beta=20_000 + seed,
                             name=f"{family}/{strings}/seed{seed}")


def init_params(seed=0):
    """He-style first two layers, zero last layer: the untrained network is the diagonal
    preconditioner exactly."""
    rng = np.random.default_rng(seed)
    H, NIN = 64, 130
    p = np.zeros(MLP_NPARAM, np.float32)
    p[:NIN * H] = rng.standard_normal(NIN * H) * np.sqrt(2.0 / 24.0)
    off_w2 = NIN * H + H
    p[off_w2:off_w2 + H * H] = rng.standard_normal(H * H) * np.sqrt(2.0 / H)
    return p


def solve(eng, dW, **kw):
    opts = dict(SOLVE, max_iteration=2000, eps=1e-8)
    opts.update(kw)
    return eng.davidson_device(dW, **opts)


def collect(eng, target, seed, max_snapshots=40, stop_residual=1e-6, log=None):
    """Training pairs of one instance; returns (samples sorted by determinant, info)."""
    import torch
    N = eng.ndet
    xs = torch.zeros(N, dtype=torch.float64, device="cuda")
    ref = solve(eng, xs.data_ptr())
    x = torch.zeros(N, dtype=torch.float64, device="cuda")
    out, steps = [], 0
    for j in range(max_snapshots):
        r = solve(eng, x.data_ptr(), max_iteration=1, eps=0.0, init=0 if j == 0 else 1)
        steps += r["steps"]
        if r["residual"] < stop_residual:
            break
        s, cnt, theta = eng.mlp_record(x.data_ptr(), xs.data_ptr(), target, seed=seed * 1000 + j)
        out.append(s)
        if log:
            log(f"  snapshot {j}: E {r['energy']:.10f} residual {r['residual']:.3e} kept {s.size}")
    samples = np.concatenate(out) if out else np.zeros(0, dtype=MLP_SAMPLE)
    samples = np.sort(samples, order=["alpha", "beta"], kind="stable")
    return samples, dict(energy=ref["energy"], converged=ref["converged"], snapshots=len(out))


def evaluate(eng, params, energy_ref, tol=1e-6, max_time=600.0, **kw):
    """Steps and seconds to |E - energy_ref| < tol with the diagonal (0) and learned (1)
    preconditioners."""
    import torch
    res = {}
    eng.mlp_set(params)
    for pc in (0, 1):
        x = torch.zeros(eng.ndet, dtype=torch.float64, device="cuda")
        r = solve(eng, x.data_ptr(), preconditioner=pc, energy_ref=energy_ref, energy_tol=tol,
                  stop_at_tol=1, eps=1e-12, max_time=max_time, **kw)
        res[pc] = dict(steps_to_tol=r["steps_to_tol"], seconds_to_tol=r["seconds_to_tol"],
                       energy=r["energy"], fallback=r["precond_fallback"])
    eng.mlp_set(None)
    return res


def seeds(spec):
    a, _, b = spec.partition("-")
    return list(range(int(a), int(b or a) + 1))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--family", default="c3", choices=sorted(FAMILIES))
    ap.add_argument("--strings", type=int, default=2000)
    ap.add_argument("--train", default="100-115")
    ap.add_argument("--eval", default="200-203")
    ap.add_argument("--eval-strings", type=int, default=None,
                    help="strings per spin of the held-out instances (default: --strings)")
    ap.add_argument("--target", type=float, default=50000.0, help="samples per snapshot")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    args = ap.parse_args(argv)
    import torch
    torch.zeros(1, device="cuda")  # device memory comes from torch: initialise it first

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    pairs = []
    for sd in seeds(args.train):
        inst = family_instance(args.family, args.strings, sd)
        eng = Engine.from_instance(inst)
        s, info = collect(eng, args.target, sd, log=log)
        log(f"[train seed {sd}] {inst.ndet} dets, E {info['energy']:.10f}, "
            f"{info['snapshots']} snapshots, {s.size} pairs")
        pairs.append(s)
        del eng
    samples = np.concatenate(pairs)
    t_collect = time.perf_counter() - t0
    p0 = init_params(0)
    t1 = time.perf_counter()
    params, trace = mlp_train(samples, p0, epochs=args.epochs, batch=args.batch, lr=args.lr,
                              seed=1, verbose=0)
    t_train = time.perf_counter() - t1
    log(f"[train] {samples.size} pairs, loss {trace[0]:.4e} -> {trace[-1]:.4e} in {t_train:.1f}s")
    report = dict(family=args.family, strings=args.strings, pairs=int(samples.size),
                  collect_s=round(t_collect, 2), train_s=round(t_train, 2),
                  loss=[float(v) for v in trace], eval={})
    for sd in seeds(args.eval):
        inst = family_instance(args.family, args.eval_strings or args.strings, sd)
        eng = Engine.from_instance(inst)
        import torch
        x = torch.zeros(eng.ndet, dtype=torch.float64, device="cuda")
        ref = solve(eng, x.data_ptr(), eps=1e-9)
        ev = evaluate(eng, params, ref["energy"])
        report["eval"][sd] = dict(ndet=inst.ndet, energy=ref["energy"], diagonal=ev[0], learned=ev[1])
        log(f"[eval seed {sd}] diagonal {ev[0]['steps_to_tol']} steps {ev[0]['seconds_to_tol']:.2f}s, "
            f"learned {ev[1]['steps_to_tol']} steps {ev[1]['seconds_to_tol']:.2f}s "
            f"(fallback at {ev[1]['fallback']})")
        del eng
    if args.out:
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        np.save(args.out, params)
    print(json.dumps(report))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
