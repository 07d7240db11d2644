#!/usr/bin/env python3
"""Benchmark: the selected-basis engine at 4e8 determinants on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--prec fp64|fp32] [--no-solve]

Workload (synthetic, seeded; sbd_amd/synth.py "c2"): L = 36 orbitals, 8 alpha + 8 beta
electrons, rank-72 integrals, 20,000 x 20,000 Fermi-sampled strings = 4e8 determinants.
A step is one sigma build H C over the whole 4e8-element vector, with C resident in HBM.

The JSON line carries:
  * value          -- sigma builds/s (FP64 unless --prec fp32);
  * roofline       -- the dominant kernel of the build (by mean duration, HIP events on its
                      launch stream) against HBM peak, from its compulsory bytes (every operand
                      read once from HBM, every output written once); `traffic` is its
                      counter-measured HBM-side bytes per launch, and `l2` puts the same kernel
                      against the L2 -> L1 ceiling measured live by a probe, from its
                      counter-measured L2 -> L1 request bytes per launch;
  * sigma_roofline -- the whole build, compulsory bytes against HBM peak;
  * solve          -- time to 1e-6 Ha of the converged FP64 energy from the HF vector: mixed
                      precision block Davidson with the learned preconditioner, the same
                      without it, and the plain FP64 Davidson (the reference's algorithm), the
                      last as a bounded run's per-step time x the full run's step count;
  * cpu_baseline   -- the reference's CPU sigma (compiled sbd, MPI world size 1 + OpenMP) on
                      three reduced instances of the same family, the element rate fitted over
                      them and extrapolated by exact off-diagonal element counts; the oracle
                      port stands in when the compiled reference is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("time-to-1e-6 Ha & sigma builds/s at 4e8 dets, 1 GPU: plain FP64 vs "
          "mixed+block+MLP")
HBM_PEAK_GBS = 8000.0
L2_PEAK_BYTES_PER_CLK = 8 * 16 * 128  # XCDs x L2 channels x bytes per channel per clock

# Davidson variants of --solve: the reference algorithm in FP64, and the block thick-restart
# solver in FP64 or mixed precision (FP32 corrections, FP64 anchor) with 1 or 2 vectors a block
SOLVERS = {
    "plain": dict(method=0, num_block=10),
    "block_fp64": dict(method=1, precision=0, block=1, num_block=10),
    "mixed": dict(method=1, precision=1, block=1, num_block=10),
    "mixed_block2": dict(method=1, precision=1, block=2, num_block=8),
    "mixed_block4": dict(method=1, precision=1, block=4, num_block=6),
    "mixed_block2_nb12": dict(method=1, precision=1, block=2, num_block=12),
    "mixed_block2_nb16": dict(method=1, precision=1, block=2, num_block=16),
    "mixed_block2_nb20": dict(method=1, precision=1, block=2, num_block=20),
    "mixed_block2_nb24": dict(method=1, precision=1, block=2, num_block=24),
    # FP32 sigma with compensated two-sum FP32 partials instead of FP64 accumulation of the
    # exact FP32 products
    "mixed_block2_nb24_twosum": dict(method=1, precision=1, block=2, num_block=24, fp32_acc=1),
    "mixed_block2_nb28": dict(method=1, precision=1, block=2, num_block=28),
    "mixed_block2_nb32": dict(method=1, precision=1, block=2, num_block=32),
    "block2_fp64_nb16": dict(method=1, precision=0, block=2, num_block=16),
    # the learned preconditioner trained on 36-orbital family instances (sbd_amd/mlp_train.py)
    "mixed_block2_nb16_mlp": dict(method=1, precision=1, block=2, num_block=16, preconditioner=1,
                                  mlp="sbd_amd/weights/c2_family.npy"),
    "mixed_block2_nb24_mlp": dict(method=1, precision=1, block=2, num_block=24, preconditioner=1,
                                  mlp="sbd_amd/weights/c2_family.npy"),
    # the network only after two diagonal restart cycles
    "mixed_block2_nb24_mlp_w2": dict(method=1, precision=1, block=2, num_block=24, preconditioner=1,
                                     precond_warmup=2, mlp="sbd_amd/weights/c2_family.npy"),
    # eight lowest roots in one block of eight
    "mixed_roots8": dict(method=1, precision=1, nroots=8, num_block=4),
    # the same with 5 blocks (fits 288 GB since blocks of 8 are built as two-vector sub-blocks)
    "mixed_roots8_nb5": dict(method=1, precision=1, nroots=8, num_block=5),
}


# the default run's solves: mixed precision block Davidson with and without the learned
# preconditioner, and with the FP32 sigma's two-sum accumulation (the plain FP64 solver is timed
# by a bounded run; see plain_full_run)
DEFAULT_SOLVES = ("mixed_block2_nb16_mlp", "mixed_block2_nb24_twosum", "mixed_block2_nb24")
DEFAULT_SOLVE_TIME = 180.0


def plain_full_run(workload):
    """Step count to 1e-6 Ha of the plain FP64 solver's full run on the workload
    (profiles/energies.json), or None."""
    path = os.path.join(ROOT, "profiles", "energies.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get(workload, {}).get("plain")
    return rec


def reference_energy(workload, strings):
    """Converged FP64 ground-state energy of the workload (profiles/energies.json), or None."""
    path = os.path.join(ROOT, "profiles", "energies.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    key = workload if strings is None else f"{workload}/{strings}"
    e = table.get(key)
    return None if e is None else float(e["energy"])


def log(*a):
    print(*a, file=sys.stderr, flush=True)


PHASE_KERNELS = {
    "alpha_rows": "k_sigma_rows<alpha> (diagonal + alpha singles/doubles)",
    "mixed": "k_sigma_mixed (alpha single x beta single)",
    "transpose": "k_transpose",
    "beta_rows": "k_sigma_rows<beta> (beta singles/doubles, transposed)",
    "add_transposed": "k_add_transposed",
}


def nnz_parts(eng):
    """Exact off-diagonal element counts of one sigma (same-spin alpha, same-spin beta, mixed)."""
    sa, da = eng.excitation_counts(0)
    sb, db = eng.excitation_counts(1)
    na, nb = eng.alpha.shape[0], eng.beta.shape[0]
    return dict(aa=nb * (sa + da), bb=na * (sb + db), ab=sa * sb)


def compulsory_bytes(eng, w):
    """Bytes each phase of one K = 1 sigma build must move between HBM and the chip at least:
    every operand read once and every output written once.  w: vector word size (8 FP64, 4
    FP32; H_ii, the list values and the integrals are held in the same precision)."""
    sa, da = eng.excitation_counts(0)
    sb, db = eng.excitation_counts(1)
    na, nb = eng.alpha.shape[0], eng.beta.shape[0]
    N = na * nb
    V = w * eng.norb ** 4  # dense (pq|rs)

    def lists(n, s, d):  # offsets, single kets + packed words + values, double kets + values
        return 16 * (n + 1) + (8 + w) * s + (4 + w) * d

    return {
        "alpha_rows": 3 * w * N + lists(na, sa, da) + 8 * nb + V,  # x, sigma, H_ii, lists, strings
        "mixed": 3 * w * N + 8 * sb + 8 * (na + 1) + 8 * sa + V,   # x, sigma read+write, tiles
        "transpose": 2 * w * N,
        "beta_rows": 2 * w * N + lists(nb, sb, db) + 8 * na + V,
        "add_transposed": 3 * w * N,
    }


def gathered_bytes(parts, N, w):
    """One vector word per off-diagonal element plus the streamed vectors: what each phase
    would move if no gathered word were ever served twice from a cache."""
    return {
        "alpha_rows": w * (parts["aa"] + 3 * N),
        "mixed": w * (parts["ab"] + 3 * N),
        "transpose": 2 * w * N,
        "beta_rows": w * (parts["bb"] + 2 * N),
        "add_transposed": 3 * w * N,
    }


def counters(workload, strings, prec):
    """Committed rocprofv3 PMC figures per launch of each phase's kernel (profiles/counters.json)
    for this workload and precision, or {}."""
    path = os.path.join(ROOT, "profiles", "counters.json")
    if strings is not None or not os.path.exists(path):
        return {}
    return json.load(open(path)).get(f"{workload}/{prec}", {})


def cpu_threads():
    """Host threads the CPU baseline uses: OMP_NUM_THREADS where it is set (the host cores
    allotted to this GPU's jobs), else every core."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_sigma_time(inst, threads):
    """One reference CPU sigma on a reduced instance: (seconds, off-diagonal elements, kind)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    so = O.System.from_instance(inst)
    sa = int(so.conn_a[0][-1]); da = int(so.conn_a[2][-1])
    sb = int(so.conn_b[0][-1]); db = int(so.conn_b[2][-1])
    nnz = inst.beta.shape[0] * (sa + da) + inst.alpha.shape[0] * (sb + db) + sa * sb
    if O.have_reference():
        try:
            line, _ = O.run_reference(inst, "timemult", threads=threads, extra=["--reps", "1"])
            return line["best_s"], nnz, "reference"
        except (subprocess.CalledProcessError, OSError) as e:  # reference binary unusable here
            log(f"compiled reference unavailable ({e}); timing the oracle port")
    O.set_threads(threads)
    hii = so.hdiag()
    x = np.random.default_rng(7).uniform(-1, 1, so.ndet)
    t0 = time.perf_counter()
    so.mult(x, hii)
    return time.perf_counter() - t0, nnz, "port"


def cpu_baseline(family, nnz_target, sizes=(600, 1000, 1400)):
    """Reference CPU sigma on reduced instances of the same family (n x n strings), the element
    rate fitted over them (t = t0 + elements / rate, least squares) and extrapolated to the
    workload's exact element count."""
    from sbd_amd import synth
    threads = cpu_threads()
    rows = []
    for n in sizes:
        inst = synth.make_instance(family, strings=n)
        t, nnz, kind = cpu_sigma_time(inst, threads)
        rows.append((n, nnz, t, kind))
        log(f"[bench] cpu baseline {n}x{n}: {nnz} elements, {t:.2f} s ({kind}, {threads} threads)")
    e = np.array([r[1] for r in rows], float)
    t = np.array([r[2] for r in rows], float)
    slope, t0 = np.polyfit(e, t, 1)
    rate = 1.0 / slope if slope > 0 else float(e.sum() / t.sum())
    kinds = {r[3] for r in rows}
    sample = "; ".join(f"{r[0]}x{r[0]} strings ({r[1]} elements) {r[2]:.2f} s" for r in rows)
    return {
        "value": rate / nnz_target,
        "unit": "sigma builds/s (element rate fitted over three sizes, extrapolated by element count)",
        "cores": threads,
        "kind": "reference" if kinds == {"reference"} else "port",
        "sample": f"{family} family, one sigma each: {sample}; fit {rate:.3e} elements/s, "
                  f"intercept {t0:.2f} s",
        "elements_per_s": rate,
        "nproc": os.cpu_count(),
    }


def rdm_energy(inst, one, two):
    """E = I0 + sum_ij h_ij (g_a + g_b)[i + L j] + 1/2 sum (ik|jl) G[i + L j + L^2 k + L^3 l]
    from the RDMs in the reference's layout."""
    L = inst.norb
    g = one[0].reshape(L, L, order="F") + one[1].reshape(L, L, order="F")
    G = two.sum(axis=0).reshape(L, L, L, L, order="F")
    eri = inst.ints.eri
    return float(inst.ints.ecore + np.sum(inst.ints.h1 * g)
                 + 0.5 * np.tensordot(eri.transpose(0, 2, 1, 3), G, axes=4))


def rdm_after_solve(eng, inst, W, name, energy):
    """1- and 2-RDM of a solve's vector (device resident), timed; with the checks the sizes
    allow at any scale: the traces of the 1-RDM (electron counts) and the energy identity."""
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    one, two = eng.rdm_device(W.data_ptr())
    dt = time.perf_counter() - t
    L = inst.norb
    tr = [float(np.trace(one[s].reshape(L, L))) for s in (0, 1)]
    e = rdm_energy(inst, one, two)
    rec = dict(vector=name, seconds=round(dt, 3), trace_alpha=tr[0], trace_beta=tr[1],
               electrons=[inst.n_alpha_el, inst.n_beta_el], energy_from_rdm=e,
               energy_of_solve=energy, energy_identity_error=abs(e - energy))
    log(f"[bench] rdm: {rec}")
    return rec


def rdm_random(eng, inst, g):
    """1- and 2-RDM of a random normalized vector (device resident), timed; the energy identity
    against <W|H|W> from one FP64 sigma of the same vector holds for any normalized W."""
    import torch
    N = eng.ndet
    W = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    W /= torch.linalg.vector_norm(W)
    y = torch.empty_like(W)
    eng.sigma_device(0, 1, W.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()  # the build runs on the context's stream
    e_hw = float(torch.dot(W, y))
    del y
    torch.cuda.synchronize()
    t = time.perf_counter()
    one, two = eng.rdm_device(W.data_ptr())
    dt = time.perf_counter() - t
    L = inst.norb
    tr = [float(np.trace(one[s].reshape(L, L))) for s in (0, 1)]
    e = rdm_energy(inst, one, two)
    rec = dict(vector="random, normalized", seconds=round(dt, 3), trace_alpha=tr[0], trace_beta=tr[1],
               electrons=[inst.n_alpha_el, inst.n_beta_el], energy_from_rdm=e, energy_hw=e_hw,
               energy_identity_error=abs(e - e_hw))
    log(f"[bench] rdm (random vector): {rec}")
    del W
    torch.cuda.empty_cache()
    return rec


def precision_report(eng, max_time):
    """The workload solved to residual 1e-8 twice -- the reference algorithm in FP64 and the
    mixed-precision block solver (FP32 corrections, FP64 anchor refinement) -- and the errors of
    the mixed run's energy, 1-RDM and 2-RDM against the FP64 run."""
    import torch
    N = eng.ndet
    out = {}
    vecs = {}
    for name, opts in (("fp64_reference_algorithm", dict(method=0, num_block=10)),
                       ("mixed_block2_nb16", SOLVERS["mixed_block2_nb16"])):
        W = torch.zeros(N * 2, dtype=torch.float64, device="cuda")
        t = time.perf_counter()
        r = eng.davidson_device(W.data_ptr(), **dict(opts, max_iteration=100000, eps=1e-8,
                                                     max_time=max_time))
        dt = time.perf_counter() - t
        w = W[:N].cpu().numpy()
        t = time.perf_counter()
        one, two = eng.rdm(w)
        t_rdm = time.perf_counter() - t
        vecs[name] = (w, one, two)
        out[name] = dict(energy=r["energy"], residual=r["residual"], converged=r["converged"],
                         steps=r["steps"], seconds=round(dt, 3), rdm_seconds=round(t_rdm, 3))
        log(f"[bench] precision report {name}: {out[name]}")
        del W
        torch.cuda.empty_cache()
    (w0, o0, t0_), (w1, o1, t1_) = vecs["fp64_reference_algorithm"], vecs["mixed_block2_nb16"]
    out["energy_error"] = abs(out["mixed_block2_nb16"]["energy"] - out["fp64_reference_algorithm"]["energy"])
    out["one_rdm_max_error"] = float(np.max(np.abs(o1 - o0)))
    out["two_rdm_max_error"] = float(np.max(np.abs(t1_ - t0_)))
    out["vector_overlap"] = float(abs(w0 @ w1))
    return out


def batch_bench(count, prec):
    """Many small problems (norb 10..20, 1e3..1e6 determinants, log-uniform) solved to 1e-8
    one after another with the block solver, then together with the batched graph-replayed
    solver; energies must agree to 1e-9 Ha."""
    import torch
    from sbd_amd import synth
    from sbd_amd._lib import Batch, Engine
    insts = synth.make_batch_instances(count, seed=0)
    engs = [Engine.from_instance(i) for i in insts]
    opts = dict(num_block=10, max_iteration=2000, eps=1e-8, precision=prec)
    for e in engs:  # warm every problem's tables and kernels
        e.davidson(method=1, max_iteration=1, num_block=2, eps=0.0, precision=prec)
    torch.cuda.synchronize()
    t = time.perf_counter()
    single = [e.davidson(method=1, **opts) for e in engs]
    t_seq = time.perf_counter() - t
    b = Batch(engs)
    b.davidson(**dict(opts, max_iteration=1))  # capture the step graphs once
    t = time.perf_counter()
    res = b.davidson(**opts)
    t_bat = time.perf_counter() - t
    b.close()
    diff = max(abs(r["energy"] - s["energy"]) for r, s in zip(res, single))
    rec = dict(count=count, dets_total=int(sum(i.ndet for i in insts)),
               dets_min=int(min(i.ndet for i in insts)), dets_max=int(max(i.ndet for i in insts)),
               precision="fp32" if prec else "fp64", sequential_s=round(t_seq, 3),
               batched_s=round(t_bat, 3), speedup=round(t_seq / t_bat, 3),
               converged=int(sum(r["converged"] for r in res)),
               max_energy_diff_vs_single=diff)
    log(f"[bench] batch: {rec}")
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--strings", type=int, default=None, help="reduced string count (tests)")
    ap.add_argument("--prec", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--fp32-acc", default="fp64", choices=["fp64", "twosum"],
                    help="accumulation of FP32 sigma builds: exact products summed in FP64, or "
                         "compensated FP32 partials (two-sum)")
    ap.add_argument("--solve", nargs="*", default=None, choices=sorted(SOLVERS),
                    help="also time these Davidson variants to 1e-6 Ha of the reference energy")
    ap.add_argument("--solve-time", type=float, default=240.0)
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the default time-to-1e-6 Ha solves (c2, FP64 runs only)")
    ap.add_argument("--no-rdm", action="store_true",
                    help="skip the 1- and 2-RDM of the last solve's vector")
    ap.add_argument("--rdm-random", action="store_true",
                    help="time the 1- and 2-RDM of a random normalized vector (energy identity "
                         "against <W|H|W>)")
    ap.add_argument("--plain-steps", type=int, default=20,
                    help="steps of the bounded plain FP64 run that gives its seconds per step")
    ap.add_argument("--converge", action="store_true",
                    help="--solve runs to residual 1e-8 instead of stopping at the reference energy")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one plain 1 GiB device copy (rocprof traffic calibration)")
    ap.add_argument("--precision-report", action="store_true",
                    help="solve to 1e-8 in FP64 and in mixed precision; energy and RDM errors")
    ap.add_argument("--batch", type=int, default=0,
                    help="also solve this many small problems one by one and as one batch")
    ap.add_argument("--batch-prec", type=int, default=1, choices=[0, 1])
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the mixed pass of each --block width before timing it")
    ap.add_argument("--block", type=int, nargs="*", default=[],
                    help="also time the fused sigma of blocks of these widths (1, 2, 4, 8)")
    ap.add_argument("--variant", type=int, default=0,
                    help="mixed-kernel variant of the timed one-vector build (sbdx_set_mix_variant; 0: default)")
    ap.add_argument("--rename", type=int, nargs=2, default=None, metavar=("NORB", "SHIFT"),
                    help="the same problem with orbital p renamed p + SHIFT among NORB orbitals "
                         "(NORB > 64: two-word strings; same H, same sigma); times the sigma "
                         "build only, unless solves are named with --solve")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design, not a build)"}))
        sys.exit(2)

    import torch
    from sbd_amd import synth
    from sbd_amd._lib import Engine, SBDX_FP32, SBDX_FP64, bandwidth_probe

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(2)

    t = time.perf_counter()
    inst = synth.make_instance(args.workload, strings=args.strings)
    renamed = ""
    if args.rename:
        L0 = inst.norb
        inst = synth.embed(inst, args.rename[0], args.rename[1])
        renamed = f", orbitals renamed p -> p + {args.rename[1]} among {inst.norb} (from {L0})"
    t_gen = time.perf_counter() - t
    log(f"[bench] instance {inst.name}: L={inst.norb} {inst.alpha.shape[0]}x{inst.beta.shape[0]} "
        f"= {inst.ndet} dets (generated in {t_gen:.1f}s)")
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng = Engine.from_instance(inst)
    eng.set_fp32_accumulation(1 if args.fp32_acc == "twosum" else 0)
    if args.variant:
        eng.set_mix_variant(SBDX_FP64 if args.prec == "fp64" else SBDX_FP32, 1, args.variant)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t
    parts = nnz_parts(eng)
    nnz = sum(parts.values())
    N = eng.ndet
    log(f"[bench] setup {t_setup:.2f}s; off-diagonal elements per sigma {nnz:.4e} "
        f"(aa {parts['aa']:.3e}, bb {parts['bb']:.3e}, ab {parts['ab']:.3e})")

    prec = SBDX_FP64 if args.prec == "fp64" else SBDX_FP32
    dt = torch.float64 if args.prec == "fp64" else torch.float32
    w = 8 if args.prec == "fp64" else 4
    g = torch.Generator(device="cuda").manual_seed(1234)
    x = torch.randn(N, dtype=dt, device="cuda", generator=g)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream().cuda_stream

    if args.calibrate_copy:
        # counter calibration: one plain 1 GiB device copy (HBM bytes) and the L2 probe (L2 ->
        # L1 request bytes of a known byte count), each a kernel of its own under rocprofv3
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    for _ in range(args.warmup):
        eng.sigma_device(prec, 1, x.data_ptr(), y.data_ptr(), stream)
    torch.cuda.synchronize()
    eng.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        eng.sigma_device(prec, 1, x.data_ptr(), y.data_ptr(), stream)
        log(f"[bench] step {k + 1}/{args.steps} done")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases, calls = eng.profile_read()
    eng.profile(False)
    sec = (t1 - t0) / args.steps
    value = 1.0 / sec

    phase_ms = {k: v / max(calls, 1) for k, v in phases.items()}
    comp = compulsory_bytes(eng, w)
    gath = gathered_bytes(parts, N, w)
    dom = max(phase_ms, key=phase_ms.get)
    dom_s = phase_ms[dom] * 1e-3
    pmc = {} if args.rename else counters(args.workload, args.strings, args.prec).get(dom, {})
    # live ceiling of HBM: a STREAM-style copy of 4 GiB
    hbm_probe = bandwidth_probe(0, 4 << 30, 5)
    log(f"[bench] probe: HBM copy {hbm_probe:.0f} GB/s")
    l2_bytes = pmc.get("l2_to_l1_bytes")
    roofline = {
        "kernel": PHASE_KERNELS[dom],
        "bound": "hbm",
        "achieved": round(comp[dom] / dom_s / 1e9, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(comp[dom] / dom_s / 1e9 / HBM_PEAK_GBS, 6),
        "traffic": pmc.get("hbm_bytes"),
        "algorithmic_bytes": comp[dom],
        "algorithmic_model": "compulsory: every operand read once from HBM, every output written once",
        "mean_ms": round(phase_ms[dom], 3),
        "hbm_probe_gbs": round(hbm_probe, 1),
        # the level the gathers are served from: counter-measured L2 -> L1 bytes per launch
        # against the L2 read peak (16 channels x 128 B per clock per XCD, 8 XCDs, at the
        # kernel's measured clock)
        "l2": None if l2_bytes is None else {
            "achieved": round(l2_bytes / dom_s / 1e9, 1),
            "peak": round(L2_PEAK_BYTES_PER_CLK * pmc["clock_ghz"], 1),
            "unit": "GB/s",
            "frac": round(l2_bytes / dom_s / 1e9 / (L2_PEAK_BYTES_PER_CLK * pmc["clock_ghz"]), 4),
            "traffic": l2_bytes,
        },
        # the binding resource named by the counters: the vector-memory address unit
        "ta_busy": pmc.get("ta_busy"),
        "gathered_bytes": gath[dom],
        "gathered_rate_gbs": round(gath[dom] / dom_s / 1e9, 1),
        "counters": pmc.get("source"),
    }
    sigma_comp = sum(comp.values())
    out = {
        "metric": METRIC,
        "value": round(value, 6),
        "unit": "sigma builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.prec == "fp64" else "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{inst.name}: L={inst.norb}, {inst.n_alpha_el}a{inst.n_beta_el}b, "
                         f"rank-{inst.meta['rank']} integrals, {inst.alpha.shape[0]}x{inst.beta.shape[0]}"
                         f" Fermi-sampled strings, {N} determinants{renamed}"),
            "ndet": N,
            "offdiag_elements": nnz,
            "precision": args.prec,
            "fp32_accumulation": args.fp32_acc,
            **({"mix_variant": args.variant} if args.variant else {}),
        },
        "roofline": roofline,
        "sigma_roofline": {
            "bound": "hbm",
            "achieved": round(sigma_comp / sec / 1e9, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(sigma_comp / sec / 1e9 / HBM_PEAK_GBS, 6),
            "algorithmic_bytes": sigma_comp,
            "gathered_bytes": sum(gath.values()),
        },
        "phase_ms": {k: round(v, 3) for k, v in phase_ms.items()},
        "setup_s": round(t_setup, 3),
    }

    if args.block:
        del x, y
        torch.cuda.empty_cache()
        out["block_sigma"] = {}
        for k in args.block:
            tuned = None
            if args.autotune:
                chosen, cand = eng.autotune(prec, k, reps=1)
                tuned = {"variant": chosen, "candidates_ms": cand}
                log(f"[bench] autotune block {k}: {tuned}")
            xb = torch.randn(N * k, dtype=dt, device="cuda", generator=g)
            yb = torch.empty_like(xb)
            eng.sigma_block_device(prec, k, xb.data_ptr(), yb.data_ptr(), stream)
            torch.cuda.synchronize()
            eng.profile(True)
            t0 = time.perf_counter()
            reps = max(1, args.steps // 2)
            for _ in range(reps):
                eng.sigma_block_device(prec, k, xb.data_ptr(), yb.data_ptr(), stream)
            torch.cuda.synchronize()
            tb = (time.perf_counter() - t0) / reps
            ph, cl = eng.profile_read()
            eng.profile(False)
            out["block_sigma"][str(k)] = {
                "ms_per_block": round(tb * 1e3, 3),
                "vector_sigmas_per_s": round(k / tb, 4),
                "phase_ms": {p: round(v / max(cl, 1), 3) for p, v in ph.items()},
                "autotune": tuned,
            }
            log(f"[bench] block {k}: {tb * 1e3:.1f} ms, {k / tb:.3f} vector sigmas/s")
            del xb, yb
            torch.cuda.empty_cache()
        x = y = None

    # (a renamed workload is a sigma-build measurement: solves only when named with --solve)
    default_solve = (args.solve is None and not args.no_solve and args.workload == "c2"
                     and args.strings is None and args.prec == "fp64" and not args.rename)
    if args.solve is not None or default_solve:
        names = args.solve if args.solve is not None else list(DEFAULT_SOLVES)
        names = names or ["plain", "mixed_block2"]
        del x, y
        torch.cuda.empty_cache()
        e_ref = None if args.converge else reference_energy(args.workload, args.strings)
        out["solve"] = {"energy_ref": e_ref, "energy_tol": 1e-6}
        solve_time = args.solve_time if args.solve is not None else DEFAULT_SOLVE_TIME
        want_rdm = not args.no_rdm
        for name in names:
            opts = dict(SOLVERS[name], max_iteration=100000, eps=1e-8, max_time=solve_time,
                        verbose=int(args.verbose))
            weights = opts.pop("mlp", None)
            eng.mlp_set(np.load(os.path.join(ROOT, weights)) if weights else None)
            acc = opts.pop("fp32_acc", None)
            if acc is not None:
                eng.set_fp32_accumulation(acc)
            if e_ref is not None:
                opts.update(energy_ref=e_ref, energy_tol=1e-6, stop_at_tol=1)
            k = max(1, opts.get("nroots", 1))
            W = torch.empty(N * k, dtype=torch.float64, device="cuda")
            log(f"[bench] solve {name}: {opts}")
            r = eng.davidson_device(W.data_ptr(), **opts)
            tr = r["trace"]
            out["solve"][name] = dict(
                seconds_to_tol=r["seconds_to_tol"], steps_to_tol=r["steps_to_tol"],
                steps=r["steps"], seconds=round(r["seconds"], 3), converged=r["converged"],
                residual=r["residual"], energy=r["energy"],
                min_ritz=float(tr.min()) if tr.size else None)
            log(f"[bench] solve {name}: {out['solve'][name]}")
            if acc is not None:
                eng.set_fp32_accumulation(1 if args.fp32_acc == "twosum" else 0)
            if want_rdm and name == names[-1]:
                out["rdm"] = rdm_after_solve(eng, inst, W, name, r["energy"])
            del W
            torch.cuda.empty_cache()
        if "plain" not in names and e_ref is not None and args.plain_steps > 0:
            # plain FP64 Davidson (the reference's algorithm): a bounded run gives the seconds
            # per step; the step count to 1e-6 Ha is its full run's (profiles/energies.json)
            full = plain_full_run(args.workload)
            # method 0's max_iteration counts restart cycles of num_block steps each
            cycles = max(1, -(-args.plain_steps // SOLVERS["plain"]["num_block"]))
            opts = dict(SOLVERS["plain"], max_iteration=cycles, eps=1e-8,
                        max_time=solve_time, verbose=int(args.verbose))
            W = torch.empty(N, dtype=torch.float64, device="cuda")
            log(f"[bench] solve plain (bounded, {args.plain_steps} steps): {opts}")
            r = eng.davidson_device(W.data_ptr(), **opts)
            per_step = r["seconds"] / max(r["steps"], 1)
            rec = dict(steps=r["steps"], seconds=round(r["seconds"], 3),
                       seconds_per_step=round(per_step, 4))
            if full:
                rec.update(steps_to_tol=full["steps_to_tol"],
                           seconds_to_tol=round(per_step * full["steps_to_tol"], 2),
                           derived=("seconds per step of this bounded run x the step count to "
                                    "1e-6 Ha of the full run"),
                           full_run=full.get("source"))
            out["solve"]["plain"] = rec
            log(f"[bench] solve plain: {rec}")
            del W
            torch.cuda.empty_cache()
            names = ["plain"] + names
        t = {n: out["solve"][n].get("seconds_to_tol") for n in names}
        if t.get("plain") and t["plain"] > 0:
            out["solve"]["speedup_vs_plain"] = {n: round(t["plain"] / v, 3) for n, v in t.items()
                                                if v and v > 0 and n != "plain"}

    if args.rdm_random:
        out["rdm_random"] = rdm_random(eng, inst, g)

    if args.precision_report:
        out["precision_report"] = precision_report(eng, args.solve_time)

    if args.batch:
        out["batch"] = batch_bench(args.batch, args.batch_prec)

    if not args.no_cpu_baseline and not args.rename:
        fam = args.workload if args.workload in ("c2", "c3") else "c2"
        out["cpu_baseline"] = cpu_baseline(fam, nnz)

    print(json.dumps(out))


if __name__ == "__main__":
    main()
