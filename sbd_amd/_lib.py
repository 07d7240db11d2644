"""ctypes binding of libsbdx.so (the C ABI declared in include/sbdx.h).

The library is built in-tree (``python -m sbd_amd.build``); importing this module never
falls back to any other implementation: a missing library is an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsbdx.so")

SBDX_FP64 = 0
SBDX_FP32 = 1

EXPORTS = [
    "sbdx_create", "sbdx_create_words", "sbdx_set_integrals", "sbdx_excitation_counts", "sbdx_get_excitations",
    "sbdx_ndet", "sbdx_hdiag", "sbdx_sigma", "sbdx_sigma_device", "sbdx_sigma_block_device",
    "sbdx_davidson_defaults",
    "sbdx_davidson", "sbdx_davidson_device", "sbdx_rdm", "sbdx_occupation", "sbdx_profile",
    "sbdx_profile_read", "sbdx_destroy", "sbdx_last_error",
    "sbdx_mlp_train_defaults", "sbdx_mlp_record", "sbdx_mlp_loss_grad", "sbdx_mlp_forward",
    "sbdx_mlp_train", "sbdx_mlp_set", "sbdx_mlp_apply", "sbdx_autotune", "sbdx_set_mix_variant", "sbdx_set_fp32_accumulation",
    "sbdx_batch_create", "sbdx_batch_davidson", "sbdx_batch_destroy", "sbdx_store_hamiltonian",
    "sbdx_bandwidth_probe", "sbdx_rdm_device", "sbdx_sigma_device_noalloc",
]
MLP_NPARAM = 12609
MLP_SAMPLE = np.dtype([("alpha", "<u8"), ("beta", "<u8"), ("xh", "<f4"), ("xr", "<f4"),
                       ("u", "<f4"), ("v", "<f4")])
PHASES = ["alpha_rows", "mixed", "transpose", "beta_rows", "add_transposed"]


class DavidsonOpts(C.Structure):
    _fields_ = [("max_iteration", C.c_int), ("num_block", C.c_int), ("eps", C.c_double),
                ("max_time", C.c_double), ("init", C.c_int), ("precision", C.c_int),
                ("nroots", C.c_int), ("preconditioner", C.c_int), ("verbose", C.c_int),
                ("energy_ref", C.c_double), ("energy_tol", C.c_double), ("method", C.c_int),
                ("block", C.c_int), ("stop_at_tol", C.c_int), ("precond_warmup", C.c_int)]


class DavidsonResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("residual", C.c_double), ("steps", C.c_int),
                ("converged", C.c_int), ("seconds", C.c_double), ("seconds_to_tol", C.c_double),
                ("steps_to_tol", C.c_int), ("precond_fallback", C.c_int)]


class MlpTrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_double), ("beta1", C.c_double),
                ("beta2", C.c_double), ("eps", C.c_double), ("weight_decay", C.c_double),
                ("seed", C.c_uint64), ("verbose", C.c_int)]


class SbdxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sbdx error {code}: {msg}")
        self.code = code


_lib = None

f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m sbd_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.sbdx_create.argtypes = [C.c_int, C.c_int64, u64p, C.c_int64, u64p, C.POINTER(vp)]
    L.sbdx_create_words.argtypes = [C.c_int, C.c_int, C.c_int64, u64p, C.c_int64, u64p, C.POINTER(vp)]
    L.sbdx_set_integrals.argtypes = [vp, C.c_double, f64p, f64p, C.c_int64]
    L.sbdx_excitation_counts.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.sbdx_get_excitations.argtypes = [vp, C.c_int, i64p, i64p, i64p, i64p]
    L.sbdx_ndet.argtypes = [vp]
    L.sbdx_ndet.restype = C.c_int64
    L.sbdx_hdiag.argtypes = [vp, f64p]
    L.sbdx_sigma.argtypes = [vp, C.c_int, f64p, f64p]
    L.sbdx_sigma_device.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
    L.sbdx_sigma_block_device.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
    L.sbdx_sigma_device_noalloc.argtypes = [vp, C.c_int, vp, vp, vp]
    L.sbdx_davidson_defaults.argtypes = [C.POINTER(DavidsonOpts)]
    L.sbdx_davidson_defaults.restype = None
    L.sbdx_davidson.argtypes = [vp, C.POINTER(DavidsonOpts), f64p, C.POINTER(DavidsonResult),
                                vp, vp, C.c_int]
    L.sbdx_davidson_device.argtypes = [vp, C.POINTER(DavidsonOpts), vp, C.POINTER(DavidsonResult),
                                       vp, vp, C.c_int]
    L.sbdx_rdm.argtypes = [vp, f64p, f64p, f64p]
    L.sbdx_rdm_device.argtypes = [vp, vp, f64p, f64p]
    L.sbdx_occupation.argtypes = [vp, f64p, f64p]
    L.sbdx_profile.argtypes = [vp, C.c_int]
    L.sbdx_profile_read.argtypes = [vp, f64p, C.c_int, C.POINTER(C.c_int64)]
    L.sbdx_destroy.argtypes = [vp]
    L.sbdx_destroy.restype = None
    L.sbdx_last_error.restype = C.c_char_p
    f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    smp = np.ctypeslib.ndpointer(dtype=MLP_SAMPLE, flags="C_CONTIGUOUS")
    L.sbdx_mlp_train_defaults.argtypes = [C.POINTER(MlpTrainOpts)]
    L.sbdx_mlp_train_defaults.restype = None
    L.sbdx_mlp_record.argtypes = [vp, vp, vp, C.c_double, C.c_uint64, smp, C.c_int64,
                                  C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.sbdx_mlp_loss_grad.argtypes = [smp, C.c_int64, f32p, C.POINTER(C.c_double), f32p]
    L.sbdx_mlp_forward.argtypes = [smp, C.c_int64, f32p, f32p]
    L.sbdx_mlp_train.argtypes = [smp, C.c_int64, C.POINTER(MlpTrainOpts), f32p, vp, C.c_int]
    L.sbdx_mlp_set.argtypes = [vp, vp]
    L.sbdx_mlp_apply.argtypes = [vp, C.c_double, f64p, f64p]
    i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    L.sbdx_autotune.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), i32p, f64p,
                                C.c_int, C.POINTER(C.c_int)]
    L.sbdx_set_mix_variant.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.sbdx_set_fp32_accumulation.argtypes = [vp, C.c_int]
    L.sbdx_batch_create.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(vp)]
    L.sbdx_batch_davidson.argtypes = [vp, C.POINTER(DavidsonOpts), C.POINTER(vp),
                                      C.POINTER(DavidsonResult)]
    L.sbdx_batch_destroy.argtypes = [vp]
    L.sbdx_store_hamiltonian.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
    L.sbdx_batch_destroy.restype = None
    L.sbdx_bandwidth_probe.argtypes = [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    for name in EXPORTS:
        if name not in ("sbdx_ndet", "sbdx_davidson_defaults", "sbdx_destroy", "sbdx_last_error",
                        "sbdx_mlp_train_defaults", "sbdx_batch_destroy"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().sbdx_last_error().decode(errors="replace")
        raise SbdxError(rc, msg)
    return rc


def davidson_defaults(**kw) -> DavidsonOpts:
    o = DavidsonOpts()
    lib().sbdx_davidson_defaults(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def mlp_loss_grad(samples, params):
    """Mean loss and gradient of the learned preconditioner's network over samples (GPU)."""
    s = np.ascontiguousarray(samples, dtype=MLP_SAMPLE)
    p = np.ascontiguousarray(params, dtype=np.float32)
    loss = C.c_double()
    g = np.zeros(MLP_NPARAM, np.float32)
    check(lib().sbdx_mlp_loss_grad(s, s.size, p, C.byref(loss), g))
    return loss.value, g


def mlp_forward(samples, params):
    s = np.ascontiguousarray(samples, dtype=MLP_SAMPLE)
    m = np.zeros(s.size, np.float32)
    check(lib().sbdx_mlp_forward(s, s.size, np.ascontiguousarray(params, dtype=np.float32), m))
    return m


def mlp_train_defaults(**kw) -> MlpTrainOpts:
    o = MlpTrainOpts()
    lib().sbdx_mlp_train_defaults(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def mlp_train(samples, params, **opts):
    """Adam training on the GPU; returns (trained FP32 parameters, per-epoch mean loss)."""
    o = mlp_train_defaults(**opts)
    s = np.ascontiguousarray(samples, dtype=MLP_SAMPLE)
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    trace = np.zeros(max(1, o.epochs))
    check(lib().sbdx_mlp_train(s, s.size, C.byref(o), p, trace.ctypes.data, trace.size))
    return p, trace[:o.epochs].copy()


def bandwidth_probe(kind, nbytes, reps):
    """GB/s of a memory level of the current device (sbdx_bandwidth_probe): kind 0 HBM copy,
    kind 1 L2 -> L1 read of an L2-resident buffer."""
    g = C.c_double()
    check(lib().sbdx_bandwidth_probe(int(kind), int(nbytes), int(reps), C.byref(g)))
    return g.value


class Engine:
    """One selected-basis problem resident on the current GPU (an ``sbdx_ctx``)."""

    def __init__(self, norb, alpha, beta, I0=None, I1=None, I2=None):
        """alpha, beta: uint64 strings, shape (n,) up to 64 orbitals or (n, words) with the least
        significant word first (sbdx_create_words, up to 128 orbitals)."""
        self.norb = int(norb)
        self.alpha = np.ascontiguousarray(alpha, dtype=np.uint64)
        self.beta = np.ascontiguousarray(beta, dtype=np.uint64)
        h = C.c_void_p()
        if self.alpha.ndim == 1 and self.beta.ndim == 1:
            check(lib().sbdx_create(self.norb, self.alpha.size, self.alpha, self.beta.size, self.beta,
                                    C.byref(h)))
        else:
            words = self.alpha.shape[1]
            if self.alpha.ndim != 2 or self.beta.ndim != 2 or self.beta.shape[1] != words:
                raise ValueError("alpha and beta must both be (n, words) arrays")
            check(lib().sbdx_create_words(self.norb, words, self.alpha.shape[0], self.alpha,
                                          self.beta.shape[0], self.beta, C.byref(h)))
        self._h = h
        if I1 is not None:
            self.set_integrals(I0, I1, I2)

    @classmethod
    def from_instance(cls, inst):
        """Build from a synthetic Instance, with the integrals in the reference's storage."""
        L = inst.norb
        I1 = np.zeros((2 * L, 2 * L))
        h = inst.ints.h1
        I1[0::2, 0::2] = h
        I1[1::2, 1::2] = h
        return cls(L, inst.alpha, inst.beta, inst.ints.ecore, I1.ravel(), inst.ints.packed_eri())

    @property
    def handle(self):
        return self._h

    def set_integrals(self, I0, I1, I2):
        I1 = np.ascontiguousarray(I1, dtype=np.float64).ravel()
        I2 = np.ascontiguousarray(I2, dtype=np.float64).ravel()
        check(lib().sbdx_set_integrals(self._h, float(I0), I1, I2, I2.size))

    @property
    def ndet(self):
        return int(lib().sbdx_ndet(self._h))

    def excitation_counts(self, spin):
        ns, nd = C.c_int64(), C.c_int64()
        check(lib().sbdx_excitation_counts(self._h, spin, C.byref(ns), C.byref(nd)))
        return ns.value, nd.value

    def excitations(self, spin):
        n = self.alpha.shape[0] if spin == 0 else self.beta.shape[0]
        ns, nd = self.excitation_counts(spin)
        off_s = np.zeros(n + 1, np.int64)
        off_d = np.zeros(n + 1, np.int64)
        idx_s = np.zeros(max(ns, 1), np.int64)
        idx_d = np.zeros(max(nd, 1), np.int64)
        check(lib().sbdx_get_excitations(self._h, spin, off_s, idx_s, off_d, idx_d))
        return off_s, idx_s[:ns], off_d, idx_d[:nd]

    def hdiag(self):
        out = np.zeros(self.ndet)
        check(lib().sbdx_hdiag(self._h, out))
        return out

    def sigma(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        nvec = x.size // self.ndet
        y = np.zeros_like(x)
        check(lib().sbdx_sigma(self._h, nvec, x, y))
        return y

    def sigma_device(self, prec, nvec, dC, dS, stream=None):
        check(lib().sbdx_sigma_device(self._h, prec, nvec, C.c_void_p(dC), C.c_void_p(dS),
                                      C.c_void_p(stream) if stream else None))

    def sigma_device_noalloc(self, prec, dC, dS, stream=None):
        """sbdx_sigma_device_noalloc: one vector, no allocation, enqueued on `stream`."""
        check(lib().sbdx_sigma_device_noalloc(self._h, prec, C.c_void_p(dC), C.c_void_p(dS),
                                              C.c_void_p(stream) if stream else None))

    def sigma_block_device(self, prec, k, dC, dS, stream=None):
        """Fused sigma of k interleaved vectors (element [det * k + v])."""
        check(lib().sbdx_sigma_block_device(self._h, prec, k, C.c_void_p(dC), C.c_void_p(dS),
                                            C.c_void_p(stream) if stream else None))

    def davidson(self, W=None, trace_cap=4096, **opts):
        o = davidson_defaults(**opts)
        nr = max(1, o.nroots)
        if W is None:
            W = np.zeros(self.ndet * nr)
            o.init = 0
        else:
            W0 = np.ascontiguousarray(W, dtype=np.float64).ravel()
            W = np.zeros(self.ndet * nr)
            W[:W0.size] = W0[:W.size]
            o.init = 1
        res = DavidsonResult()
        trace = np.zeros(trace_cap)
        resid = np.zeros(trace_cap)
        check(lib().sbdx_davidson(self._h, C.byref(o), W, C.byref(res), trace.ctypes.data,
                                  resid.ctypes.data, trace_cap))
        n = min(res.steps, trace_cap)
        return dict(W=W, energy=res.energy, residual=res.residual, steps=res.steps,
                    converged=bool(res.converged), seconds=res.seconds,
                    seconds_to_tol=res.seconds_to_tol, steps_to_tol=res.steps_to_tol,
                    precond_fallback=res.precond_fallback,
                    trace=trace[:n].copy(), resid=resid[:n].copy())

    def davidson_device(self, dW, trace_cap=4096, **opts):
        o = davidson_defaults(**opts)
        res = DavidsonResult()
        trace = np.zeros(trace_cap)
        resid = np.zeros(trace_cap)
        check(lib().sbdx_davidson_device(self._h, C.byref(o), C.c_void_p(dW), C.byref(res),
                                         trace.ctypes.data, resid.ctypes.data, trace_cap))
        n = min(res.steps, trace_cap)
        return dict(energy=res.energy, residual=res.residual, steps=res.steps,
                    converged=bool(res.converged), seconds=res.seconds,
                    seconds_to_tol=res.seconds_to_tol, steps_to_tol=res.steps_to_tol,
                    precond_fallback=res.precond_fallback,
                    trace=trace[:n].copy(), resid=resid[:n].copy())

    def mlp_set(self, params):
        """Network of preconditioner 1 (None removes it)."""
        if params is None:
            check(lib().sbdx_mlp_set(self._h, None))
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == MLP_NPARAM
        check(lib().sbdx_mlp_set(self._h, p.ctypes.data))

    def store_hamiltonian(self, enable=True):
        """Build (or drop) the stored CSR Hamiltonian; returns the stored element count."""
        nnz = C.c_int64()
        check(lib().sbdx_store_hamiltonian(self._h, int(bool(enable)), C.byref(nnz)))
        return nnz.value

    def autotune(self, prec, k, reps=2):
        """Time the mixed-pass variants for (prec, k) and keep the fastest; returns
        (chosen variant, {variant: mean ms}), variant = 1000000 (stages - 1) + 100 * threads
        per workgroup + entries per thread, or for k = 4 and 8 a sub-block width (2 or 4: the
        block built as k / width narrower blocks)."""
        cand = np.zeros(16, np.int32)
        ms = np.zeros(16)
        chosen, nc = C.c_int(), C.c_int()
        check(lib().sbdx_autotune(self._h, prec, k, reps, C.byref(chosen), cand, ms, len(cand), C.byref(nc)))
        return chosen.value, {int(cand[i]): float(ms[i]) for i in range(nc.value)}

    def set_fp32_accumulation(self, mode):
        """0: exact FP32 products accumulated in FP64 (default); 1: compensated FP32 (two-sum)."""
        check(lib().sbdx_set_fp32_accumulation(self._h, int(mode)))

    def set_mix_variant(self, prec, k, variant):
        """Pin the mixed-pass variant for (prec, k) (one of autotune's candidates; 0: default)."""
        check(lib().sbdx_set_mix_variant(self._h, prec, k, int(variant)))

    def mlp_apply(self, theta, r):
        t = np.zeros(self.ndet)
        check(lib().sbdx_mlp_apply(self._h, float(theta), np.ascontiguousarray(r, np.float64), t))
        return t

    def mlp_record(self, dX, dXstar, target, seed=0, cap=None):
        """Training pairs at iterate dX against the converged dXstar (device pointers)."""
        cap = int(cap if cap is not None else max(16, 4 * target))
        out = np.zeros(cap, dtype=MLP_SAMPLE)
        cnt = C.c_int64()
        theta = C.c_double()
        check(lib().sbdx_mlp_record(self._h, C.c_void_p(dX), C.c_void_p(dXstar), float(target),
                                    int(seed), out, cap, C.byref(cnt), C.byref(theta)))
        return out[:min(cnt.value, cap)].copy(), cnt.value, theta.value

    def rdm(self, W):
        L = self.norb
        one = np.zeros(2 * L * L)
        two = np.zeros(4 * L ** 4)
        check(lib().sbdx_rdm(self._h, np.ascontiguousarray(W, np.float64), one, two))
        return one.reshape(2, L * L), two.reshape(4, L ** 4)

    def rdm_device(self, dW):
        """RDMs of a device vector (sbdx_rdm_device): dW is a device pointer to ndet doubles."""
        L = self.norb
        one = np.zeros(2 * L * L)
        two = np.zeros(4 * L ** 4)
        check(lib().sbdx_rdm_device(self._h, C.c_void_p(dW), one, two))
        return one.reshape(2, L * L), two.reshape(4, L ** 4)

    def occupation(self, W):
        d = np.zeros(2 * self.norb)
        check(lib().sbdx_occupation(self._h, np.ascontiguousarray(W, np.float64), d))
        return d

    def profile(self, enable=True):
        check(lib().sbdx_profile(self._h, int(enable)))

    def profile_read(self):
        ms = np.zeros(len(PHASES))
        calls = C.c_int64()
        check(lib().sbdx_profile_read(self._h, ms, ms.size, C.byref(calls)))
        return dict(zip(PHASES, ms.tolist())), calls.value

    def close(self):
        if getattr(self, "_h", None):
            lib().sbdx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several Engines (same device) solved together by the lockstep graph-replayed Davidson."""

    def __init__(self, engines):
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        h = C.c_void_p()
        check(lib().sbdx_batch_create(len(self.engines), arr, C.byref(h)))
        self._h = h

    def davidson(self, want_vectors=False, **opts):
        o = davidson_defaults(**opts)
        cnt = len(self.engines)
        res = (DavidsonResult * cnt)()
        W = [np.zeros(e.ndet) for e in self.engines] if want_vectors else None
        ptrs = (C.c_void_p * cnt)(*[w.ctypes.data for w in W]) if W else None
        check(lib().sbdx_batch_davidson(self._h, C.byref(o), ptrs, res))
        out = []
        for i in range(cnt):
            r = res[i]
            d = dict(energy=r.energy, residual=r.residual, steps=r.steps, converged=bool(r.converged),
                     seconds=r.seconds, seconds_to_tol=r.seconds_to_tol, steps_to_tol=r.steps_to_tol)
            if W:
                d["W"] = W[i]
            out.append(d)
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().sbdx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
