"""ctypes front end of the CPU oracle (oracle/sbdo.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by bench.py's
cpu_baseline leg, never by the product package.  See sbdo.h for what each function restates
and how it is pinned.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsbdo.so")
REF_BIN = os.path.join(HERE, "_ref", "sbd_ref")

_lib = None

f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Ints(C.Structure):
    _fields_ = [("L", C.c_int), ("I0", C.c_double), ("I1", C.c_void_p), ("I2", C.c_void_p),
                ("Dm", C.c_void_p), ("Xm", C.c_void_p)]


class Conn(C.Structure):
    _fields_ = [("off_s", C.c_void_p), ("idx_s", C.c_void_p), ("off_d", C.c_void_p),
                ("idx_d", C.c_void_p)]


class Basis(C.Structure):
    _fields_ = [("L", C.c_int), ("na", C.c_int64), ("nb", C.c_int64), ("a", C.c_void_p),
                ("b", C.c_void_p), ("ca", Conn), ("cb", Conn), ("W", C.c_int)]


DW = 4  # uint64 words of an interleaved determinant (SBDO_DW)


def words_of(strings):
    """(uint64 array of the strings, words per string): (n,) arrays hold one word per string,
    (n, W) arrays W words, least significant first."""
    s = np.ascontiguousarray(strings, dtype=np.uint64)
    return s, (1 if s.ndim == 1 else s.shape[1])


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.sbdo_setup_integrals.argtypes = [C.c_int, C.c_int64, f64p, i32p, C.POINTER(C.c_double),
                                           f64p, f64p, f64p, f64p]
        L.sbdo_setup_integrals.restype = C.c_int
        L.sbdo_value.argtypes = [C.POINTER(Ints), C.c_int, C.c_int, C.c_int, C.c_int]
        L.sbdo_value.restype = C.c_double
        L.sbdo_det_from_ab.argtypes = [u64p, u64p, C.c_int, u64p]
        L.sbdo_parity.argtypes = [u64p, C.c_int, C.c_int]
        L.sbdo_parity.restype = C.c_int
        L.sbdo_zero_excite.argtypes = [u64p, C.POINTER(Ints)]
        L.sbdo_zero_excite.restype = C.c_double
        L.sbdo_hij.argtypes = [u64p, u64p, C.POINTER(Ints)]
        L.sbdo_hij.restype = C.c_double
        L.sbdo_count_excitations.argtypes = [u64p, C.c_int64, C.c_int, C.c_int, i64p, i64p]
        L.sbdo_fill_excitations.argtypes = [u64p, C.c_int64, C.c_int, C.c_int, i64p, i64p, i64p, i64p]
        L.sbdo_hdiag.argtypes = [C.POINTER(Basis), C.POINTER(Ints), f64p]
        L.sbdo_mult.argtypes = [C.POINTER(Basis), C.POINTER(Ints), f64p, f64p, f64p]
        L.sbdo_inner.argtypes = [f64p, f64p, C.c_int64]
        L.sbdo_inner.restype = C.c_double
        L.sbdo_syev.argtypes = [C.c_int, f64p, C.c_int, f64p]
        L.sbdo_davidson.argtypes = [C.POINTER(Basis), C.POINTER(Ints), f64p, f64p, C.c_int,
                                    C.c_int, C.c_double, C.c_double, f64p, f64p, C.c_int]
        L.sbdo_davidson.restype = C.c_int
        L.sbdo_correlation.argtypes = [C.POINTER(Basis), f64p, f64p, f64p]
        L.sbdo_occupation.argtypes = [C.POINTER(Basis), f64p, f64p]
        L.sbdo_set_threads.argtypes = [C.c_int]
        L.sbdo_get_threads.restype = C.c_int
        _lib = L
    return _lib


def setup_integrals(norb, vals, idx):
    L = norb
    B = L * (L + 1) // 2
    I0 = C.c_double(0.0)
    I1 = np.zeros(4 * L * L)
    I2 = np.zeros(B * (B + 1) // 2)
    Dm = np.zeros(L * L)
    Xm = np.zeros(L * L)
    rc = lib().sbdo_setup_integrals(L, len(vals), np.ascontiguousarray(vals, np.float64),
                                    np.ascontiguousarray(idx, np.int32).reshape(-1), C.byref(I0),
                                    I1, I2, Dm, Xm)
    if rc != 0:
        raise ValueError("bad FCIDump entry")
    return I0.value, I1, I2, Dm, Xm


def excitations(strings, norb):
    s, W = words_of(strings)
    n = s.shape[0]
    s = s.reshape(-1)
    ns = np.zeros(n, np.int64)
    nd = np.zeros(n, np.int64)
    lib().sbdo_count_excitations(s, n, W, norb, ns, nd)
    off_s = np.zeros(n + 1, np.int64)
    off_d = np.zeros(n + 1, np.int64)
    np.cumsum(ns, out=off_s[1:])
    np.cumsum(nd, out=off_d[1:])
    idx_s = np.zeros(max(int(off_s[-1]), 1), np.int64)
    idx_d = np.zeros(max(int(off_d[-1]), 1), np.int64)
    lib().sbdo_fill_excitations(s, n, W, norb, off_s, idx_s, off_d, idx_d)
    return off_s, idx_s[: off_s[-1]], off_d, idx_d[: off_d[-1]]


class System:
    """One selected-basis problem held by the oracle: integrals + strings + excitation lists."""

    def __init__(self, norb, alpha, beta, vals, idx):
        self.L = int(norb)
        self.alpha, self.W = words_of(alpha)
        self.beta, wb = words_of(beta)
        if wb != self.W:
            raise ValueError("alpha and beta strings must have the same word count")
        self.I0, self.I1, self.I2, self.Dm, self.Xm = setup_integrals(self.L, vals, idx)
        self.conn_a = excitations(self.alpha, self.L)
        self.conn_b = (self.conn_a if self.beta is self.alpha or np.array_equal(self.alpha, self.beta)
                       else excitations(self.beta, self.L))
        self._ints = Ints(self.L, self.I0, self.I1.ctypes.data, self.I2.ctypes.data,
                          self.Dm.ctypes.data, self.Xm.ctypes.data)
        ca = Conn(*[x.ctypes.data if x.size else np.zeros(1, np.int64).ctypes.data
                    for x in self.conn_a])
        cb = Conn(*[x.ctypes.data if x.size else np.zeros(1, np.int64).ctypes.data
                    for x in self.conn_b])
        self._keep = [np.zeros(1, np.int64) for _ in range(8)]
        self._basis = Basis(self.L, self.alpha.shape[0], self.beta.shape[0], self.alpha.ctypes.data,
                            self.beta.ctypes.data, ca, cb, self.W)

    @classmethod
    def from_instance(cls, inst):
        vals, idx = inst.ints.fcidump_entries()
        return cls(inst.norb, inst.alpha, inst.beta, vals, idx)

    @property
    def ndet(self):
        return self.alpha.shape[0] * self.beta.shape[0]

    def value(self, i, j, k, l):
        return lib().sbdo_value(C.byref(self._ints), i, j, k, l)

    def det(self, a, b):
        """Interleaved determinant (DW words) of an alpha and a beta string (an int, or W words)."""
        d = np.zeros(DW, np.uint64)
        nw = (self.L + 63) // 64
        lib().sbdo_det_from_ab(string_words(a, nw), string_words(b, nw), self.L, d)
        return d

    def hij(self, bra, ket):
        return lib().sbdo_hij(np.ascontiguousarray(bra, np.uint64),
                              np.ascontiguousarray(ket, np.uint64), C.byref(self._ints))

    def hdiag(self):
        out = np.zeros(self.ndet)
        lib().sbdo_hdiag(C.byref(self._basis), C.byref(self._ints), out)
        return out

    def mult(self, x, hii=None, out=None):
        hii = self.hdiag() if hii is None else hii
        y = np.zeros(self.ndet) if out is None else out
        lib().sbdo_mult(C.byref(self._basis), C.byref(self._ints), np.ascontiguousarray(hii),
                        np.ascontiguousarray(x, np.float64), y)
        return y

    def davidson(self, max_it=100, num_block=10, eps=1e-8, max_time=86400.0, W=None, hii=None):
        hii = self.hdiag() if hii is None else hii
        if W is None:
            W = np.zeros(self.ndet)
            W[0] = 1.0
        W = np.ascontiguousarray(W, np.float64).copy()
        cap = max(1, max_it * num_block)
        trace = np.zeros(cap)
        resid = np.zeros(cap)
        n = lib().sbdo_davidson(C.byref(self._basis), C.byref(self._ints), hii, W, max_it,
                                num_block, eps, max_time, trace, resid, cap)
        sigma = self.mult(W, hii)
        energy = float(np.dot(W, sigma))
        return dict(W=W, energy=energy, trace=trace[:n], resid=resid[:n], steps=n)

    def correlation(self, W):
        L = self.L
        one = np.zeros(2 * L * L)
        two = np.zeros(4 * L ** 4)
        lib().sbdo_correlation(C.byref(self._basis), np.ascontiguousarray(W, np.float64), one, two)
        return one.reshape(2, L * L), two.reshape(4, L ** 4)

    def occupation(self, W):
        d = np.zeros(2 * self.L)
        lib().sbdo_occupation(C.byref(self._basis), np.ascontiguousarray(W, np.float64), d)
        return d


def string_words(x, nw):
    """One string as nw uint64 words, least significant first (from an int of any width, or an
    array of words, padded with zero words)."""
    if isinstance(x, (int, np.integer)):
        x = int(x)
        return np.array([(x >> (64 * k)) & 0xFFFFFFFFFFFFFFFF for k in range(nw)], np.uint64)
    w = np.ascontiguousarray(x, np.uint64).reshape(-1)
    return np.concatenate([w, np.zeros(max(0, nw - w.size), np.uint64)])


def set_threads(n):
    lib().sbdo_set_threads(int(n))


def get_threads():
    return lib().sbdo_get_threads()


# ------------------------------------------------------------- compiled reference driver

def write_instance(path, norb, nel_a, nel_b, alpha, beta, vals, idx):
    alpha = np.ascontiguousarray(alpha, np.uint64)
    beta = np.ascontiguousarray(beta, np.uint64)
    vals = np.ascontiguousarray(vals, np.float64)
    idx = np.ascontiguousarray(idx, np.int32).reshape(-1)
    # version 2: (n, words) string arrays, least significant word first
    ver, words = (1, 0) if alpha.ndim == 1 else (2, alpha.shape[1])
    with open(path, "wb") as f:
        f.write(b"SBDI")
        f.write(struct.pack("<I4i3q", ver, norb, nel_a, nel_b, words, alpha.shape[0], beta.shape[0], vals.size))
        f.write(alpha.tobytes())
        f.write(beta.tobytes())
        f.write(vals.tobytes())
        f.write(idx.tobytes())


def have_reference():
    return os.path.exists(REF_BIN)


def run_reference(inst, op, workdir=None, vec=None, threads=None, extra=()):
    """Run the compiled reference driver on an Instance; returns (json line or None, outprefix)."""
    wd = workdir or tempfile.mkdtemp(prefix="sbdref_")
    ipath = os.path.join(wd, "inst.bin")
    vals, idx = inst.ints.fcidump_entries()
    write_instance(ipath, inst.norb, inst.n_alpha_el, inst.n_beta_el, inst.alpha, inst.beta,
                   vals, idx)
    args = [REF_BIN, "--in", ipath, "--out", os.path.join(wd, "ref"), "--op", op]
    if vec is not None:
        vpath = os.path.join(wd, "vec.f64")
        np.ascontiguousarray(vec, np.float64).tofile(vpath)
        args += ["--vec", vpath]
    args += list(extra)
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    out = subprocess.run(args, check=True, capture_output=True, text=True, env=env).stdout
    line = None
    for ln in out.splitlines():
        if ln.startswith("{"):
            line = json.loads(ln)
    return line, os.path.join(wd, "ref")
