"""ctypes access to the C oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it is loaded here, by __graft_entry__.smoke
and by bench.py's CPU-baseline leg only, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)
f64p = C.POINTER(C.c_double)


class OrcCase(C.Structure):
    _fields_ = [("N", C.c_int), ("F", C.c_int), ("M", C.c_int), ("nlod", C.c_int),
                ("mother", i32p), ("father", i32p), ("typed", i32p), ("geno", u8p),
                ("disease_prob", f64p), ("geno_prob", f64p), ("legal", u8p),
                ("theta", f64p), ("partial_theta", f64p), ("minor", f64p),
                ("n_ops", C.c_int), ("op_type", i32p), ("op_peel", i32p),
                ("cut_off", i32p), ("cut", i32p), ("prev_off", i32p), ("prev", i32p),
                ("child_off", i32p), ("child", i32p)]


def _load():
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"])
    lib = C.CDLL(LIB)
    P = C.POINTER(OrcCase)
    lib.orc_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.orc_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
    lib.orc_uniform.restype = C.c_double
    lib.orc_trait_prob.argtypes = [P]
    lib.orc_trait_prob.restype = C.c_double
    lib.orc_trait_likelihood.argtypes = [P, i32p, C.c_int, C.c_int]
    lib.orc_trait_likelihood.restype = C.c_double
    lib.orc_lod_terms.argtypes = [P, i32p, C.c_int, f64p]
    lib.orc_lod_terms_scaled.argtypes = [P, i32p, C.c_int, f64p]
    lib.orc_trait_likelihood_scaled.argtypes = [P, i32p, C.c_int, C.c_int, C.POINTER(C.c_int)]
    lib.orc_trait_likelihood_scaled.restype = C.c_double
    lib.orc_trait_grad.argtypes = [P, i32p, C.c_int, C.c_int, f64p, f64p]
    lib.orc_log_sum.argtypes = [C.c_double, C.c_double]
    lib.orc_log_sum.restype = C.c_double
    lib.orc_lod_accumulate.argtypes = [P, i32p, C.c_int, f64p]
    lib.orc_lstep.argtypes = [P, i32p, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(C.c_int)]
    lib.orc_impute.argtypes = [P, i32p, C.c_uint64, C.c_uint32, C.c_int, f64p]
    lib.orc_impute_from.argtypes = [P, i32p, C.c_uint64, C.c_uint32, C.c_int, f64p]
    lib.orc_fag.argtypes = [P, i32p, C.c_int]
    lib.orc_locus_likelihood.argtypes = [P, i32p, C.c_int, C.c_int, C.c_int]
    lib.orc_locus_likelihood.restype = C.c_double
    lib.orc_fag.restype = C.c_double
    lib.orc_msweep.argtypes = [P, i32p, i32p, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32]
    lib.orc_meiosis_list.argtypes = [P, i32p]
    lib.orc_chain.argtypes = [P, i32p, C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_double, f64p,
                              C.POINTER(C.c_int), i32p]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """Binds one swiftlink_amd.case.Case to the C oracle."""

    def __init__(self, case):
        self.case = case
        c = case
        self._keep = [np.ascontiguousarray(x) for x in (
            c.mother.astype(np.int32), c.father.astype(np.int32), c.typed.astype(np.int32),
            c.geno.astype(np.uint8), c.disease_prob.astype(np.float64), c.geno_prob.astype(np.float64),
            c.legal.astype(np.uint8), c.theta.astype(np.float64), c.partial_theta.astype(np.float64),
            c.minor.astype(np.float64), c.ops.type.astype(np.int32), c.ops.peel.astype(np.int32),
            c.ops.cut_off.astype(np.int32), c.ops.cut.astype(np.int32) if c.ops.cut.size else np.zeros(1, np.int32),
            c.ops.prev_off.astype(np.int32), c.ops.prev.astype(np.int32) if c.ops.prev.size else np.zeros(1, np.int32),
            c.ops.child_off.astype(np.int32),
            c.ops.child.astype(np.int32) if c.ops.child.size else np.zeros(1, np.int32))]
        k = self._keep
        self.s = OrcCase(c.N, c.F, c.M, c.nlod, _p(k[0], i32p), _p(k[1], i32p), _p(k[2], i32p), _p(k[3], u8p),
                         _p(k[4], f64p), _p(k[5], f64p), _p(k[6], u8p), _p(k[7], f64p), _p(k[8], f64p),
                         _p(k[9], f64p), c.ops.n, _p(k[10], i32p), _p(k[11], i32p), _p(k[12], i32p),
                         _p(k[13], i32p), _p(k[14], i32p), _p(k[15], i32p), _p(k[16], i32p), _p(k[17], i32p))
        self.ref = C.byref(self.s)

    def trait_prob(self) -> float:
        return lib().orc_trait_prob(self.ref)

    def trait_likelihood(self, dg, locus, offset) -> float:
        d = None if dg is None else np.ascontiguousarray(dg, np.int32)
        return lib().orc_trait_likelihood(self.ref, None if d is None else _p(d, i32p), locus, offset)

    def trait_grad(self, dg, locus, offset, jac) -> np.ndarray:
        """[L, dL/d(pen0, pen1, pen2, freq, theta_L, theta_R)] by forward mode
        (P(T) when dg is None); jac[N][4][4] = d disease_prob / d(pen, freq)."""
        out = np.zeros(7)
        j = np.ascontiguousarray(jac, np.float64)
        d = None if dg is None else np.ascontiguousarray(dg, np.int32)
        rc = lib().orc_trait_grad(self.ref, None if d is None else _p(d, i32p), locus, offset, _p(j, f64p),
                                  _p(out, f64p))
        if rc:
            raise MemoryError("orc_trait_grad")
        return out

    def lod_terms(self, dg, locus) -> np.ndarray:
        out = np.zeros(self.case.nlod)
        d = np.ascontiguousarray(dg, np.int32)
        rc = lib().orc_lod_terms(self.ref, _p(d, i32p), locus, _p(out, f64p))
        if rc:
            raise RuntimeError("zero likelihood")
        return out

    def trait_likelihood_scaled(self, dg, locus, offset):
        """(rescaled L, e): the trait peel with power-of-two table scaling,
        L = rescaled L * 2^e."""
        d = None if dg is None else np.ascontiguousarray(dg, np.int32)
        e = C.c_int(0)
        v = lib().orc_trait_likelihood_scaled(self.ref, None if d is None else _p(d, i32p), locus, offset,
                                              C.byref(e))
        return v, e.value

    def lod_terms_scaled(self, dg, locus) -> np.ndarray:
        """lod_terms through the rescaled peel (no FP64 underflow)."""
        out = np.zeros(self.case.nlod)
        d = np.ascontiguousarray(dg, np.int32)
        rc = lib().orc_lod_terms_scaled(self.ref, _p(d, i32p), locus, _p(out, f64p))
        if rc:
            raise RuntimeError("zero likelihood")
        return out

    def lod_accumulate(self, graphs) -> np.ndarray:
        g = np.ascontiguousarray(graphs, np.int32)
        G = g.shape[0]
        acc = np.zeros((self.case.M - 1) * self.case.nlod)
        rc = lib().orc_lod_accumulate(self.ref, _p(g, i32p), G, _p(acc, f64p))
        if rc:
            raise RuntimeError("zero likelihood")
        return acc

    def lstep(self, dg, locus, seed, chain, sweep):
        d = np.ascontiguousarray(dg, np.int32).copy()
        nd = C.c_int(0)
        rc = lib().orc_lstep(self.ref, _p(d, i32p), locus, seed, chain, sweep, C.byref(nd))
        if rc:
            raise RuntimeError("zero likelihood")
        return d, nd.value

    def impute(self, seed, chain, by_locus=False):
        """One sequential-imputation replicate: (graph [M][2N], weight)."""
        d = np.zeros((self.case.M, 2 * self.case.N), np.int32)
        w = np.zeros(1)
        rc = lib().orc_impute(self.ref, _p(d, i32p), seed, chain, 1 if by_locus else 0, _p(w, f64p))
        if rc:
            raise RuntimeError("zero likelihood")
        return d, float(w[0])

    def impute_from(self, seed, chain, start):
        d = np.zeros((self.case.M, 2 * self.case.N), np.int32)
        w = np.zeros(1)
        if lib().orc_impute_from(self.ref, _p(d, i32p), seed, chain, start, _p(w, f64p)):
            raise RuntimeError("zero likelihood")
        return d, float(w[0])

    def locus_likelihood(self, dg, locus, use_left=True, use_right=True) -> float:
        d = np.ascontiguousarray(dg, np.int32)
        return lib().orc_locus_likelihood(self.ref, _p(d, i32p), locus, int(use_left), int(use_right))

    def impute_weight(self, dg, start) -> float:
        """The weight of a replicate that started at `start` and ended at dg:
        its locus likelihoods recomputed along the same walk."""
        M = self.case.M
        w = np.log(self.locus_likelihood(dg, start, False, False))
        for l in range(start - 1, -1, -1):
            w += np.log(self.locus_likelihood(dg, l, False, True))
        for l in range(start + 1, M):
            w += np.log(self.locus_likelihood(dg, l, True, False))
        return float(w)

    def fag(self, dg, locus) -> float:
        d = np.ascontiguousarray(dg, np.int32)
        return lib().orc_fag(self.ref, _p(d, i32p), locus)

    def msweep(self, dg, order, seed, chain, sweep):
        d = np.ascontiguousarray(dg, np.int32).copy()
        o = np.ascontiguousarray(order, np.int32)
        rc = lib().orc_msweep(self.ref, _p(d, i32p), _p(o, i32p), len(o), seed, chain, sweep)
        if rc:
            raise RuntimeError("illegal graph")
        return d

    def meioses(self) -> np.ndarray:
        out = np.zeros(2 * (self.case.N - self.case.F) + 1, np.int32)
        n = lib().orc_meiosis_list(self.ref, _p(out, i32p))
        return out[:n]

    def chain(self, dg, seed, chain, sweeps, burnin, period, lprob):
        d = np.ascontiguousarray(dg, np.int32).copy()
        acc = np.zeros((self.case.M - 1) * self.case.nlod)
        cnt = C.c_int(0)
        kinds = np.zeros(max(sweeps, 1), np.int32)
        rc = lib().orc_chain(self.ref, _p(d, i32p), seed, chain, sweeps, burnin, period, lprob, _p(acc, f64p),
                             C.byref(cnt), _p(kinds, i32p))
        if rc:
            raise RuntimeError("chain failed")
        return d, acc, cnt.value, kinds[:sweeps]


def uniform(seed, chain, sweep, step, draw) -> float:
    return lib().orc_uniform(seed, chain, sweep, step, draw)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)
