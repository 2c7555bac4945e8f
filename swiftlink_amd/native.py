"""ctypes binding of libswiftlink_amd.so (the C ABI in include/swiftlink_amd.h).

There is no fallback: if the library is missing or a call fails, this module
raises.  The library is built in-tree by `make -C swiftlink_amd` (or
__graft_entry__.build()).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from .case import Case

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libswiftlink_amd.so")

SL_FP64 = 0
SL_FP32_RESCALED = 1
SL_MSAMPLER_AUTO, SL_MSAMPLER_SPECULATIVE, SL_MSAMPLER_TWO_LAUNCH = 0, 1, 2
SL_MSAMPLER_LDS_GRAPHS = 4  # or-ed into a mode: founder-allele graphs built in LDS
SL_LOD_AUTO, SL_LOD_WORKGROUP, SL_LOD_LANES = 0, 1, 2
SL_KERNEL_LOD, SL_KERNEL_LSAMPLER, SL_KERNEL_MSAMPLER, SL_KERNEL_MSAMPLER_FB = 0, 1, 2, 3
SL_OK, SL_E_INVALID, SL_E_UNSUPPORTED, SL_E_DEVICE = 0, -1, -2, -3
SL_E_ZERO_LIKELIHOOD, SL_E_ILLEGAL_GRAPH, SL_E_CAPACITY = -4, -5, -6

i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
f64p = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)


class PedigreeDesc(C.Structure):
    _fields_ = [("n_people", C.c_int32), ("n_founders", C.c_int32), ("n_markers", C.c_int32),
                ("mother", i32p), ("father", i32p), ("typed", i32p), ("genotypes", u8p),
                ("disease_prob", f64p), ("geno_prob", f64p), ("legal", u8p)]


class MapDesc(C.Structure):
    _fields_ = [("theta", f64p), ("partial_theta", f64p), ("minor_freq", f64p), ("lods_per_interval", C.c_int32)]


class PeelDesc(C.Structure):
    _fields_ = [("n_ops", C.c_int32), ("type", i32p), ("peel_node", i32p), ("cut_offsets", i32p),
                ("cutset", i32p), ("prev_offsets", i32p), ("prev", i32p), ("child_offsets", i32p),
                ("children", i32p)]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("device", C.c_int32), ("block_threads", C.c_int32),
                ("graphs_per_block", C.c_int32), ("msampler_mode", C.c_int32), ("lod_kernel", C.c_int32)]


class Schedule(C.Structure):
    _fields_ = [("sweeps", C.c_int32), ("burnin", C.c_int32), ("period", C.c_int32), ("lprob", C.c_double),
                ("seed", C.c_uint64), ("first_sweep", C.c_uint32)]


class McmcStats(C.Structure):
    _fields_ = [("l_sweeps", C.c_int64), ("m_sweeps", C.c_int64), ("scored", C.c_int64)]


class TraitModel(C.Structure):
    _fields_ = [("penetrance", C.c_double * 3), ("disease_freq", C.c_double), ("affection", C.POINTER(C.c_int8))]


class SlError(RuntimeError):
    pass


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "sl_create": (C.c_int, [C.POINTER(PedigreeDesc), C.POINTER(MapDesc), C.POINTER(PeelDesc), C.POINTER(Options),
                            C.POINTER(C.c_void_p)]),
    "sl_destroy": (None, [C.c_void_p]),
    "sl_last_error": (C.c_char_p, [C.c_void_p]),
    "sl_trait_prob": (C.c_int, [C.c_void_p, f64p]),
    "sl_lod_reset": (C.c_int, [C.c_void_p]),
    "sl_lod_accumulate": (C.c_int, [C.c_void_p, i32p, C.c_int32]),
    "sl_lod_accumulate_packed": (C.c_int, [C.c_void_p, u32p, C.c_int32]),
    "sl_lod_accumulate_packed_noalloc": (C.c_int, [C.c_void_p, u32p, C.c_int32]),
    "sl_sync": (C.c_int, [C.c_void_p]),
    "sl_lod_results": (C.c_int, [C.c_void_p, f64p, i64p, f64p]),
    "sl_lod_terms": (C.c_int, [C.c_void_p, i32p, C.c_int32, f64p]),
    "sl_profile": (C.c_int, [C.c_void_p, C.c_int32]),
    "sl_profile_read": (C.c_int, [C.c_void_p, C.c_int32, f64p, i64p]),
    "sl_lod_work": (C.c_int, [C.c_void_p, f64p, f64p]),
    "sl_lod_kernel_for": (C.c_int, [C.c_void_p, C.c_int32, i32p]),
    "sl_program_export": (C.c_int, [C.POINTER(PedigreeDesc), C.POINTER(PeelDesc), C.c_int32, i32p, C.c_int64,
                                    i64p]),
    "sl_mcmc_init": (C.c_int, [C.c_void_p, C.c_int32, i32p]),
    "sl_mcmc_get_graphs": (C.c_int, [C.c_void_p, i32p]),
    "sl_mcmc_status": (C.c_int, [C.c_void_p, i32p]),
    "sl_sequential_imputation": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_int32, i32p, f64p, i32p,
                                           C.POINTER(C.c_int32)]),
    "sl_elod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int64, f64p]),
    "sl_person_order": (C.c_int, [C.c_int32, C.POINTER(C.c_uint8), i32p]),
    "sl_peel_sequence": (C.c_int, [C.c_int32, i32p, i32p, C.c_int64, C.c_uint64, C.c_uint32, C.c_int32, i32p,
                                   C.c_int64, i64p, i64p]),
    "sl_lsweep": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_int32]),
    "sl_msweep": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, i32p, C.c_int32]),
    "sl_mcmc_run": (C.c_int, [C.c_void_p, C.POINTER(Schedule), C.POINTER(McmcStats)]),
    "sl_mcmc_likelihood": (C.c_int, [C.c_void_p, f64p]),
    "sl_mcmc_near_edge": (C.c_int, [C.c_void_p, i64p]),
    "sl_mcmc_near_edge_loci": (C.c_int, [C.c_void_p, u32p]),
    "sl_mcmc_trace": (C.c_int, [C.c_void_p, C.c_int32]),
    "sl_mcmc_trace_read": (C.c_int, [C.c_void_p, C.c_int64, i64p, i32p, f64p]),
    "sl_set_options": (C.c_int, [C.c_void_p, C.POINTER(Options)]),
    "sl_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]),
    "sl_batch_run": (C.c_int, [C.c_void_p, C.POINTER(Schedule), C.POINTER(McmcStats)]),
    "sl_batch_destroy": (None, [C.c_void_p]),
    "sl_trait_model_tables": (C.c_int, [C.POINTER(TraitModel), C.c_int32, C.c_int32, f64p, f64p]),
    "sl_set_trait_model": (C.c_int, [C.c_void_p, C.POINTER(TraitModel)]),
    "sl_trait_grad": (C.c_int, [C.c_void_p, i32p, C.c_int32, f64p]),
    "sl_trait_prob_grad": (C.c_int, [C.c_void_p, f64p]),
    "sl_lod_grad": (C.c_int, [C.c_void_p, u32p, C.c_int32, f64p, f64p]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise SlError(f"{LIBPATH} is not built (run `make -C {HERE}` or __graft_entry__.build())")
        # torch ships its own HIP runtime under the file name libamdhip64.so;
        # loading it first makes the library bind to that same copy (same
        # soname), so the process holds one HIP runtime.  Loaded the other way
        # round, the process would open the GPU through two runtimes and
        # torch's own device initialisation fails.
        import torch  # noqa: F401
        l = C.CDLL(LIBPATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def exported_symbols():
    return list(_SIGS)


def _trait_model(penetrance, freq, affection):
    aff = np.ascontiguousarray(affection, np.int8)
    m = TraitModel((C.c_double * 3)(*[float(x) for x in penetrance]), float(freq),
                   aff.ctypes.data_as(C.POINTER(C.c_int8)))
    return m, aff


def trait_model_tables(penetrance, freq, affection, n_founders: int):
    """Disease probabilities [N][4] of a model and their derivatives [N][4][4]
    with respect to (penetrance 0..2, disease frequency); host only."""
    m, keep = _trait_model(penetrance, freq, affection)
    N = keep.shape[0]
    dp = np.zeros((N, 4))
    jac = np.zeros((N, 4, 4))
    rc = lib().sl_trait_model_tables(C.byref(m), N, n_founders, _ptr(dp, f64p), _ptr(jac, f64p))
    if rc:
        raise SlError(f"sl_trait_model_tables failed ({rc})")
    return dp, jac


def _ptr(a, t):
    return a.ctypes.data_as(t)


class _Descs:
    """Holds numpy buffers alive while descriptors point at them."""

    def __init__(self, case: Case):
        c = case
        z = lambda a, dt: np.ascontiguousarray(a, dtype=dt) if np.size(a) else np.zeros(1, dt)
        self.bufs = dict(
            mother=z(c.mother, np.int32), father=z(c.father, np.int32), typed=z(c.typed, np.int32),
            geno=z(c.geno, np.uint8), dprob=z(c.disease_prob, np.float64), gprob=z(c.geno_prob, np.float64),
            legal=z(c.legal, np.uint8), theta=z(c.theta, np.float64), ptheta=z(c.partial_theta, np.float64),
            minor=z(c.minor, np.float64), type=z(c.ops.type, np.int32), peel=z(c.ops.peel, np.int32),
            cut_off=z(c.ops.cut_off, np.int32), cut=z(c.ops.cut, np.int32), prev_off=z(c.ops.prev_off, np.int32),
            prev=z(c.ops.prev, np.int32), child_off=z(c.ops.child_off, np.int32), child=z(c.ops.child, np.int32))
        b = self.bufs
        self.ped = PedigreeDesc(c.N, c.F, c.M, _ptr(b["mother"], i32p), _ptr(b["father"], i32p),
                                _ptr(b["typed"], i32p), _ptr(b["geno"], u8p), _ptr(b["dprob"], f64p),
                                _ptr(b["gprob"], f64p), _ptr(b["legal"], u8p))
        self.map = MapDesc(_ptr(b["theta"], f64p), _ptr(b["ptheta"], f64p), _ptr(b["minor"], f64p), c.nlod)
        self.peel = PeelDesc(c.ops.n, _ptr(b["type"], i32p), _ptr(b["peel"], i32p), _ptr(b["cut_off"], i32p),
                             _ptr(b["cut"], i32p), _ptr(b["prev_off"], i32p), _ptr(b["prev"], i32p),
                             _ptr(b["child_off"], i32p), _ptr(b["child"], i32p))


def program_export(case: Case, npos: Optional[int] = None) -> np.ndarray:
    """Compiled LOD program as int32 words (host only, no GPU needed)."""
    d = _Descs(case)
    need = C.c_int64(0)
    npos = case.nlod if npos is None else npos
    rc = lib().sl_program_export(C.byref(d.ped), C.byref(d.peel), npos, None, 0, C.byref(need))
    if rc:
        raise SlError(lib().sl_last_error(None).decode())
    buf = np.zeros(need.value, np.int32)
    rc = lib().sl_program_export(C.byref(d.ped), C.byref(d.peel), npos, _ptr(buf, i32p), buf.size, C.byref(need))
    if rc:
        raise SlError(lib().sl_last_error(None).decode())
    return buf


class Context:
    """One sl_ctx: a pedigree, its map and peel program on one GPU."""

    def __init__(self, case: Case, precision: int = SL_FP64, device: int = 0, block: int = 0, gpw: int = 0,
                 msampler_mode: int = SL_MSAMPLER_AUTO, lod_kernel: int = SL_LOD_AUTO):
        self.case = case
        self._d = _Descs(case)
        opts = Options(precision, device, block, gpw, msampler_mode, lod_kernel)
        h = C.c_void_p()
        rc = lib().sl_create(C.byref(self._d.ped), C.byref(self._d.map), C.byref(self._d.peel), C.byref(opts),
                             C.byref(h))
        if rc:
            raise SlError(f"sl_create failed ({rc}): {lib().sl_last_error(None).decode()}")
        self.h = h
        self.precision, self.device = precision, device
        self.options = {"block": block, "gpw": gpw, "msampler_mode": msampler_mode, "lod_kernel": lod_kernel}

    def set_options(self, block: int = 0, gpw: int = 0, msampler_mode: int = SL_MSAMPLER_AUTO,
                    lod_kernel: int = SL_LOD_AUTO):
        """New launch shapes (sl_set_options); chains and accumulators are kept."""
        opts = Options(self.precision, self.device, block, gpw, msampler_mode, lod_kernel)
        self._check(lib().sl_set_options(self.h, C.byref(opts)), "sl_set_options")
        self.options = {"block": block, "gpw": gpw, "msampler_mode": msampler_mode, "lod_kernel": lod_kernel}

    def _check(self, rc, what):
        if rc:
            raise SlError(f"{what} failed ({rc}): {lib().sl_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            lib().sl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- LOD ----
    def trait_prob(self) -> float:
        v = C.c_double(0)
        self._check(lib().sl_trait_prob(self.h, C.byref(v)), "sl_trait_prob")
        return v.value

    def lod_reset(self):
        self._check(lib().sl_lod_reset(self.h), "sl_lod_reset")

    def lod_accumulate(self, graphs: np.ndarray):
        g = np.ascontiguousarray(graphs, np.int32).reshape(-1, self.case.M, 2 * self.case.N)
        self._check(lib().sl_lod_accumulate(self.h, _ptr(g, i32p), g.shape[0]), "sl_lod_accumulate")

    def lod_accumulate_packed(self, dev_ptr: int, n: int):
        self._check(lib().sl_lod_accumulate_packed(self.h, C.cast(C.c_void_p(dev_ptr), u32p), n),
                    "sl_lod_accumulate_packed")

    def lod_accumulate_packed_noalloc(self, dev_ptr: int, n: int):
        """sl_lod_accumulate_packed_noalloc: the same scoring with no device
        allocation (partials in the library's fixed per-device array)."""
        self._check(lib().sl_lod_accumulate_packed_noalloc(self.h, C.cast(C.c_void_p(dev_ptr), u32p), n),
                    "sl_lod_accumulate_packed_noalloc")

    def sync(self):
        self._check(lib().sl_sync(self.h), "sl_sync")

    def lod_results(self):
        n = (self.case.M - 1) * self.case.nlod
        acc = np.zeros(n)
        lod = np.zeros(n)
        cnt = C.c_int64(0)
        self._check(lib().sl_lod_results(self.h, _ptr(acc, f64p), C.byref(cnt), _ptr(lod, f64p)), "sl_lod_results")
        return acc, cnt.value, lod

    def set_trait_model(self, penetrance, freq, affection=None):
        """Disease model (penetrance[3], disease allele frequency, per-person
        affection 0/1/2, by default the case's)."""
        aff = self.case.affection if affection is None else affection
        m, keep = _trait_model(penetrance, freq, aff)
        self._check(lib().sl_set_trait_model(self.h, C.byref(m)), "sl_set_trait_model")

    def trait_grad(self, graphs: np.ndarray) -> np.ndarray:
        """[G][M-1][n][7]: ln L - R - T, then d ln L / d(pen0, pen1, pen2, freq, theta_L, theta_R)."""
        g = np.ascontiguousarray(graphs, np.int32).reshape(-1, self.case.M, 2 * self.case.N)
        out = np.zeros((g.shape[0], self.case.M - 1, self.case.nlod, 7))
        self._check(lib().sl_trait_grad(self.h, _ptr(g, i32p), g.shape[0], _ptr(out, f64p)), "sl_trait_grad")
        return out

    def trait_prob_grad(self) -> np.ndarray:
        """[ln P(T), d/d pen0, pen1, pen2, freq]."""
        out = np.zeros(5)
        self._check(lib().sl_trait_prob_grad(self.h, _ptr(out, f64p)), "sl_trait_prob_grad")
        return out

    def lod_grad_packed(self, dev_ptr: int, n: int):
        """LOD [M-1][n] and its derivatives [M-1][n][6] over n packed device-resident graphs."""
        lod = np.zeros((self.case.M - 1, self.case.nlod))
        dlod = np.zeros((self.case.M - 1, self.case.nlod, 6))
        self._check(lib().sl_lod_grad(self.h, C.cast(C.c_void_p(dev_ptr), u32p), n, _ptr(lod, f64p),
                                      _ptr(dlod, f64p)), "sl_lod_grad")
        return lod, dlod

    def lod_terms(self, graphs: np.ndarray) -> np.ndarray:
        g = np.ascontiguousarray(graphs, np.int32).reshape(-1, self.case.M, 2 * self.case.N)
        out = np.zeros((g.shape[0], self.case.M - 1, self.case.nlod))
        self._check(lib().sl_lod_terms(self.h, _ptr(g, i32p), g.shape[0], _ptr(out, f64p)), "sl_lod_terms")
        return out

    def profile(self, enable: bool):
        self._check(lib().sl_profile(self.h, 1 if enable else 0), "sl_profile")

    def profile_read(self, kernel: int = SL_KERNEL_LOD):
        ms = C.c_double(0)
        n = C.c_int64(0)
        self._check(lib().sl_profile_read(self.h, kernel, C.byref(ms), C.byref(n)), "sl_profile_read")
        return ms.value, n.value

    # ---- chains ----
    def mcmc_init(self, graphs: np.ndarray):
        g = np.ascontiguousarray(graphs, np.int32).reshape(-1, self.case.M, 2 * self.case.N)
        self.n_chains = g.shape[0]
        self._check(lib().sl_mcmc_init(self.h, g.shape[0], _ptr(g, i32p)), "sl_mcmc_init")

    def mcmc_graphs(self) -> np.ndarray:
        out = np.zeros((self.n_chains, self.case.M, 2 * self.case.N), np.int32)
        self._check(lib().sl_mcmc_get_graphs(self.h, _ptr(out, i32p)), "sl_mcmc_get_graphs")
        return out

    def mcmc_status(self) -> np.ndarray:
        out = np.zeros(self.n_chains, np.int32)
        self._check(lib().sl_mcmc_status(self.h, _ptr(out, i32p)), "sl_mcmc_status")
        return out

    def sequential_imputation(self, seed: int, iterations: int, n_chains: int = 0, all_graphs: bool = False):
        """SequentialImputation::parallel_run on the device: `iterations`
        replicates (0: one locus-by-locus pass).  Returns (best graph [M][2N],
        weights, best index, every replicate's graph or None); with n_chains
        > 0 the chains start from the best graph."""
        R = max(iterations, 1)
        best = np.zeros((self.case.M, 2 * self.case.N), np.int32)
        w = np.zeros(R)
        allg = np.zeros((R, self.case.M, 2 * self.case.N), np.int32) if all_graphs else None
        bi = C.c_int32(-1)
        self._check(lib().sl_sequential_imputation(self.h, seed, iterations, n_chains, _ptr(best, i32p),
                                                   _ptr(w, f64p), _ptr(allg, i32p) if all_graphs else None,
                                                   C.byref(bi)), "sl_sequential_imputation")
        if n_chains > 0:
            self.n_chains = n_chains
        return best, w, bi.value, allg

    def lsweep(self, seed: int, sweep: int, first_parity: int = 0):
        self._check(lib().sl_lsweep(self.h, seed, sweep, first_parity), "sl_lsweep")

    def msweep(self, seed: int, sweep: int, order=None):
        """One M-sweep on every chain; order=None shuffles the non-ignorable meioses per chain."""
        if order is None:
            self._check(lib().sl_msweep(self.h, seed, sweep, None, 0), "sl_msweep")
        else:
            o = np.ascontiguousarray(order, np.int32)
            self._check(lib().sl_msweep(self.h, seed, sweep, _ptr(o, i32p), o.size), "sl_msweep")

    def mcmc_run(self, sweeps: int, burnin: int, period: int, lprob: float, seed: int, first_sweep: int = 0):
        """Runs every chain through the sweep schedule, scoring into the LOD accumulators."""
        sch = Schedule(sweeps, burnin, period, lprob, seed, first_sweep)
        st = McmcStats()
        self._check(lib().sl_mcmc_run(self.h, C.byref(sch), C.byref(st)), "sl_mcmc_run")
        return {"l_sweeps": st.l_sweeps, "m_sweeps": st.m_sweeps, "scored": st.scored}

    def mcmc_likelihood(self) -> np.ndarray:
        out = np.zeros((self.n_chains, self.case.M))
        self._check(lib().sl_mcmc_likelihood(self.h, _ptr(out, f64p)), "sl_mcmc_likelihood")
        return out

    def mcmc_near_edge(self) -> np.ndarray:
        out = np.zeros(self.n_chains, np.int64)
        self._check(lib().sl_mcmc_near_edge(self.h, _ptr(out, i64p)), "sl_mcmc_near_edge")
        return out

    def mcmc_trace(self, enable: bool = True):
        """Records the CODA likelihood of every chain at every scored sweep."""
        self._check(lib().sl_mcmc_trace(self.h, 1 if enable else 0), "sl_mcmc_trace")

    def mcmc_trace_read(self):
        """(iterations [R], loglik [R][C]) of the recorded trace."""
        n = C.c_int64(0)
        self._check(lib().sl_mcmc_trace_read(self.h, 0, C.byref(n), None, None), "sl_mcmc_trace_read")
        it = np.zeros(n.value, np.int32)
        ll = np.zeros((n.value, self.n_chains))
        if n.value:
            self._check(lib().sl_mcmc_trace_read(self.h, n.value, C.byref(n), _ptr(it, i32p), _ptr(ll, f64p)),
                        "sl_mcmc_trace_read")
        return it, ll

    def mcmc_near_edge_loci(self) -> np.ndarray:
        """[C][M] booleans: loci where the L-sampler drew near a CDF edge."""
        nw = (self.case.M + 31) // 32
        bits = np.zeros((self.n_chains, nw), np.uint32)
        self._check(lib().sl_mcmc_near_edge_loci(self.h, _ptr(bits, u32p)), "sl_mcmc_near_edge_loci")
        b = np.unpackbits(bits.view(np.uint8), axis=1, bitorder="little")
        return b[:, :self.case.M].astype(bool)

    def lod_kernel_for(self, n_graphs: int) -> int:
        """SL_LOD_WORKGROUP or SL_LOD_LANES: the kernel that scores n_graphs."""
        k = C.c_int32(0)
        self._check(lib().sl_lod_kernel_for(self.h, n_graphs, C.byref(k)), "sl_lod_kernel_for")
        return k.value

    def lod_work(self):
        f = C.c_double(0)
        b = C.c_double(0)
        self._check(lib().sl_lod_work(self.h, C.byref(f), C.byref(b)), "sl_lod_work")
        return f.value, b.value


def person_order(founder) -> np.ndarray:
    """sl_person_order: file positions of the reference's internal order."""
    f = np.ascontiguousarray(founder, np.uint8)
    out = np.zeros(len(f), np.int32)
    rc = lib().sl_person_order(len(f), f.ctypes.data_as(C.POINTER(C.c_uint8)), _ptr(out, i32p))
    if rc:
        raise SlError(f"sl_person_order failed ({rc})")
    return out


def peel_sequence(mother, father, iterations: int = 1000000, seed: int = 12345, rand_seed: int = 12345,
                  threads: int = 0):
    """sl_peel_sequence: the reference's peel-sequence search on the host.
    Returns (PeelOps, cost)."""
    from .case import PeelOps
    m = np.ascontiguousarray(mother, np.int32)
    f = np.ascontiguousarray(father, np.int32)
    need, cost = C.c_int64(0), C.c_int64(0)
    buf = np.zeros(1 + len(m) * (5 + 3 * 16), np.int32)  # enough unless cutsets exceed 16
    rc = lib().sl_peel_sequence(len(m), _ptr(m, i32p), _ptr(f, i32p), iterations, seed, rand_seed, threads,
                                _ptr(buf, i32p), len(buf), C.byref(need), C.byref(cost))
    if rc == SL_E_CAPACITY:  # search again into a buffer of the size it reported
        buf = np.zeros(need.value, np.int32)
        rc = lib().sl_peel_sequence(len(m), _ptr(m, i32p), _ptr(f, i32p), iterations, seed, rand_seed, threads,
                                    _ptr(buf, i32p), len(buf), C.byref(need), C.byref(cost))
    if rc:
        raise SlError(f"sl_peel_sequence failed ({rc}): {lib().sl_last_error(None).decode()}")
    n, pos = int(buf[0]), 1
    types, peels, cuts, prevs, kids = [], [], [], [], []
    for _ in range(n):
        types.append(int(buf[pos])); peels.append(int(buf[pos + 1])); k = int(buf[pos + 2]); pos += 3
        cuts.append([int(x) for x in buf[pos:pos + k]]); pos += k
        npv = int(buf[pos]); pos += 1
        prevs.append([int(x) for x in buf[pos:pos + npv]]); pos += npv
        nc = int(buf[pos]); pos += 1
        kids.append([int(x) for x in buf[pos:pos + nc]]); pos += nc
    return PeelOps.from_lists(types, peels, cuts, prevs, kids), cost.value


def run_elod(sampler: "Context", scorer: "Context", seed: int, replicates: int) -> float:
    """sl_elod: the ELOD of `replicates` replicates; the scorer keeps the
    accumulators."""
    out = C.c_double(0.0)
    sampler._check(lib().sl_elod(sampler.h, scorer.h, seed, replicates, C.byref(out)), "sl_elod")
    return out.value


class Batch:
    """Contexts (one per pedigree) whose sweeps run as one captured hipGraph
    (sl_batch_*).  Destroy the batch before its contexts, or use it as a
    context manager."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = C.c_void_p()
        rc = lib().sl_batch_create(arr, len(self.contexts), C.byref(h))
        if rc:
            raise SlError(f"sl_batch_create failed ({rc}): {lib().sl_last_error(None).decode()}")
        self.h = h

    def run(self, sweeps: int, burnin: int, period: int, lprob: float, seed: int, first_sweep: int = 0):
        """Same schedule on every context; returns one stats dict per context."""
        sch = Schedule(sweeps, burnin, period, lprob, seed, first_sweep)
        st = (McmcStats * len(self.contexts))()
        rc = lib().sl_batch_run(self.h, C.byref(sch), st)
        if rc:
            msgs = [lib().sl_last_error(c.h).decode() for c in self.contexts]
            raise SlError(f"sl_batch_run failed ({rc}): {[m for m in msgs if m]}")
        return [{"l_sweeps": x.l_sweeps, "m_sweeps": x.m_sweeps, "scored": x.scored} for x in st]

    def close(self):
        if self.h:
            lib().sl_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
