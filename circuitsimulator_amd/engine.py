"""ctypes binding of libcse.so (include/cse.h): the Python face of the gfx950
engine, used by the tests, bench.py and anyone driving the engine from Python.

There is no CPU path here: if libcse.so is missing or no gfx950 GPU is
present, the calls raise.
"""
import atexit
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CSE_LIBCSE", os.path.join(HERE, "libcse.so"))
BACKEND_PATH = os.path.join(HERE, "cse_backend")

NMAX = 64
HIST_BINS = 129 * 8
U64 = ctypes.c_uint64
P = ctypes.POINTER

# every entry point declared in include/cse.h
ABI_SYMBOLS = [
    "cse_create", "cse_destroy", "cse_error", "cse_device_info", "cse_inner_products",
    "cse_prep_states", "cse_random_states", "cse_exact_projector", "cse_sampled_projector",
    "cse_batch_exact", "cse_decompose", "cse_hist_value", "cse_sampling_kept", "cse_projector_norm", "cse_projector_norms",
    "cse_lspace_weights", "cse_batch_sampled", "cse_decompose_many", "cse_exact_projector_noalloc",
    "cse_set_graphs", "cse_set_sweep_kernel",
]
ERR_NODEV = -5


class Projector(ctypes.Structure):
    """struct cse_projector (comms.h:4-11 semantics)"""
    _fields_ = [("nstabs", ctypes.c_int), ("nqubits", ctypes.c_int),
                ("phase", ctypes.c_int * NMAX), ("xs", U64 * NMAX), ("zs", U64 * NMAX)]

    @classmethod
    def make(cls, phases, xs, zs, nqubits):
        p = cls()
        p.nstabs = len(phases)
        p.nqubits = nqubits
        for g, (ph, x, z) in enumerate(zip(phases, xs, zs)):
            p.phase[g] = int(ph) & 3
            p.xs[g] = int(x)
            p.zs[g] = int(z)
        return p


class ExactResult(ctypes.Structure):
    _fields_ = [("pmin", ctypes.c_int), ("hist", P(ctypes.c_int64)),
                ("re", ctypes.c_double), ("im", ctypes.c_double), ("value", ctypes.c_double),
                ("kahan_re", ctypes.c_double), ("kahan_im", ctypes.c_double),
                ("terms", ctypes.c_int64), ("ips", ctypes.c_int64),
                ("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


class SampledResult(ctypes.Structure):
    _fields_ = [("mean", ctypes.c_double), ("mean_kahan", ctypes.c_double),
                ("ips", ctypes.c_int64), ("vanished", ctypes.c_int64),
                ("kernel_ms", ctypes.c_double), ("prep_ms", ctypes.c_double),
                ("zsum", P(ctypes.c_int64)), ("sexp", P(ctypes.c_int32)),
                ("kahan", P(ctypes.c_double)), ("terms", P(ctypes.c_int32))]


class Instance(ctypes.Structure):
    _fields_ = [("P", Projector), ("exact", ctypes.c_int), ("k", ctypes.c_int), ("L", U64 * NMAX)]


# the same layout as a numpy record: arrays of instances are filled with
# array operations and passed to cse_batch_exact by pointer
INSTANCE_DTYPE = np.dtype(Instance)


def instance_array(n):
    """n zeroed struct cse_instance records as a numpy array."""
    return np.zeros(n, dtype=INSTANCE_DTYPE)


class SampledInstance(ctypes.Structure):
    _fields_ = [("P", Projector), ("exact", ctypes.c_int), ("k", ctypes.c_int), ("L", U64 * NMAX),
                ("seed", U64), ("first", ctypes.c_int64), ("count", ctypes.c_int64)]


SAMPLED_INSTANCE_DTYPE = np.dtype(SampledInstance)


def sampled_instance_array(n):
    """n zeroed struct cse_sampled_instance records as a numpy array."""
    return np.zeros(n, dtype=SAMPLED_INSTANCE_DTYPE)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run circuitsimulator_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        L.cse_create.argtypes = [ctypes.c_int, P(ctypes.c_void_p)]
        L.cse_destroy.argtypes = [ctypes.c_void_p]
        L.cse_error.argtypes = [ctypes.c_void_p]
        L.cse_error.restype = ctypes.c_char_p
        L.cse_device_info.argtypes = [ctypes.c_void_p, P(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
        L.cse_set_graphs.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.cse_set_sweep_kernel.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.cse_inner_products.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, P(U64), P(U64),
                                         P(ctypes.c_int32), ctypes.c_int, P(ctypes.c_double)]
        L.cse_prep_states.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, P(U64), ctypes.c_int,
                                      P(U64), ctypes.c_size_t, P(U64)]
        L.cse_random_states.argtypes = [ctypes.c_void_p, ctypes.c_int, U64, ctypes.c_int64, ctypes.c_size_t,
                                        P(Projector), P(U64), P(ctypes.c_int32)]
        L.cse_exact_projector.argtypes = [ctypes.c_void_p, P(Projector), P(U64), ctypes.c_int, ctypes.c_int,
                                          P(ExactResult)]
        L.cse_exact_projector_noalloc.argtypes = [ctypes.c_void_p, P(Projector), P(U64), ctypes.c_int,
                                                  ctypes.c_int, P(ExactResult)]
        L.cse_sampled_projector.argtypes = [ctypes.c_void_p, P(Projector), P(U64), ctypes.c_int, ctypes.c_int,
                                            U64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, P(SampledResult)]
        L.cse_batch_exact.argtypes = [ctypes.c_void_p, P(Instance), ctypes.c_int, ctypes.c_int, P(ExactResult),
                                      P(ctypes.c_int32), P(ctypes.c_double)]
        L.cse_batch_sampled.argtypes = [ctypes.c_void_p, P(SampledInstance), ctypes.c_int, ctypes.c_int,
                                        P(SampledResult), P(ctypes.c_double)]
        L.cse_decompose.argtypes = [ctypes.c_int, U64, P(U64), P(ctypes.c_double), P(ctypes.c_int),
                                    P(ctypes.c_int), ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]
        L.cse_decompose_many.argtypes = [ctypes.c_int, P(U64), ctypes.c_int, P(U64), ctypes.c_int,
                                         P(ctypes.c_double), P(ctypes.c_int), P(ctypes.c_int), ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.cse_hist_value.argtypes = [P(ctypes.c_int64), ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double)]
        L.cse_sampling_kept.argtypes = [ctypes.c_int] * 7
        L.cse_lspace_weights.argtypes = [ctypes.c_void_p, P(U64), ctypes.c_int, P(ctypes.c_int64)]
        L.cse_projector_norm.argtypes = [P(ctypes.c_void_p), ctypes.c_int, P(Projector), P(U64), ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         U64, P(ctypes.c_double)]
        L.cse_projector_norms.argtypes = [P(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, P(Projector), P(U64),
                                          P(U64), ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]
        _lib = L
    return _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(P(ctype))


class EngineError(RuntimeError):
    pass


class Engine:
    """One engine context (one HIP stream) on one gfx950 device."""

    def __init__(self, device=0):
        self._L = lib()
        self.sweep_kernels = [0, 0]
        h = ctypes.c_void_p()
        rc = self._L.cse_create(device, ctypes.byref(h))
        if rc != 0:
            raise EngineError(f"cse_create failed ({rc}): {self._L.cse_error(None).decode()}")
        self.h = h

    def close(self):
        if self.h:
            self._L.cse_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise EngineError(f"engine error {rc}: {self._L.cse_error(self.h).decode()}")

    def set_graphs(self, enable=True):
        """replay batched calls of a repeated shape from captured HIP graphs
        (cse_set_graphs)"""
        self._check(self._L.cse_set_graphs(self.h, 1 if enable else 0))

    def set_sweep_kernel(self, kernel, mode=0):
        """sweep kernel of one mode (cse_set_sweep_kernel): sampled (mode 0):
        0 the engine's choice (2 for calls of |H^t> instances only, else 1),
        1 the eight-wave kernel with the two-column asm path, 2 the general
        one; exact (mode 1): 0 the engine's choice (2 for launches of at
        least 16 slices per resident wave, else 1), 1 seven waves per SIMD,
        2 eight"""
        self._check(self._L.cse_set_sweep_kernel(self.h, int(mode), int(kernel)))
        self.sweep_kernels[int(mode)] = int(kernel)

    @property
    def sweep_kernel(self):
        return self.sweep_kernels[0]

    def device_info(self):
        cus = ctypes.c_int()
        name = ctypes.create_string_buffer(128)
        self._check(self._L.cse_device_info(self.h, ctypes.byref(cus), name, 128))
        return cus.value, name.value.decode()

    def inner_products(self, n, s1, s2, one_launch_per_pair=False):
        """s1, s2: (npairs, 3n+4) uint64 packed states -> (npairs, 3) int32, ms"""
        s1 = np.ascontiguousarray(s1, dtype=np.uint64)
        s2 = np.ascontiguousarray(s2, dtype=np.uint64)
        npairs = s1.shape[0]
        out = np.zeros((npairs, 3), dtype=np.int32)
        ms = ctypes.c_double()
        self._check(self._L.cse_inner_products(self.h, n, npairs, _ptr(s1, U64), _ptr(s2, U64),
                                               _ptr(out, ctypes.c_int32), 1 if one_launch_per_pair else 0,
                                               ctypes.byref(ms)))
        return out, ms.value

    def prep_states(self, exact, t, L, idx):
        L = np.ascontiguousarray(np.array(L if len(L) else [0], dtype=np.uint64))
        idx = np.ascontiguousarray(np.array(idx, dtype=np.uint64))
        out = np.zeros((len(idx), 3 * t + 4), dtype=np.uint64)
        self._check(self._L.cse_prep_states(self.h, int(exact), t, _ptr(L, U64), 0 if exact else len(L),
                                            _ptr(idx, U64), len(idx), _ptr(out, U64)))
        return out

    def random_states(self, n, seed, first, count, proj=None, want_states=True):
        out = np.zeros((count, 3 * n + 4), dtype=np.uint64) if want_states else None
        sexp = np.zeros(count, dtype=np.int32)
        self._check(self._L.cse_random_states(self.h, n, seed, first, count,
                                              ctypes.byref(proj) if proj is not None else None,
                                              _ptr(out, U64) if out is not None else None,
                                              _ptr(sexp, ctypes.c_int32)))
        return out, sexp

    def exact_projector(self, proj, L, k, exact, want_hist=False):
        L = np.ascontiguousarray(np.array(list(L) + [0] * (NMAX - len(L)), dtype=np.uint64))
        r = ExactResult()
        hist = None
        if want_hist:
            hist = np.zeros(HIST_BINS, dtype=np.int64)
            r.hist = _ptr(hist, ctypes.c_int64)
        self._check(self._L.cse_exact_projector(self.h, ctypes.byref(proj), _ptr(L, U64), k, int(exact),
                                                ctypes.byref(r)))
        return r, hist

    def exact_projector_noalloc(self, proj, L, k, exact, want_hist=False):
        """exact_projector through cse_exact_projector_noalloc: the same
        result, computed in the library's static device array (chi <= 2^13)."""
        L = np.ascontiguousarray(np.array(list(L) + [0] * (NMAX - len(L)), dtype=np.uint64))
        r = ExactResult()
        hist = None
        if want_hist:
            hist = np.zeros(HIST_BINS, dtype=np.int64)
            r.hist = _ptr(hist, ctypes.c_int64)
        self._check(self._L.cse_exact_projector_noalloc(self.h, ctypes.byref(proj), _ptr(L, U64), k, int(exact),
                                                        ctypes.byref(r)))
        return r, hist

    def sampled_projector(self, proj, L, k, exact, seed, first, count, chunk=0, want_terms=False,
                          want_samples=True):
        L = np.ascontiguousarray(np.array(list(L) + [0] * (NMAX - len(L)), dtype=np.uint64))
        t = proj.nqubits
        chi = (1 << ((t + 1) // 2)) if exact else (1 << k)
        r = SampledResult()
        extra = {}
        if want_samples:
            extra["zsum"] = np.zeros((count, 4), dtype=np.int64)
            extra["sexp"] = np.zeros(count, dtype=np.int32)
            extra["kahan"] = np.zeros((count, 2), dtype=np.float64)
            r.zsum = _ptr(extra["zsum"], ctypes.c_int64)
            r.sexp = _ptr(extra["sexp"], ctypes.c_int32)
            r.kahan = _ptr(extra["kahan"], ctypes.c_double)
        if want_terms:
            extra["terms"] = np.zeros((count, chi), dtype=np.int32)
            r.terms = _ptr(extra["terms"], ctypes.c_int32)
        self._check(self._L.cse_sampled_projector(self.h, ctypes.byref(proj), _ptr(L, U64), k, int(exact),
                                                  seed, first, count, chunk, ctypes.byref(r)))
        return r, extra

    def lspace_weights(self, L, k):
        """hist[w] = number of the 2^k combinations of the rows of L with
        Hamming weight w (65 entries)."""
        Larr = (U64 * NMAX)(*(list(L) + [0] * (NMAX - len(L))))
        hist = np.zeros(65, dtype=np.int64)
        self._check(self._L.cse_lspace_weights(self.h, Larr, int(k), _ptr(hist, ctypes.c_int64)))
        return hist

    def batch_exact(self, instances, chunk=0, want_terms=False, nterms=0):
        """instances: a list of Instance or an instance_array.  chunk: the
        claim cap in j (0: the engine's default; "auto": the autotuner's
        choice of kernel and cap for this call)."""
        res, terms, ms = self._batch_exact(instances, chunk, want_terms, nterms)
        return list(res), terms, ms

    def batch_exact_values(self, instances, chunk=0):
        """(values, kernel ms) of batch_exact: the exact FP64 value of every
        instance as one array, read from the results in place."""
        res, _, ms = self._batch_exact(instances, chunk, False, 0)
        n = len(res)
        vals = np.ndarray(shape=(n,), dtype=np.float64, buffer=res, offset=ExactResult.value.offset,
                          strides=(ctypes.sizeof(ExactResult),)).copy() if n else np.zeros(0)
        return vals, ms

    def _batch_exact(self, instances, chunk, want_terms, nterms):
        kernel = None
        if isinstance(chunk, str):
            if chunk != "auto":
                raise ValueError(f"chunk must be an int or 'auto', not {chunk!r}")
            from .autotune import tune_exact
            kernel, chunk = tune_exact(self, instances)
        if isinstance(instances, np.ndarray):
            if instances.dtype != INSTANCE_DTYPE or not instances.flags["C_CONTIGUOUS"]:
                raise ValueError("instances must be a contiguous instance_array")
            arr = instances.ctypes.data_as(P(Instance))
        else:
            arr = (Instance * len(instances))(*instances)
        res = (ExactResult * len(instances))()
        terms = np.zeros(nterms, dtype=np.int32) if want_terms else None
        ms = ctypes.c_double()
        keep = self.sweep_kernels[1]
        if kernel is not None and kernel != keep:
            self.set_sweep_kernel(kernel, mode=1)
        try:
            self._check(self._L.cse_batch_exact(self.h, arr, len(instances), chunk, res,
                                                _ptr(terms, ctypes.c_int32) if want_terms else None,
                                                ctypes.byref(ms)))
        finally:
            if kernel is not None and kernel != keep:
                self.set_sweep_kernel(keep, mode=1)
        return res, terms, ms.value

    def batch_sampled(self, instances, chunk=0, want_samples=False):
        """instances: a contiguous sampled_instance_array -> (means, results,
        kernel ms, samples); means[i] is cse_sampled_projector's mean for
        instance i; samples[i] (want_samples) its per-sample zsum and sexp.
        chunk: the item size in j (0: the engine's default; "auto": the
        autotuner's choice of kernel and item size for this call)."""
        if instances.dtype != SAMPLED_INSTANCE_DTYPE or not instances.flags["C_CONTIGUOUS"]:
            raise ValueError("instances must be a contiguous sampled_instance_array")
        kernel = None
        if isinstance(chunk, str):
            if chunk != "auto":
                raise ValueError(f"chunk must be an int or 'auto', not {chunk!r}")
            from .autotune import tune_sampled
            kernel, chunk = tune_sampled(self, instances)
        n = len(instances)
        res = (SampledResult * n)()
        samples = []
        if want_samples:
            for i in range(n):
                cnt = int(instances["count"][i])
                extra = {"zsum": np.zeros((cnt, 4), dtype=np.int64), "sexp": np.zeros(cnt, dtype=np.int32)}
                res[i].zsum = _ptr(extra["zsum"], ctypes.c_int64)
                res[i].sexp = _ptr(extra["sexp"], ctypes.c_int32)
                samples.append(extra)
        ms = ctypes.c_double()
        keep = self.sweep_kernel
        if kernel is not None and kernel != keep:
            self.set_sweep_kernel(kernel)
        try:
            self._check(self._L.cse_batch_sampled(self.h, instances.ctypes.data_as(P(SampledInstance)), n, chunk,
                                                  res, ctypes.byref(ms)))
        finally:
            if kernel is not None and kernel != keep:
                self.set_sweep_kernel(keep)
        return np.array([r.mean for r in res]), list(res), ms.value, samples


# one engine context per (device, host thread) for callers that evaluate many
# batches (sampleShots): device buffers and pinned staging stay allocated
# between calls; destroyed at exit
_shared = {}


def shared_engine(device=0):
    """The process's Engine for this device and host thread, created on first
    use (a context serves one host thread at a time)."""
    key = (int(device), threading.get_ident())
    e = _shared.get(key)
    if e is None or not e.h:
        e = _shared[key] = Engine(device)
        # its batches repeat their shape round after round: replay them from
        # captured graphs unless CSE_GRAPHS=0
        e.set_graphs(os.environ.get("CSE_GRAPHS", "1") != "0")
    return e


def _close_shared():
    for e in _shared.values():
        e.close()
    _shared.clear()


atexit.register(_close_shared)


def decompose(t, seed, exact=0, k=0, fidbound=1e-5, rank=0, fidelity=0, forceL=0, verbose=0, quiet=1):
    L = (U64 * NMAX)()
    norm = ctypes.c_double()
    ex = ctypes.c_int(exact)
    kk = ctypes.c_int(k)
    rc = lib().cse_decompose(t, seed, L, ctypes.byref(norm), ctypes.byref(ex), ctypes.byref(kk), fidbound,
                             rank, fidelity, forceL, verbose, quiet)
    if rc != 0:
        raise EngineError(f"cse_decompose failed ({rc})")
    return [int(L[i]) for i in range(kk.value if not ex.value else 0)], norm.value, ex.value, kk.value


def sampling_kept(noapprox, forceSample, exact, t, k, samples, bins):
    """The sampling-disable rule (probability.c:162-174)."""
    return bool(lib().cse_sampling_kept(int(bool(noapprox)), int(bool(forceSample)), int(bool(exact)), int(t),
                                        int(k), int(samples), int(bins)))


# one engine context per device for the projector norms, created by the
# library on the first call that needs the GPU and kept for the process
_norm_ctx = {}


def _destroy_norm_contexts():
    for h in _norm_ctx.values():
        if h:
            lib().cse_destroy(h)
    _norm_ctx.clear()


atexit.register(_destroy_norm_contexts)


def projector_norm(proj, L, k, exact, norm, sampled, samples=0, bins=1, seed=0, device=0):
    """||P|H^t>||^2 as the backend computes it (cse_projector_norm):
    exact, or the median over bins of sampled means."""
    h = _norm_ctx.setdefault(device, ctypes.c_void_p())
    Larr = (U64 * NMAX)(*(list(L) + [0] * (NMAX - len(L))))
    v = ctypes.c_double()
    rc = lib().cse_projector_norm(ctypes.byref(h), device, ctypes.byref(proj), Larr, int(k), int(bool(exact)),
                                  float(norm), int(bool(sampled)), int(samples), int(bins), seed & (2 ** 64 - 1),
                                  ctypes.byref(v))
    if rc == ERR_NODEV:
        raise EngineError(f"no usable gfx950 device {device} for the engine")
    if rc != 0:
        raise EngineError(f"engine error {rc}: {lib().cse_error(h).decode()}")
    return v.value


def projector_norms_batch(projs, seeds, L, k, exact, norm, sampled, samples=0, bins=1, device=0):
    """[projector_norm(projs[i], ..., seeds[i]) ...] in one cse_projector_norms
    call: exact pair sums of every projector in one batched launch."""
    h = _norm_ctx.setdefault(device, ctypes.c_void_p())
    n = len(projs)
    arr = (Projector * n)(*projs)
    sd = (U64 * max(1, n))(*[s & (2 ** 64 - 1) for s in seeds])
    Larr = (U64 * NMAX)(*(list(L) + [0] * (NMAX - len(L))))
    vals = (ctypes.c_double * max(1, n))()
    rc = lib().cse_projector_norms(ctypes.byref(h), device, n, arr, sd, Larr, int(k), int(bool(exact)),
                                   float(norm), int(bool(sampled)), int(samples), int(bins), vals)
    if rc == ERR_NODEV:
        raise EngineError(f"no usable gfx950 device {device} for the engine")
    if rc != 0:
        raise EngineError(f"engine error {rc}: {lib().cse_error(h).decode()}")
    return [vals[i] for i in range(n)]


def hist_value(hist, pmin=-128):
    h = np.ascontiguousarray(hist, dtype=np.int64)
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().cse_hist_value(_ptr(h, ctypes.c_int64), pmin, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def pack_state(n, k, Q, h, D1, D2, G, Gb, J):
    """Packed layout of include/cse.h from uint64 rows."""
    w = np.zeros(3 * n + 4, dtype=np.uint64)
    w[0] = n | (k << 8) | ((Q & 7) << 16)
    w[1], w[2], w[3] = h, D1, D2
    w[4:4 + n] = np.array(G[:n], dtype=np.uint64)
    w[4 + n:4 + 2 * n] = np.array(Gb[:n], dtype=np.uint64)
    w[4 + 2 * n:4 + 3 * n] = np.array(J[:n], dtype=np.uint64)
    return w


def unpack_state(w):
    """(n, k, Q, h, D1, D2, G rows, Gbar rows, J rows) from the packed layout"""
    w = [int(v) for v in w]
    n, k, Q = w[0] & 0xFF, (w[0] >> 8) & 0xFF, (w[0] >> 16) & 7
    return n, k, Q, w[1], w[2], w[3], w[4:4 + n], w[4 + n:4 + 2 * n], w[4 + 2 * n:4 + 3 * n]
