"""ctypes view of the CPU oracle (oracle/liborc.so) and of oracle/_ref/libref.so.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this.  Also holds a brute-force state-vector model of
the representation (semantics of the reference's stabilizer.py:54-75 unpack)
used to pin the oracle for small n.
"""
import ctypes
import gzip
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")

U64 = ctypes.c_uint64
I64 = ctypes.c_int64


class OrcState(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("k", ctypes.c_int), ("Q", ctypes.c_int),
        ("h", U64), ("D1", U64), ("D2", U64),
        ("G", U64 * 64), ("Gb", U64 * 64), ("J", U64 * 64),
    ]

    def copy(self):
        c = OrcState()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(OrcState))
        return c

    def D(self, a):
        return 2 * ((self.D1 >> a) & 1) + 4 * ((self.D2 >> a) & 1)

    def Dlist(self, upto=None):
        upto = self.n if upto is None else upto
        return [self.D(a) for a in range(upto)]

    def rows(self, name, nrows=None, ncols=None):
        nrows = self.n if nrows is None else nrows
        ncols = self.n if ncols is None else ncols
        arr = getattr(self, name)
        return np.array([[(arr[i] >> j) & 1 for j in range(ncols)] for i in range(nrows)], dtype=np.int64)


class OrcProjector(ctypes.Structure):
    _fields_ = [
        ("nstabs", ctypes.c_int), ("nqubits", ctypes.c_int),
        ("phase", ctypes.c_int * 64),
        ("xs", U64 * 64), ("zs", U64 * 64),
    ]


class IpRes(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_int), ("p", ctypes.c_int), ("m", ctypes.c_int)]


_lib = None
_ref = {}


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liborc.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ORACLE_DIR, "liborc.so")
        if not os.path.exists(path):
            build_oracle()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        L.orc_state_init.argtypes = [P(OrcState), ctypes.c_int, ctypes.c_int]
        L.orc_updateQD.argtypes = [P(OrcState), U64]
        L.orc_updateDJ.argtypes = [P(OrcState), P(U64)]
        L.orc_exponentialSum.argtypes = [P(OrcState)] + [P(ctypes.c_int)] * 3
        L.orc_shrink.argtypes = [P(OrcState), U64, ctypes.c_int, ctypes.c_int]
        L.orc_shrink.restype = ctypes.c_int
        L.orc_innerProduct.argtypes = [P(OrcState), P(OrcState)] + [P(ctypes.c_int)] * 3
        L.orc_extend.argtypes = [P(OrcState), U64]
        L.orc_measurePauli.argtypes = [P(OrcState), ctypes.c_int, U64, U64]
        L.orc_measurePauli.restype = ctypes.c_int
        L.orc_prepH.argtypes = [P(OrcState), U64, ctypes.c_int]
        L.orc_prepL.argtypes = [P(OrcState), U64, ctypes.c_int, P(U64), ctypes.c_int]
        L.orc_philox4x32_10.argtypes = [P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]
        L.orc_rand32.argtypes = [P(ctypes.c_uint32), U64, ctypes.c_uint32, U64]
        L.orc_rand32.restype = ctypes.c_uint32
        L.orc_dimension_thresholds.argtypes = [ctypes.c_int, P(U64)]
        L.orc_random_state.argtypes = [P(OrcState), ctypes.c_int, P(U64), P(ctypes.c_uint32), U64]
        L.orc_random_state.restype = ctypes.c_int
        L.orc_project.argtypes = [P(OrcState), P(OrcProjector)]
        L.orc_project.restype = ctypes.c_int
        L.orc_exact_term.argtypes = [U64, P(OrcProjector), P(U64), ctypes.c_int, ctypes.c_int,
                                     P(U64), P(U64), P(IpRes), P(ctypes.c_int)]
        L.orc_exact_term.restype = ctypes.c_int
        L.orc_zomega_add.argtypes = [P(I64), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_ip_wordops.argtypes = [P(OrcState), P(OrcState)]
        L.orc_ip_wordops.restype = ctypes.c_long
        L.orc_sampled_sweep.argtypes = [P(OrcState), ctypes.c_int, ctypes.c_int, P(U64), ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, P(I64), P(ctypes.c_int32)]
        L.orc_sampled_sweep.restype = ctypes.c_long
        _lib = L
    return _lib


def ref_available(name="libref.so"):
    return os.path.exists(os.path.join(ORACLE_DIR, "_ref", name))


def ref(name="libref.so"):
    """the reference C built from its sources: libref.so (-O2) or
    libref_O0.so (-O0, the reference Makefile's level)"""
    if name not in _ref:
        R = ctypes.CDLL(os.path.join(ORACLE_DIR, "_ref", name))
        P = ctypes.POINTER
        R.ref_innerProduct.argtypes = [P(OrcState), P(OrcState)] + [P(ctypes.c_int)] * 3
        R.ref_exponentialSum.argtypes = [P(OrcState)] + [P(ctypes.c_int)] * 3
        R.ref_shrink.argtypes = [P(OrcState), U64, ctypes.c_int, ctypes.c_int]
        R.ref_shrink.restype = ctypes.c_int
        R.ref_measurePauli.argtypes = [P(OrcState), ctypes.c_int, U64, U64]
        R.ref_measurePauli.restype = ctypes.c_double
        R.ref_prepH.argtypes = [P(OrcState), ctypes.c_int, ctypes.c_int]
        R.ref_prepL.argtypes = [P(OrcState), ctypes.c_int, ctypes.c_int, P(U64), ctypes.c_int]
        R.ref_sampled_sweep.argtypes = [P(OrcState), ctypes.c_int, ctypes.c_int, P(U64), ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, P(I64)]
        R.ref_sampled_sweep.restype = ctypes.c_long
        R.ref_sampled_sweep_terms.argtypes = [P(OrcState), ctypes.c_int, ctypes.c_int, P(U64), ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_int, P(I64),
                                              P(ctypes.c_int32)]
        R.ref_sampled_sweep_terms.restype = ctypes.c_long
        _ref[name] = R
    return _ref[name]


# ---------------------------------------------------------------- helpers

def rand_bits(rng, n):
    """uniform n-bit integer (n <= 64)"""
    v = (int(rng.integers(0, 2 ** 32)) << 32) | int(rng.integers(0, 2 ** 32))
    return v & ((1 << n) - 1)


def bits_to_u64(bits):
    v = 0
    for j, b in enumerate(np.atleast_1d(np.asarray(bits)).reshape(-1)):
        if int(b):
            v |= 1 << j
    return v


def state_from_arrays(n, k, Q, h, D, G, Gb, J):
    """G, Gb, J: n x n 0/1 (J>0 treated as 1); D: list of values (len <= n)."""
    s = OrcState()
    lib().orc_state_init(ctypes.byref(s), n, k)
    s.Q = int(Q)
    G = np.atleast_2d(np.asarray(G, dtype=float))
    Gb = np.atleast_2d(np.asarray(Gb, dtype=float))
    J = np.atleast_2d(np.asarray(J, dtype=float))
    D = np.atleast_1d(np.asarray(D, dtype=float)).reshape(-1)
    s.h = bits_to_u64(h)
    d1 = d2 = 0
    for a, v in enumerate(D):
        v = int(v) % 8
        d1 |= ((v >> 1) & 1) << a
        d2 |= ((v >> 2) & 1) << a
    s.D1, s.D2 = d1, d2
    for i in range(n):
        s.G[i] = bits_to_u64(G[i])
        s.Gb[i] = bits_to_u64(Gb[i])
        s.J[i] = bits_to_u64([1 if x > 0 else 0 for x in J[i]])
    return s


def exp_sum(s):
    e, p, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().orc_exponentialSum(ctypes.byref(s), ctypes.byref(e), ctypes.byref(p), ctypes.byref(m))
    return e.value, p.value, m.value


def inner(s1, s2):
    e, p, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().orc_innerProduct(ctypes.byref(s1), ctypes.byref(s2), ctypes.byref(e), ctypes.byref(p), ctypes.byref(m))
    return e.value, p.value, m.value


def ref_inner(s1, s2):
    e, p, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    ref().ref_innerProduct(ctypes.byref(s1), ctypes.byref(s2), ctypes.byref(e), ctypes.byref(p), ctypes.byref(m))
    return e.value, p.value, m.value


def prepH(i, t):
    s = OrcState()
    lib().orc_prepH(ctypes.byref(s), i, t)
    return s


def prepL(i, t, L):
    s = OrcState()
    arr = (U64 * len(L))(*L)
    lib().orc_prepL(ctypes.byref(s), i, t, arr, len(L))
    return s


def key2(seed):
    return (ctypes.c_uint32 * 2)(seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)


def thresholds(n):
    thr = (U64 * 65)()
    lib().orc_dimension_thresholds(n, thr)
    return thr


def random_state(n, seed, stream, thr=None):
    thr = thresholds(n) if thr is None else thr
    s = OrcState()
    w = lib().orc_random_state(ctypes.byref(s), n, thr, key2(seed), stream)
    assert w >= 0
    return s


def make_projector(phases, xs, zs, nqubits):
    P = OrcProjector()
    P.nstabs = len(phases)
    P.nqubits = nqubits
    for g in range(len(phases)):
        P.phase[g] = int(phases[g])
        P.xs[g] = xs[g]
        P.zs[g] = zs[g]
    return P


def project(s, P):
    return lib().orc_project(ctypes.byref(s), ctypes.byref(P))


def same_state(a, b, full=True):
    """Compare two oracle states; full=True compares n x n G, Gbar, J like the
    reference's shrink test (stabtests.c:215-230); D on [0,k)."""
    if (a.n, a.k, a.Q, a.h) != (b.n, b.k, b.Q, b.h):
        return False
    if a.Dlist(a.k) != b.Dlist(b.k):
        return False
    n = a.n
    for i in range(n):
        if a.G[i] != b.G[i] or a.Gb[i] != b.Gb[i]:
            return False
    if full:
        return all(a.J[i] == b.J[i] for i in range(n))
    km = (1 << a.k) - 1
    return all((a.J[i] & km) == (b.J[i] & km) for i in range(a.k))


# ------------------------------------------------- brute-force state vectors

OMEGA = np.exp(1j * np.pi / 4)


def unpack(s):
    """2^n amplitudes of |K,q>.  Amplitude index = sum_j x_j 2^j (x[0] is the
    least significant bit), the convention the reference's Unpack fixtures
    (tests-python/Unpack*.txt) hold; stabilizer.py:54-75 orders the bits the
    other way round, which is only a relabelling of qubits."""
    n, k = s.n, s.k
    psi = np.zeros(2 ** n, dtype=complex)
    G = [s.G[a] for a in range(k)]
    Joff = [[((s.J[a] >> b) & 1) if a != b else 0 for b in range(k)] for a in range(k)]
    D = s.Dlist(k)
    for z in range(2 ** k):
        x = s.h
        ph = s.Q
        zb = [(z >> a) & 1 for a in range(k)]
        for a in range(k):
            if zb[a]:
                x ^= G[a]
                ph += D[a]
        for a in range(k):
            if zb[a]:
                for b in range(a + 1, k):
                    if zb[b] and Joff[a][b]:
                        ph += 4
        psi[x] += OMEGA ** (ph % 8)
    return psi / 2 ** (k / 2)


def pauli_matrix_action(psi, n, m, zeta, xi):
    """P psi for P = i^m Z(zeta) X(xi) (stabilizer.py:454 convention):
    |x> -> i^m (-1)^{zeta.(x+xi)} |x+xi>, index = x as an integer."""
    out = np.zeros_like(psi)
    for x in range(2 ** n):
        y = x ^ xi
        sign = -1 if bin(y & zeta).count("1") % 2 else 1
        out[y] += (1j ** m) * sign * psi[x]
    return out


def value(eps, p, m):
    return eps * 2 ** (p / 2) * OMEGA ** (m % 8)


# ------------------------------------------------------------ fixtures

def read_fixture_lines(name):
    with gzip.open(os.path.join(GOLDEN, "ref", name), "rt") as f:
        return [ln.strip() for ln in f.read().split("\n")]


def parse_list(line):
    return [float(x) for x in line.split(",") if x != ""]


def load_json_fixture(name):
    with gzip.open(os.path.join(GOLDEN, "ref", name), "rt") as f:
        return json.load(f)


def sampled_sweep(thetas, t, L, k, nphi=None, threads=1, terms=False, lib_ref=None):
    """Per-theta Z[w] sums (scale 2^32) of sum_{i < nphi} <prepL(i)|theta'> for
    projected thetas, by the oracle (lib_ref None) or by the reference C
    (lib_ref = a libref name); with terms, also every packed term.
    Returns (csum as an (ntheta, 4) int64 array, terms as (ntheta, nphi)
    int32 or None, inner products computed)."""
    nphi = (1 << k) if nphi is None else int(nphi)
    n = len(thetas)
    arr = (OrcState * n)(*thetas)
    Larr = (U64 * max(1, k))(*L)
    cs = np.zeros((n, 4), dtype=np.int64)
    tm = np.zeros((n, nphi), dtype=np.int32) if terms else None
    tp = tm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if terms else None
    cp = cs.ctypes.data_as(ctypes.POINTER(I64))
    if lib_ref is None:
        cnt = lib().orc_sampled_sweep(arr, n, t, Larr, k, 32, threads, nphi, cp, tp)
    else:
        cnt = ref(lib_ref).ref_sampled_sweep_terms(arr, n, t, Larr, k, 32, threads, nphi, cp, tp)
    return cs, tm, cnt
