"""HIP engine vs the CPU oracle on the same seeded inputs, through the C ABI
(libcse.so).  Integer work must match bit for bit: per-term (eps, p, m mod 8)
for every inner product, full states where the device mirrors the reference's
row bookkeeping, exact per-sample sums and histograms."""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

import orc
from gpu_util import (from_packed, full_equal, oracle_exact_terms, oracle_sampled, pack_term,
                      random_commuting_projector, term_value, to_packed)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from circuitsimulator_amd import Engine
    e = Engine(0)
    yield e
    e.close()


def _mixed_states(n, seed, count=10):
    rng = np.random.default_rng(seed)
    thr = orc.thresholds(n)
    states = []
    for st in range(count):
        s = orc.random_state(n, seed, st, thr)
        if st % 2:
            ph, xs, zs = random_commuting_projector(rng, n, int(rng.integers(1, max(2, n // 2))))
            t = s.copy()
            if orc.project(t, orc.make_projector(ph, xs, zs, n)) >= 0:
                s = t
        states.append(s)
    L = [orc.rand_bits(rng, n) for _ in range(min(6, n))]
    states += [orc.prepL(int(rng.integers(0, 2 ** len(L))), n, L) for _ in range(3)]
    if n <= 32:
        states += [orc.prepH(int(rng.integers(0, 2 ** ((n + 1) // 2))), n) for _ in range(3)]
    return states


@pytest.mark.parametrize("n", [1, 3, 8, 24, 48, 59, 60, 61, 64])
def test_inner_products_all_pairs(E, n):
    # n <= 60 takes the packed restrictions and loop (row 60 the end row),
    # n > 60 the general compiled path
    states = _mixed_states(n, 40 + n)
    pk = [to_packed(s) for s in states]
    A = np.array([pk[a] for a in range(len(pk)) for b in range(len(pk))])
    B = np.array([pk[b] for a in range(len(pk)) for b in range(len(pk))])
    out, _ = E.inner_products(n, A, B)
    idx = 0
    for a in states:
        for b in states:
            eps, p, m = orc.inner(a, b)
            g = out[idx]
            assert g[0] == eps, (n, idx)
            if eps:
                assert (g[1], g[2] % 8) == (p, m % 8), (n, idx)
            idx += 1


def test_inner_products_one_launch_per_pair(E):
    n = 48
    states = _mixed_states(n, 7, 6)
    pk = [to_packed(s) for s in states]
    A = np.array(pk[:-1])
    B = np.array(pk[1:])
    out1, _ = E.inner_products(n, A, B, one_launch_per_pair=True)
    out2, _ = E.inner_products(n, A, B, one_launch_per_pair=False)
    assert np.array_equal(out1, out2)
    for i in range(len(states) - 1):
        eps, p, m = orc.inner(states[i], states[i + 1])
        assert out1[i][0] == eps and (not eps or (out1[i][1], out1[i][2] % 8) == (p, m % 8))


@pytest.mark.parametrize("t", [1, 2, 5, 6, 11, 24, 48, 64])
def test_prep_states_full_state(E, t):
    rng = np.random.default_rng(t)
    k = min(t, 10)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    idx = [int(v) for v in rng.integers(0, 2 ** k, size=8)]
    got = E.prep_states(0, t, L, idx)
    for c, i in enumerate(idx):
        assert full_equal(from_packed(got[c]), orc.prepL(i, t, L)), (t, i)
    if t <= 40:
        idx = [int(v) for v in rng.integers(0, 2 ** ((t + 1) // 2), size=8)]
        got = E.prep_states(1, t, [], idx)
        for c, i in enumerate(idx):
            assert full_equal(from_packed(got[c]), orc.prepH(i, t)), (t, i)


@pytest.mark.parametrize("n", [1, 2, 5, 24, 48, 64])
def test_random_states_full_state(E, n):
    seed = 0xC0FFEE + n
    got, sexp = E.random_states(n, seed, 100, 24)
    thr = orc.thresholds(n)
    for c in range(24):
        want = orc.random_state(n, seed, 100 + c, thr)
        assert full_equal(from_packed(got[c]), want), (n, c)


@pytest.mark.parametrize("n,r", [(6, 3), (24, 12), (48, 24)])
def test_random_states_projected_full_state(E, n, r):
    from circuitsimulator_amd.engine import Projector
    rng = np.random.default_rng(n)
    ph, xs, zs = random_commuting_projector(rng, n, r)
    seed = 77 + n
    got, sexp = E.random_states(n, seed, 0, 20, Projector.make(ph, xs, zs, n))
    thr = orc.thresholds(n)
    Po = orc.make_projector(ph, xs, zs, n)
    for c in range(20):
        want = orc.random_state(n, seed, c, thr)
        e = orc.project(want, Po)
        assert sexp[c] == e, (n, c)
        if e >= 0:
            assert full_equal(from_packed(got[c]), want), (n, c)


def _exact_mean(t, sexp, zsum):
    """exact sampled estimator as a Fraction-free float via integers: sum of
    2^{t-s-64} (A + B sqrt2) / count, evaluated with Python integers"""
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    tot = Decimal(0)
    r2 = Decimal(2).sqrt()
    for s, c in zip(sexp, zsum):
        if s < 0:
            continue
        c0, c1, c2, c3 = (int(v) for v in c)
        A = c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3
        B = c0 * c1 - c0 * c3 + c1 * c2 + c2 * c3
        tot += (Decimal(A) + Decimal(B) * r2) * (Decimal(2) ** (t - int(s) - 64))
    return float(tot / len(sexp))


@pytest.mark.parametrize("t,r,k,exact", [(6, 3, 3, 0), (8, 4, 4, 0), (7, 2, 0, 1), (10, 5, 5, 0), (12, 6, 0, 1),
                                         (14, 6, 0, 1), (15, 5, 0, 1)])
def test_sampled_small_all_terms(E, t, r, k, exact):
    from circuitsimulator_amd.engine import Projector
    rng = np.random.default_rng(100 * t + r)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    chi = (1 << ((t + 1) // 2)) if exact else (1 << k)
    seed, first, count = 1234 + t, 50, 48
    res, ex = E.sampled_projector(Projector.make(ph, xs, zs, t), L, k, exact, seed, first, count,
                                  chunk=max(1, chi // 3), want_terms=True)
    se, zsum, terms = oracle_sampled(t, k, L, exact, ph, xs, zs, seed, first, count, chi)
    assert np.array_equal(ex["sexp"], se)
    assert np.array_equal(ex["terms"], terms)
    assert np.array_equal(ex["zsum"], zsum)
    assert res.ips == int((se >= 0).sum()) * chi
    want = _exact_mean(t, se, zsum)
    assert abs(res.mean - want) <= 1e-15 * max(1.0, abs(want))
    # FP32-Kahan shadow: relative error bound of FP32 accumulation
    assert abs(res.mean_kahan - want) <= 1e-5 * max(1.0, abs(want))


def test_sampled_estimators_of_many_samples(E):
    """40,000 samples: the engine splits the estimators into parts of 8,192
    samples over host threads.  The int64 mean equals the one-pass host
    routine on the same per-sample sums bit for bit, and the exact value to
    one rounding; the Kahan mean agrees to FP32 accumulation."""
    from circuitsimulator_amd.engine import LIB_PATH, Projector
    rng = np.random.default_rng(4242)
    t, r, k, count = 16, 6, 4, 40000
    ph, xs, zs = random_commuting_projector(rng, t, r)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    res, ex = E.sampled_projector(Projector.make(ph, xs, zs, t), L, k, 0, 99, 0, count)
    P = ctypes.POINTER
    zsum = np.ascontiguousarray(ex["zsum"]).reshape(-1)
    sexp = np.ascontiguousarray(ex["sexp"])
    kz = np.ascontiguousarray(ex["kahan"]).reshape(-1)
    m1, k1 = ctypes.c_double(), ctypes.c_double()
    ctypes.CDLL(LIB_PATH).cse_sampled_estimators(
        ctypes.c_int(t), ctypes.c_int64(count), zsum.ctypes.data_as(P(ctypes.c_longlong)),
        sexp.ctypes.data_as(P(ctypes.c_int32)), kz.ctypes.data_as(P(ctypes.c_double)), ctypes.byref(m1),
        ctypes.byref(k1))
    assert res.mean == m1.value
    want = _exact_mean(t, sexp, ex["zsum"])
    assert abs(res.mean - want) <= 1e-15 * max(1.0, abs(want))
    assert abs(res.mean_kahan - want) <= 1e-5 * max(1.0, abs(want))
    assert abs(res.mean_kahan - k1.value) <= 1e-12 * abs(k1.value)


@pytest.mark.parametrize("t,r,exact,k", [(1, 1, 1, 0), (3, 2, 1, 0), (6, 3, 1, 0), (5, 2, 0, 3), (9, 4, 0, 4),
                                         (8, 8, 1, 0)])
def test_exact_all_terms(E, t, r, exact, k):
    from circuitsimulator_amd.engine import Instance, Projector
    rng = np.random.default_rng(7 * t + r + 100 * exact)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    inst = Instance()
    inst.P = Projector.make(ph, xs, zs, t)
    inst.exact = exact
    inst.k = k
    for i, v in enumerate(L):
        inst.L[i] = v
    chi = (1 << ((t + 1) // 2)) if exact else (1 << k)
    nl = chi * (chi + 1) // 2
    res, terms, _ = E.batch_exact([inst], chunk=max(1, chi // 2), want_terms=True, nterms=nl)
    want = oracle_exact_terms(t, k, L, exact, ph, xs, zs)
    assert [w[0] for w in want] == list(terms)
    tot = 0
    for (pt, s, i, j) in want:
        if pt:
            v = term_value(pt, s)
            tot += v if i == j else 2 * v.real
    assert abs(res[0].re - tot.real) <= 1e-12 * max(1, abs(tot))
    assert res[0].im == 0.0
    assert abs(res[0].value - abs(tot)) <= 1e-12 * max(1, abs(tot))
    assert abs(res[0].kahan_re - tot.real) <= 1e-5 * max(1, abs(tot))


def _exact_instance_terms(E, t, ph, xs, zs, chunk):
    from circuitsimulator_amd.engine import Instance, Projector
    inst = Instance()
    inst.P = Projector.make(ph, xs, zs, t)
    inst.exact = 1
    inst.k = 0
    chi = 1 << ((t + 1) // 2)
    nl = chi * (chi + 1) // 2
    res, terms, _ = E.batch_exact([inst], chunk=chunk, want_terms=True, nterms=nl)
    want = oracle_exact_terms(t, 0, [], 1, ph, xs, zs)
    assert [w[0] for w in want] == list(terms)
    tot = 0
    for (pt, s, i, j) in want:
        if pt:
            v = term_value(pt, s)
            tot += v if i == j else 2 * v.real
    assert abs(res[0].re - tot.real) <= 1e-12 * max(1, abs(tot))
    assert res[0].im == 0.0
    return terms


@pytest.mark.parametrize("t,r,chunk", [(13, 5, 0), (14, 6, 64), (15, 4, 0), (16, 7, 128)])
def test_exact_H_terms_with_shared_pair_operations(E, t, r, chunk):
    # |H^t> terms at t >= 13, where a batch's terms share the pair bits
    # above their six low ones: those pair operations run once per batch
    # (phiH_prefix), for aligned batches and for each record's first,
    # unaligned one (from its diagonal); every term against the oracle
    rng = np.random.default_rng(900 + t)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    print(f"exact |H^t> terms t={t} r={r} chunk={chunk}", flush=True)
    _exact_instance_terms(E, t, ph, xs, zs, chunk)


def test_exact_H_terms_when_a_shared_constraint_vanishes(E):
    # -Z_0 Z_1 fixes x_0 + x_1 = 1: every j whose highest pair bit (pair 0)
    # is set gives zero, and in aligned batches that bit is shared, so the
    # batch's shared constraint vanishes for all 64 terms
    t = 16
    terms = _exact_instance_terms(E, t, [2, 0, 0], [0, 0, 0b110000], [0b11, 0b1100, 0], 0)
    assert np.count_nonzero(terms) > 0 and np.count_nonzero(terms == 0) > 0


@pytest.mark.parametrize("kern", [1, 2])
def test_exact_H_terms_when_a_group_constraint_vanishes(E, kern):
    # -Z_8 Z_9 fixes x_8 + x_9 = 1 at t=16: pair 4 is bit 3 of j, which a
    # group of 8 terms in an aligned batch shares, so that group's shared
    # constraint vanishes while the other groups of its batch do not
    t = 16
    print(f"exact |H^t> group constraint t={t} kernel={kern}", flush=True)
    E.set_sweep_kernel(kern, mode=1)
    try:
        terms = _exact_instance_terms(E, t, [2, 0], [0, 0b11], [0b11 << 8, 0], 0)
    finally:
        E.set_sweep_kernel(0, mode=1)
    assert np.count_nonzero(terms) > 0 and np.count_nonzero(terms == 0) > 0


@pytest.mark.parametrize("kern,t,r", [(1, 17, 6), (2, 17, 6), (1, 18, 9), (2, 18, 9)])
def test_exact_H_terms_each_exact_kernel(E, kern, t, r):
    # both exact-mode sweep kernels run aligned batches as eight groups of
    # 8 terms, each from its own shared state
    rng = np.random.default_rng(1700 + 10 * t + kern)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    print(f"exact |H^t> terms t={t} r={r} kernel={kern}", flush=True)
    E.set_sweep_kernel(kern, mode=1)
    try:
        _exact_instance_terms(E, t, ph, xs, zs, 0)
    finally:
        E.set_sweep_kernel(0, mode=1)


@pytest.mark.parametrize("kern,t", [(1, 16), (2, 16), (2, 17)])
def test_sampled_H_terms_aligned_batches(E, kern, t):
    # sampled |H^t> terms in items of 128 j: aligned 64-term batches, run as
    # one batch (two-column kernel) or as eight groups (general kernel); odd
    # t has a last "pair" of one qubit (a constraint, never a coupling)
    from circuitsimulator_amd.engine import Projector
    r, k = 6, 0
    rng = np.random.default_rng(4242 + kern + 100 * t)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    chi = 1 << ((t + 1) // 2)
    seed, first, count = 77, 3, 24
    print(f"sampled |H^t> terms t={t} r={r} kernel={kern}", flush=True)
    E.set_sweep_kernel(kern)
    try:
        res, ex = E.sampled_projector(Projector.make(ph, xs, zs, t), [], k, 1, seed, first, count, chunk=128,
                                      want_terms=True)
    finally:
        E.set_sweep_kernel(0)
    se, zsum, terms = oracle_sampled(t, k, [], 1, ph, xs, zs, seed, first, count, chi)
    assert np.array_equal(ex["sexp"], se)
    assert np.array_equal(ex["terms"], terms)
    assert np.array_equal(ex["zsum"], zsum)


def test_batch_uneven_instances(E):
    from circuitsimulator_amd.engine import Instance, Projector
    rng = np.random.default_rng(2024)
    insts, specs = [], []
    for _ in range(12):
        t = int(rng.integers(2, 13))
        r = int(rng.integers(1, t // 2 + 1))
        exact = int(rng.integers(0, 2)) if t <= 10 else 0
        k = 0 if exact else int(rng.integers(1, min(t, 5) + 1))
        ph, xs, zs = random_commuting_projector(rng, t, r)
        L = [orc.rand_bits(rng, t) for _ in range(k)]
        inst = Instance()
        inst.P = Projector.make(ph, xs, zs, t)
        inst.exact = exact
        inst.k = k
        for i, v in enumerate(L):
            inst.L[i] = v
        insts.append(inst)
        specs.append((t, k, L, exact, ph, xs, zs))
    nterms = 0
    for (t, k, L, exact, *_r) in specs:
        chi = (1 << ((t + 1) // 2)) if exact else (1 << k)
        nterms += chi * (chi + 1) // 2
    res, terms, _ = E.batch_exact(insts, chunk=7, want_terms=True, nterms=nterms)
    off = 0
    for q, (t, k, L, exact, ph, xs, zs) in enumerate(specs):
        want = oracle_exact_terms(t, k, L, exact, ph, xs, zs)
        got = list(terms[off:off + len(want)])
        assert got == [w[0] for w in want], q
        off += len(want)
        assert res[q].im == 0.0


def test_exact_diagonal_exponents_above_64(E):
    # an uneven batch instance (t=48, r=14, k=7): on the diagonal terms a
    # projected |L> state meets itself, and p = n0 + nC reaches ~88
    from circuitsimulator_amd.workloads import uneven_instances
    inst, specs = uneven_instances(140, seed=1003)
    t, r, k, ph, xs, zs, L = specs[139]
    chi = 1 << k
    res, terms, _ = E.batch_exact([inst[139]], want_terms=True, nterms=chi * (chi + 1) // 2)
    diag = [i * chi - i * (i - 1) // 2 for i in range(chi)]
    want = oracle_exact_terms(t, k, [int(v) for v in L], 0, ph, xs, zs, ls=diag)
    assert [int(terms[l]) for l in diag] == [w[0] for w in want]
    assert res[0].im == 0.0


def test_batch_uneven_large_t_runs_clean(E):
    # the first 330 instances of the uneven batch workload (t up to 48, k up
    # to 14): every total exactly real and non-negative, no term out of range
    from circuitsimulator_amd.workloads import uneven_instances
    inst, specs = uneven_instances(330, seed=1003)
    res, _, _ = E.batch_exact(inst)
    assert all(r.im == 0.0 and r.re >= 0.0 for r in res)
    assert [r.terms for r in res] == [(1 << k) * ((1 << k) + 1) // 2 for (_, _, k, *_x) in specs]


def test_uneven_workload_seeded_subsample(E):
    # the uneven batch workload itself (4,096 instances, t up to 48): 64
    # seeded instances with k <= 10 in one launch, and per instance a seeded
    # subsample of its pair terms (always the first and last) against the
    # oracle
    from circuitsimulator_amd.workloads import uneven_instances
    inst, specs = uneven_instances(4096, seed=1003)
    rng = np.random.default_rng(3003)
    pick = [int(q) for q in rng.permutation(len(inst)) if specs[q][2] <= 10][:64]
    sizes = [(1 << specs[q][2]) * ((1 << specs[q][2]) + 1) // 2 for q in pick]
    res, terms, _ = E.batch_exact([inst[q] for q in pick], want_terms=True, nterms=sum(sizes))
    off = 0
    for q, nl in zip(pick, sizes):
        t, r, k, ph, xs, zs, L = specs[q]
        ls = sorted(set([0, nl - 1] + [int(v) for v in rng.integers(0, nl, size=120)]))
        want = oracle_exact_terms(t, k, [int(v) for v in L], 0, ph, xs, zs, ls=ls)
        assert [int(terms[off + l]) for l in ls] == [w[0] for w in want], q
        off += nl
    assert all(r.im == 0.0 for r in res)


def test_uneven_workload_large_k_subsample(E):
    # a seeded 3% subsample of the uneven batch workload with no filter on k:
    # instances taken in a seeded order until 40 of them have k >= 11 (those
    # hold 99.6% of the workload's terms), 123 in all, k = 4..14.  They run in
    # launches of at most 2^28 pair terms with every term written out, and
    # per instance 100 seeded terms (always the first and last) are checked
    # against the oracle.
    from concurrent.futures import ThreadPoolExecutor

    from circuitsimulator_amd.workloads import uneven_instances
    inst, specs = uneven_instances(4096, seed=1003)
    rng = np.random.default_rng(3004)
    pick, big = [], 0
    for q in rng.permutation(len(inst)):
        pick.append(int(q))
        big += specs[q][2] >= 11
        if big == 40:
            break
    assert len(pick) == 123 and max(specs[q][2] for q in pick) == 14
    size = {q: (1 << specs[q][2]) * ((1 << specs[q][2]) + 1) // 2 for q in pick}
    groups, cur = [], []
    for q in sorted(pick, key=lambda q: size[q]):
        if cur and sum(size[c] for c in cur) + size[q] > (1 << 28):
            groups.append(cur)
            cur = []
        cur.append(q)
    groups.append(cur)
    checks = []
    for g in groups:
        print(f"large-k subsample: launch of {len(g)} instances, k = {sorted({specs[q][2] for q in g})}", flush=True)
        res, terms, _ = E.batch_exact([inst[q] for q in g], want_terms=True, nterms=sum(size[q] for q in g))
        assert all(r.im == 0.0 and r.re >= 0.0 for r in res)
        assert [r.terms for r in res] == [size[q] for q in g]
        off = 0
        for q in g:
            nl = size[q]
            ls = sorted(set([0, nl - 1] + [int(v) for v in rng.integers(0, nl, size=98)]))
            checks.append((q, ls, [int(terms[off + l]) for l in ls]))
            off += nl
        del terms

    def oracle(c):
        q, ls, got = c
        t, r, k, ph, xs, zs, L = specs[q]
        want = [w[0] for w in oracle_exact_terms(t, k, [int(v) for v in L], 0, ph, xs, zs, ls=ls)]
        return q, got == want

    with ThreadPoolExecutor(8) as pool:
        bad = [q for q, ok in pool.map(oracle, checks) if not ok]
    assert not bad, bad
    assert sum(len(c[1]) for c in checks if specs[c[0]][2] >= 11) >= 40 * 99


def test_sampled_t64_unpacked_layout(E):
    # n > 61 runs the sweep with D in its own register
    from circuitsimulator_amd.engine import Projector
    t, r, k = 64, 4, 3
    rng = np.random.default_rng(64)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    res, ex = E.sampled_projector(Projector.make(ph, xs, zs, t), L, k, 0, 3, 0, 10, want_terms=True)
    se, zsum, terms = oracle_sampled(t, k, L, 0, ph, xs, zs, 3, 0, 10, 1 << k)
    assert np.array_equal(ex["sexp"], se)
    assert np.array_equal(ex["terms"], terms)
    assert np.array_equal(ex["zsum"], zsum)


@pytest.mark.parametrize("chunk", [100, 192, 64])
def test_sampled_chunks_across_batches(E, chunk):
    # chi = 256 split into items of `chunk` j: items starting at a multiple
    # of 64 build x_j from a low-bit lane table and a high part advanced
    # between batches; the others build it per lane
    from circuitsimulator_amd.engine import Projector
    t, r, k = 30, 8, 8
    rng = np.random.default_rng(700 + chunk)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    chi = 1 << k
    res, ex = E.sampled_projector(Projector.make(ph, xs, zs, t), L, k, 0, 11, 2, 4, chunk=chunk, want_terms=True)
    se, zsum, terms = oracle_sampled(t, k, L, 0, ph, xs, zs, 11, 2, 4, chi)
    assert np.array_equal(ex["sexp"], se)
    assert np.array_equal(ex["terms"], terms)
    assert np.array_equal(ex["zsum"], zsum)


@pytest.mark.parametrize("wide", [False, True])
def test_batch_sampled_against_oracle_and_single_calls(E, wide):
    # uneven sampled-mode instances (t, r, k, exact, seed, first, count) in
    # one call: each instance's samples match the oracle and its mean equals
    # cse_sampled_projector's bit for bit.  With a t = 61 instance the whole
    # batch runs the unpacked sweep.
    from circuitsimulator_amd.engine import (SAMPLED_INSTANCE_DTYPE, Projector, SampledInstance)
    specs = [(6, 3, 3, 0, 5, 0, 7), (7, 2, 0, 1, 9, 100, 5), (10, 5, 5, 0, 5, 7, 1), (30, 8, 8, 0, 11, 2, 4),
             (12, 6, 0, 1, 3, 40, 6), (24, 6, 7, 0, 5, 0, 9)]
    if wide:
        specs.insert(2, (61, 12, 3, 0, 4, 0, 2))
    rng = np.random.default_rng(808 + wide)
    arr = (SampledInstance * len(specs))()
    data = []
    for i, (t, r, k, exact, seed, first, count) in enumerate(specs):
        ph, xs, zs = random_commuting_projector(rng, t, r)
        L = [orc.rand_bits(rng, t) for _ in range(k)]
        arr[i].P = Projector.make(ph, xs, zs, t)
        arr[i].exact, arr[i].k = exact, k
        for j, v in enumerate(L):
            arr[i].L[j] = v
        arr[i].seed, arr[i].first, arr[i].count = seed, first, count
        data.append((ph, xs, zs, L))
    A = np.frombuffer(arr, dtype=SAMPLED_INSTANCE_DTYPE)
    means, res, _, samp = E.batch_sampled(A, want_samples=True)
    for i, (t, r, k, exact, seed, first, count) in enumerate(specs):
        ph, xs, zs, L = data[i]
        chi = (1 << ((t + 1) // 2)) if exact else (1 << k)
        se, zsum, _ = oracle_sampled(t, k, L, exact, ph, xs, zs, seed, first, count, chi)
        assert np.array_equal(samp[i]["sexp"], se), i
        assert np.array_equal(samp[i]["zsum"], zsum), i
        one, _ = E.sampled_projector(arr[i].P, L, k, exact, seed, first, count, want_samples=False)
        assert means[i] == one.mean, i
        assert res[i].ips == one.ips == int((se >= 0).sum()) * chi
        assert res[i].vanished == one.vanished == int((se < 0).sum())


def test_batch_sampled_rejects_bad_counts(E):
    # checked on the host before anything is uploaded or launched
    from circuitsimulator_amd.engine import EngineError, sampled_instance_array
    A = sampled_instance_array(2)
    A["P"]["nqubits"] = 8
    A["P"]["nstabs"] = 1
    A["P"]["xs"][:, 0] = 1
    A["k"] = 2
    A["L"][:, :2] = [3, 12]
    A["count"] = [5, 0]
    with pytest.raises(EngineError, match="count=0"):
        E.batch_sampled(A)
    A["count"] = [1 << 23, (1 << 23) + 1]
    with pytest.raises(EngineError, match="2\\^24"):
        E.batch_sampled(A)


def test_state_calls_reject_counts_above_the_limit(E):
    # at most 2^30 pairs or states per call, checked on the host before any
    # buffer is read or anything is launched (the buffers here hold one)
    import ctypes

    from circuitsimulator_amd.engine import U64, _ptr
    L = E._L
    big = (1 << 30) + 1
    st = np.zeros(3 * 4 + 4, dtype=np.uint64)
    out = np.zeros(16, dtype=np.int32)
    assert L.cse_inner_products(E.h, 4, big, _ptr(st, U64), _ptr(st, U64), _ptr(out, ctypes.c_int32), 0,
                                None) < 0
    assert b"2^30" in L.cse_error(E.h)
    idx = np.zeros(1, dtype=np.uint64)
    assert L.cse_prep_states(E.h, 1, 4, None, 0, _ptr(idx, U64), big, _ptr(st, U64)) < 0
    assert b"2^30" in L.cse_error(E.h)
    sexp = np.zeros(1, dtype=np.int32)
    assert L.cse_random_states(E.h, 4, 1, 0, big, None, None, _ptr(sexp, ctypes.c_int32)) < 0
    assert b"2^30" in L.cse_error(E.h)
    # the context still works
    res = E.prep_states(1, 4, [], [0, 3])
    assert res.shape == (2, 16)


def test_projector_norms_batch_equals_single_calls():
    # cse_projector_norms: exact pair sums of several projectors in one
    # launch equal cse_projector_norm's single calls bit for bit; closed
    # forms and sampled projectors (own seeds) as the single calls
    from circuitsimulator_amd.engine import Projector, projector_norm, projector_norms_batch
    from circuitsimulator_amd.workloads import sampled_config
    P1, _, L = sampled_config(14, 5, 6, seed=71)
    P2, _, _ = sampled_config(14, 4, 6, seed=72)
    empty = Projector()
    empty.nqubits = 14
    norm = 1.7
    for exact, k in [(0, 6), (1, 0)]:
        projs = [P1, P2, empty, P1]
        seeds = [11, 12, 13, 14]
        got = projector_norms_batch(projs, seeds, L, k, exact, norm, False)
        want = [projector_norm(p, L, k, exact, norm, False, seed=sd) for p, sd in zip(projs, seeds)]
        assert got == want, (exact, got, want)
        assert got[2] == norm * norm and got[0] == got[3]
    got = projector_norms_batch([P1, P2], [5, 6], L, 6, 0, norm, True, samples=64, bins=3)
    want = [projector_norm(P1, L, 6, 0, norm, True, 64, 3, 5), projector_norm(P2, L, 6, 0, norm, True, 64, 3, 6)]
    assert got == want


def test_reused_buffers_start_clean(E):
    # a large batch, then a smaller one on the same context: the one clear
    # launch resets every sum, counter, queue and flag the smaller call
    # reads, so it matches the same call on a fresh context
    from circuitsimulator_amd import Engine
    from circuitsimulator_amd.engine import sampled_instance_array
    from circuitsimulator_amd.workloads import sampled_config, uneven_instances
    big, _ = uneven_instances(30, seed=31, kmax=9)
    small, _ = uneven_instances(5, seed=32, kmax=6)

    def sampled(specs):
        A = sampled_instance_array(len(specs))
        for i, (t, r, k, count) in enumerate(specs):
            P, _, L = sampled_config(t, r, k, seed=500 + i)
            A[i]["P"] = np.frombuffer(bytes(P), dtype=A.dtype["P"])[0]
            A[i]["k"] = k
            A[i]["L"][:k] = L
            A[i]["seed"], A[i]["count"] = 40 + i, count
        return A
    sbig = sampled([(30, 8, 9, 400), (20, 6, 7, 300), (40, 10, 8, 200)])
    ssmall = sampled([(12, 4, 5, 7), (16, 5, 6, 9)])
    E.batch_exact(big)
    after, _, _ = E.batch_exact(small)
    E.batch_sampled(sbig)
    safter, sres, _, _ = E.batch_sampled(ssmall)
    with Engine(0) as F:
        fresh, _, _ = F.batch_exact(small)
        sfresh, sfres, _, _ = F.batch_sampled(ssmall)
    assert [(a.re, a.im, a.terms, a.ips) for a in after] == [(b.re, b.im, b.terms, b.ips) for b in fresh]
    assert np.array_equal(safter, sfresh)
    assert [(a.ips, a.vanished) for a in sres] == [(b.ips, b.vanished) for b in sfres]


@pytest.mark.parametrize("bins", [1, 3, 4])
def test_projector_norm_bins_in_one_call(E, bins):
    # cse_projector_norm sends every bin as one instance of one batched call;
    # the median equals that of one cse_sampled_projector call per bin
    from circuitsimulator_amd.engine import Projector, projector_norm
    t, r, k, samples, seed = 16, 6, 5, 120, 4242
    rng = np.random.default_rng(1600 + bins)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    P = Projector.make(ph, xs, zs, t)
    got = projector_norm(P, L, k, 0, 1.0, 1, samples, bins, seed)
    vals = sorted(E.sampled_projector(P, L, k, 0, seed, b * samples, samples, want_samples=False)[0].mean
                  for b in range(bins))
    want = vals[(bins - 1) // 2] if bins % 2 else (vals[bins // 2] + vals[bins // 2 - 1]) / 2
    assert got == want


@pytest.mark.parametrize("t", [60, 61])
def test_sampled_layout_limits(E, t):
    # n = 60 is the packed layout's limit (row 60 is the loop's end row);
    # n = 61 runs the unpacked sweep
    from circuitsimulator_amd.engine import Projector
    r, k = 12, 6
    rng = np.random.default_rng(600 + t)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    res, ex = E.sampled_projector(Projector.make(ph, xs, zs, t), L, k, 0, 5, 0, 6, want_terms=True)
    Po = orc.make_projector(ph, xs, zs, t)
    thr = orc.thresholds(t)
    phis = [orc.prepL(j, t, L) for j in range(1 << k)]
    for s in range(6):
        th = orc.random_state(t, 5, s, thr)
        e = orc.project(th, Po)
        assert ex["sexp"][s] == e
        if e < 0:
            continue
        for j in range(1 << k):
            eps, p, m = orc.inner(th, phis[j])
            assert ex["terms"][s][j] == pack_term(eps, p, m), (s, j)


def test_sampled_t48_workload_subsample(E):
    # the benchmark's workload shape: t=48, r=24, k=12 (chi=4096)
    from circuitsimulator_amd.workloads import sampled_config
    P, (ph, xs, zs), L = sampled_config(48, 24, 12, seed=1004)
    t, k, chi, count = 48, 12, 4096, 3
    res, ex = E.sampled_projector(P, L, k, 0, 77, 5, count, want_terms=True)
    Po = orc.make_projector(ph, xs, zs, t)
    thr = orc.thresholds(t)
    rng = np.random.default_rng(48)
    for s in range(count):
        th = orc.random_state(t, 77, 5 + s, thr)
        e = orc.project(th, Po)
        assert ex["sexp"][s] == e
        if e < 0:
            assert not ex["terms"][s].any()
            continue
        for i in rng.integers(0, chi, size=120):
            eps, p, m = orc.inner(th, orc.prepL(int(i), t, L))
            assert ex["terms"][s][i] == pack_term(eps, p, m), (s, int(i))
    assert res.ips == int((ex["sexp"] >= 0).sum()) * chi


def test_sampled_exact_decomposition_t44_subsample(E):
    # |H^t> terms at t=44 (chi = 2^22): phi_0 has every pair coupled and no
    # restriction, so with theta it can reach p = n0 + nC above 64
    from circuitsimulator_amd.engine import Projector
    t, r = 44, 10
    rng = np.random.default_rng(4401)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    chi = 1 << ((t + 1) // 2)
    count = 2
    res, ex = E.sampled_projector(Projector.make(ph, xs, zs, t), [], 0, 1, 91, 3, count, want_terms=True)
    Po = orc.make_projector(ph, xs, zs, t)
    thr = orc.thresholds(t)
    js = [0, 1, chi - 1] + [int(v) for v in rng.integers(0, chi, size=60)]
    for s in range(count):
        th = orc.random_state(t, 91, 3 + s, thr)
        e = orc.project(th, Po)
        assert ex["sexp"][s] == e
        if e < 0:
            continue
        for j in js:
            eps, p, m = orc.inner(th, orc.prepH(j, t))
            assert ex["terms"][s][j] == pack_term(eps, p, m), (s, j)
    assert res.ips == int((ex["sexp"] >= 0).sum()) * chi


def test_exact_t24_workload_subsample(E):
    # exact mode at t=24, r=12, k=11: chi(chi+1)/2 = 2,098,176 pair terms
    from circuitsimulator_amd.engine import Instance
    from circuitsimulator_amd.workloads import commuting_paulis, random_L
    rng = np.random.default_rng(1002)
    t, r, k = 24, 12, 11
    ph, xs, zs = commuting_paulis(rng, t, r)
    L = random_L(rng, t, k)
    from circuitsimulator_amd.engine import Projector
    inst = Instance()
    inst.P = Projector.make(ph, xs, zs, t)
    inst.exact = 0
    inst.k = k
    for i, v in enumerate(L):
        inst.L[i] = v
    chi = 1 << k
    nl = chi * (chi + 1) // 2
    res, terms, _ = E.batch_exact([inst], want_terms=True, nterms=nl)
    assert res[0].terms == nl
    ls = [int(v) for v in np.random.default_rng(5).integers(0, nl, size=300)]
    want = oracle_exact_terms(t, k, L, 0, ph, xs, zs, ls=ls)
    for l, w in zip(ls, want):
        assert terms[l] == w[0], l
    # the total is real: conjugate pairs and real diagonal terms
    assert res[0].im == 0.0 and res[0].re >= 0.0
    assert abs(res[0].kahan_re - res[0].re) <= 1e-5 * max(1.0, res[0].re)


@pytest.mark.parametrize("chunk", [512, 1000])
def test_exact_item_sizes_give_identical_terms(E, chunk):
    # items of 512 j (exact mode's default for large launches) and of an
    # unaligned 1000 j split the same t=24, k=11 workload differently; every
    # pair term and the exact total are the same as with 256
    from circuitsimulator_amd.engine import Instance, Projector
    from circuitsimulator_amd.workloads import commuting_paulis, random_L
    rng = np.random.default_rng(1002)
    t, r, k = 24, 12, 11
    ph, xs, zs = commuting_paulis(rng, t, r)
    L = random_L(rng, t, k)
    inst = Instance()
    inst.P = Projector.make(ph, xs, zs, t)
    inst.exact = 0
    inst.k = k
    for i, v in enumerate(L):
        inst.L[i] = v
    chi = 1 << k
    nl = chi * (chi + 1) // 2
    base, terms256, _ = E.batch_exact([inst], chunk=256, want_terms=True, nterms=nl)
    res, terms, _ = E.batch_exact([inst], chunk=chunk, want_terms=True, nterms=nl)
    assert np.array_equal(terms, terms256)
    assert (res[0].re, res[0].im, res[0].ips) == (base[0].re, base[0].im, base[0].ips)


def test_exact_terms_through_the_empty_restriction_batch(E):
    # the seed-200 instance (t=11, r=4, k=4): some terms' first restriction
    # has an empty constraint, and in one of them another column that misses
    # the free variables has right side 1; every pair term against the oracle
    from circuitsimulator_amd.engine import Instance, Projector
    from circuitsimulator_amd.workloads import commuting_paulis, random_L
    rng = np.random.default_rng(200)
    t = int(rng.integers(10, 40))
    r = int(rng.integers(1, t // 2 + 1))
    k = int(rng.integers(3, 9))
    ph, xs, zs = commuting_paulis(rng, t, r)
    L = random_L(rng, t, k)
    assert (t, r, k) == (11, 4, 4)
    inst = Instance()
    inst.P = Projector.make(ph, xs, zs, t)
    inst.exact = 0
    inst.k = k
    for i, v in enumerate(L):
        inst.L[i] = v
    chi = 1 << k
    nl = chi * (chi + 1) // 2
    res, terms, _ = E.batch_exact([inst], want_terms=True, nterms=nl)
    want = oracle_exact_terms(t, k, L, 0, ph, xs, zs)
    assert [w[0] for w in want] == list(terms)
    assert res[0].im == 0.0


@pytest.mark.parametrize("t", [61, 64])
def test_exact_unpacked_layout_all_terms(E, t):
    # exact-mode |L> records (built directly in echelon form) past the packed
    # layout's limit run the unpacked sweep; every pair term against the oracle
    from circuitsimulator_amd.engine import Instance, Projector
    rng = np.random.default_rng(6100 + t)
    r, k = 9, 4
    ph, xs, zs = random_commuting_projector(rng, t, r)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    inst = Instance()
    inst.P = Projector.make(ph, xs, zs, t)
    inst.exact = 0
    inst.k = k
    for i, v in enumerate(L):
        inst.L[i] = v
    chi = 1 << k
    nl = chi * (chi + 1) // 2
    res, terms, _ = E.batch_exact([inst], chunk=5, want_terms=True, nterms=nl)
    want = oracle_exact_terms(t, k, L, 0, ph, xs, zs)
    assert [w[0] for w in want] == list(terms)
    assert res[0].im == 0.0
