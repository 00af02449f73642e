"""The oracle against the reference's own C (stabilizer.c, matrix.c,
stateprep.c compiled from the reference checkout into oracle/_ref by
oracle/Makefile) on seeded inputs at n = 8..64, full state compared."""
import ctypes

import numpy as np
import pytest

import orc

pytestmark = pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref/libref.so not built")


def _pairs(orc, n, count, seed, project=True):
    thr = orc.thresholds(n)
    rng = np.random.default_rng(seed)
    states = []
    for st in range(count):
        s = orc.random_state(n, seed, st, thr)
        if project and st % 2:
            from test_oracle_bruteforce import random_commuting_projector
            ph, xs, zs = random_commuting_projector(rng, n, int(rng.integers(1, n // 2 + 1)))
            P = orc.make_projector(ph, xs, zs, n)
            t = s.copy()
            if orc.project(t, P) >= 0:
                s = t
        states.append(s)
    return states


@pytest.mark.parametrize("n", [8, 16, 24, 40, 48, 64])
def test_inner_product_vs_reference(orc, n):
    states = _pairs(orc, n, 10 if n <= 40 else 5, 500 + n)
    rng = np.random.default_rng(n)
    # add prepL / prepH states, the engine's other operand family
    L = [orc.rand_bits(rng, n) for _ in range(min(6, n))]
    states += [orc.prepL(int(rng.integers(0, 2 ** len(L))), n, L) for _ in range(3)]
    if n <= 24:
        states += [orc.prepH(int(rng.integers(0, 2 ** ((n + 1) // 2))), n) for _ in range(3)]
    nonzero = 0
    for a in states:
        for b in states:
            e0, p0, m0 = orc.ref_inner(a, b)
            e1, p1, m1 = orc.inner(a, b)
            assert e0 == e1
            if e0:
                nonzero += 1
                assert (p0, m0 % 8) == (p1, m1 % 8)
    assert nonzero > 0


@pytest.mark.parametrize("n", [8, 24, 48])
def test_shrink_vs_reference_full_state(orc, n):
    rng = np.random.default_rng(7 * n)
    for s in _pairs(orc, n, 12, 600 + n):
        for lazy in (0, 1):
            xi = orc.rand_bits(rng, n)
            alpha = int(rng.integers(0, 2))
            a, b = s.copy(), s.copy()
            r0 = orc.ref().ref_shrink(ctypes.byref(a), xi, alpha, lazy)
            r1 = orc.lib().orc_shrink(ctypes.byref(b), xi, alpha, lazy)
            assert r0 == r1
            assert orc.same_state(a, b, full=True)
            assert a.D1 == b.D1 and a.D2 == b.D2


@pytest.mark.parametrize("n", [8, 24, 48])
def test_measure_pauli_vs_reference_full_state(orc, n):
    rng = np.random.default_rng(11 * n)
    from test_oracle_bruteforce import _hermitian_pauli
    for s in _pairs(orc, n, 12, 700 + n, project=False):
        for _ in range(4):
            m, ze, xi = _hermitian_pauli(rng, n)
            a, b = s.copy(), s.copy()
            r0 = orc.ref().ref_measurePauli(ctypes.byref(a), m, ze, xi)
            code = orc.lib().orc_measurePauli(ctypes.byref(b), m, ze, xi)
            assert abs(r0 - {0: 0.0, 1: 1.0, 2: 2 ** -0.5}[code]) < 1e-15
            assert orc.same_state(a, b, full=True)
            if code:
                s = b


@pytest.mark.parametrize("n", [8, 24, 48])
def test_exponential_sum_vs_reference(orc, n):
    for s in _pairs(orc, n, 16, 800 + n):
        a, b = s.copy(), s.copy()
        e0, p0, m0 = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        orc.ref().ref_exponentialSum(ctypes.byref(a), ctypes.byref(e0), ctypes.byref(p0), ctypes.byref(m0))
        e1, p1, m1 = orc.exp_sum(b)
        assert (e0.value, p0.value, m0.value) == (e1, p1, m1)
        assert orc.same_state(a, b, full=True)


@pytest.mark.parametrize("t", [1, 2, 5, 6, 11, 24])
def test_prep_states_vs_reference(orc, t):
    rng = np.random.default_rng(t)
    k = min(t, 8)
    L = [int(rng.integers(0, 2 ** t)) for _ in range(k)]
    Larr = (orc.U64 * k)(*L)
    for i in rng.integers(0, 2 ** k, size=6):
        a = orc.OrcState()
        orc.ref().ref_prepL(ctypes.byref(a), int(i), t, Larr, k)
        assert orc.same_state(a, orc.prepL(int(i), t, L), full=True)
    for i in rng.integers(0, 2 ** ((t + 1) // 2), size=6):
        a = orc.OrcState()
        orc.ref().ref_prepH(ctypes.byref(a), int(i), t)
        assert orc.same_state(a, orc.prepH(int(i), t), full=True)
