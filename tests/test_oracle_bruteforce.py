"""The oracle against what the mathematics fixes, for small n: state-vector
brute force (the representation model is itself pinned by the reference's
Unpack fixtures), closed forms of the |H> and |L> decompositions, projector
identities, Philox known-answer vectors and random-state invariants."""
import ctypes
import math

import numpy as np
import pytest

import orc


def test_philox_known_answers(orc):
    # Random123 kat_vectors for philox4x32_10
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        c = (ctypes.c_uint32 * 4)(*ctr)
        k = (ctypes.c_uint32 * 2)(*key)
        out = (ctypes.c_uint32 * 4)()
        orc.lib().orc_philox4x32_10(c, k, out)
        assert tuple(out) == want


def _check_invariants(s):
    n, k = s.n, s.k
    G = s.rows("G")
    Gb = s.rows("Gb")
    assert np.array_equal((G @ Gb.T) % 2, np.eye(n, dtype=np.int64))
    J = s.rows("J", k, k)
    assert np.array_equal(J, J.T)
    for a in range(k):
        assert J[a, a] == (1 if s.D(a) in (2, 6) else 0)


@pytest.mark.parametrize("n", [1, 2, 5, 8, 13, 24, 40, 48, 64])
def test_random_state_invariants(orc, n):
    thr = orc.thresholds(n)
    for st in range(40):
        s = orc.random_state(n, 1234, st, thr)
        _check_invariants(s)


def test_random_state_dimension_law(orc):
    # d = n - k follows eta(d) (stabilizer.c:677-716): chi-square over d in {0,1,2,>=3}
    n = 10
    thr = orc.thresholds(n)
    N = 4000
    counts = np.zeros(n + 1)
    for st in range(N):
        s = orc.random_state(n, 99, st, thr)
        counts[n - s.k] += 1
    cdf = np.array([thr[d] for d in range(n + 1)], dtype=float) / 2.0 ** 53
    pd = np.diff(np.concatenate([[0.0], cdf]))
    obs = np.array([counts[0], counts[1], counts[2], counts[3:].sum()])
    exp = N * np.array([pd[0], pd[1], pd[2], pd[3:].sum()])
    chi2 = float(np.sum((obs - exp) ** 2 / exp))
    assert chi2 < 16.3  # 3 dof, p = 0.001


def test_random_state_overlap_mean(orc):
    # E |<theta|psi>|^2 = 2^-n over uniform random stabilizer states
    n = 6
    thr = orc.thresholds(n)
    rng = np.random.default_rng(5)
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    psi /= np.linalg.norm(psi)
    vals = []
    for st in range(6000):
        th = orc.unpack(orc.random_state(n, 7, st, thr))
        vals.append(abs(np.vdot(th, psi)) ** 2)
    mean = np.mean(vals)
    sem = np.std(vals) / math.sqrt(len(vals))
    assert abs(mean - 2.0 ** -n) < 5 * sem


def _random_states(orc, n, count, seed):
    thr = orc.thresholds(n)
    return [orc.random_state(n, seed, st, thr) for st in range(count)]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8, 9])
def test_inner_product_brute_force(orc, n):
    states = _random_states(orc, n, 14, 100 + n)
    vecs = [orc.unpack(s) for s in states]
    for v in vecs:
        assert abs(np.linalg.norm(v) - 1) < 1e-9
    for a in range(len(states)):
        for b in range(len(states)):
            eps, p, m = orc.inner(states[a], states[b])
            want = np.vdot(vecs[b], vecs[a])  # <s2|s1>
            got = orc.value(eps, p, m)
            assert abs(got - want) < 1e-9
            if eps:
                assert -n <= p <= 0


def _hermitian_pauli(rng, n):
    xi = orc.rand_bits(rng, n)
    ze = orc.rand_bits(rng, n)
    par = bin(xi & ze).count("1") % 2
    m = (par + 2 * int(rng.integers(0, 2))) % 4
    return m, ze, xi


@pytest.mark.parametrize("n", [1, 2, 4, 7, 9])
def test_measure_pauli_brute_force(orc, n):
    rng = np.random.default_rng(n)
    for s in _random_states(orc, n, 25, 300 + n):
        for _ in range(3):
            m, ze, xi = _hermitian_pauli(rng, n)
            v = orc.unpack(s)
            pv = orc.pauli_matrix_action(v, n, m, ze, xi)
            proj = (v + pv) / 2
            nrm = np.linalg.norm(proj)
            t = s.copy()
            code = orc.lib().orc_measurePauli(ctypes.byref(t), m, ze, xi)
            res = {0: 0.0, 1: 1.0, 2: 2 ** -0.5}[code]
            assert abs(res - nrm) < 1e-9
            if code:
                assert np.allclose(orc.unpack(t), proj / nrm, atol=1e-9)
                s = t


def test_prepH_sums_to_H_product(orc):
    c, s_ = math.cos(math.pi / 8), math.sin(math.pi / 8)
    Hv = np.array([c, s_])
    for t in range(1, 9):
        chi = 2 ** ((t + 1) // 2)
        tot = sum(orc.unpack(orc.prepH(i, t)) for i in range(chi))
        want = Hv
        for _ in range(t - 1):
            want = np.kron(want, Hv)
        norm = 2 ** (math.floor(t / 2) / 2) * (2 * c if t % 2 else 1)
        # decompose's norm (probability.c:319-323) is ||sum_i phi_i||
        assert abs(np.linalg.norm(tot) - norm) < 1e-9
        assert abs(abs(np.vdot(want, tot)) - norm) < 1e-9


def test_prepL_is_product_of_zero_and_plus(orc):
    rng = np.random.default_rng(3)
    for t in range(1, 9):
        k = int(rng.integers(1, t + 1))
        L = [int(rng.integers(0, 2 ** t)) for _ in range(k)]
        for i in range(2 ** k):
            x = 0
            for j in range(k):
                if (i >> (k - 1 - j)) & 1:
                    x ^= L[j]
            v = np.array([1.0])
            for q in reversed(range(t)):  # index = sum_q x_q 2^q: qubit 0 least significant
                f = np.array([1, 1]) / math.sqrt(2) if (x >> q) & 1 else np.array([1.0, 0.0])
                v = np.kron(v, f)
            assert np.allclose(orc.unpack(orc.prepL(i, t, L)), v)


def random_commuting_projector(rng, t, r, max_tries=10000):
    """r commuting, independent Hermitian Paulis on t qubits (as in the
    C2 workload: uniform xi, zeta, rejected unless they commute with every
    earlier generator and raise the rank)."""
    xs, zs, ph = [], [], []
    tries = 0
    while len(xs) < r and tries < max_tries:
        tries += 1
        xi = orc.rand_bits(rng, t)
        ze = orc.rand_bits(rng, t)
        if xi == 0 and ze == 0:
            continue
        ok = all((bin(xi & z2).count("1") + bin(ze & x2).count("1")) % 2 == 0 for x2, z2 in zip(xs, zs))
        if not ok:
            continue
        # rank over GF(2) of the symplectic vectors
        vecs = [(x2 << t) | z2 for x2, z2 in zip(xs, zs)] + [(xi << t) | ze]
        rows = []
        for v in vecs:
            for b in rows:
                v = min(v, v ^ b)
            if v:
                rows.append(v)
        if len(rows) < len(vecs):
            continue
        par = bin(xi & ze).count("1") % 2
        ph.append((par + 2 * int(rng.integers(0, 2))) % 4)
        xs.append(xi)
        zs.append(ze)
    return ph, xs, zs


def brute_projector_norm(psi, t, ph, xs, zs):
    v = psi
    for m, xi, ze in zip(ph, xs, zs):
        v = (v + orc.pauli_matrix_action(v, t, m, ze, xi)) / 2
    return np.vdot(psi, v).real


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("t", [2, 3, 5, 6])
def test_exact_projector_sum_brute_force(orc, t, exact):
    # |sum over i<=j terms| = <psi|Pi|psi>, psi = sum_i phi_i (docs.tex:191-193)
    rng = np.random.default_rng(10 * t + exact)
    for trial in range(3):
        r = int(rng.integers(1, t + 1))
        ph, xs, zs = random_commuting_projector(rng, t, r)
        P = orc.make_projector(ph, xs, zs, t)
        if exact:
            chi = 2 ** ((t + 1) // 2)
            phis = [orc.prepH(i, t) for i in range(chi)]
            L = [0]
            k = 0
        else:
            k = int(rng.integers(1, t + 1))
            L = [int(rng.integers(0, 2 ** t)) for _ in range(k)]
            chi = 2 ** k
            phis = [orc.prepL(i, t, L) for i in range(chi)]
        psi = sum(orc.unpack(p) for p in phis)
        want = brute_projector_norm(psi, t, ph, xs, zs)
        Larr = (orc.U64 * max(1, len(L)))(*L)
        tot = 0
        for l in range(chi * (chi + 1) // 2):
            i, j = orc.U64(), orc.U64()
            ip = orc.IpRes()
            s = ctypes.c_int()
            z = orc.lib().orc_exact_term(l, ctypes.byref(P), Larr, k, exact, ctypes.byref(i),
                                         ctypes.byref(j), ctypes.byref(ip), ctypes.byref(s))
            if z:
                continue
            v = orc.value(ip.eps, ip.p, ip.m) * 2 ** (-s.value / 2)
            tot += v if i.value == j.value else 2 * v.real
        assert abs(abs(tot) - abs(want)) < 1e-9 * max(1, abs(want))


def test_zomega_accumulator_exact(orc):
    # Z[omega] fixed point reproduces eps 2^{p/2} omega^m exactly
    rng = np.random.default_rng(1)
    E = 40
    for _ in range(200):
        c = (orc.I64 * 4)()
        tot = 0
        for _ in range(5):
            p = -int(rng.integers(0, 2 * E))
            m = int(rng.integers(0, 16))
            orc.lib().orc_zomega_add(c, E, 1, p, m)
            tot += orc.value(1, p, m)
        w = orc.OMEGA
        got = (c[0] + c[1] * w + c[2] * w ** 2 + c[3] * w ** 3) / 2 ** E
        assert abs(got - tot) < 1e-12 * max(1, abs(tot))
