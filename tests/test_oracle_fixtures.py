"""The oracle against the golden vectors the reference's own tests hold
(tests/units/tests-c/*.txt, read as in tests/units/stabtests.c:38-580, and
tests/units/tests-python/QDtests.json, Unpack*.txt).  Records are checked by
the reference test's own rules (which fields, and how strictly)."""
import ctypes

import numpy as np
import pytest

import orc


def _matrix(line, n):
    v = orc.parse_list(line)
    assert len(v) == n * n
    return [[v[i * n + j] for j in range(n)] for i in range(n)]


def test_exponential_sum_fixture(orc):
    # stabtests.c:38-103: 8 lines per record; eps, p exact, m mod 8
    L = orc.read_fixture_lines("exponentialSumTests.txt.gz")
    nrec = 0
    for r in range(len(L) // 8):
        rec = L[r * 8:(r + 1) * 8]
        n, k, Q = int(rec[0]), int(rec[1]), int(rec[2])
        D = orc.parse_list(rec[3])
        J = _matrix(rec[4], n)
        eps0, p0, m0 = int(rec[5]), int(rec[6]), int(rec[7])
        G = np.eye(n, dtype=int)
        s = orc.state_from_arrays(n, k, Q, [0] * n, D, G, G, J)
        eps, p, m = orc.exp_sum(s)
        assert (eps, p, m % 8) == (eps0, p0, m0 % 8), f"record {r}"
        nrec += 1
    assert nrec == 274


def test_exponential_sum_fixture_brute_force(orc):
    # independent check of the same records for n <= 18: direct 2^k sum
    L = orc.read_fixture_lines("exponentialSumTests.txt.gz")
    checked = 0
    for r in range(len(L) // 8):
        rec = L[r * 8:(r + 1) * 8]
        n, k, Q = int(rec[0]), int(rec[1]), int(rec[2])
        if k > 16:
            continue
        D = [int(x) for x in orc.parse_list(rec[3])]
        J = _matrix(rec[4], n)
        eps0, p0, m0 = int(rec[5]), int(rec[6]), int(rec[7])
        z = np.arange(2 ** k)
        bits = ((z[:, None] >> np.arange(k)[None, :]) & 1).astype(np.int64)
        Jm = (np.array(J, dtype=np.int64)[:k, :k] > 0).astype(np.int64)
        np.fill_diagonal(Jm, 0)
        ph = Q + bits @ np.array(D[:k]) + 2 * np.einsum("za,ab,zb->z", bits, Jm, bits)
        tot = np.sum(orc.OMEGA ** (ph % 8))
        assert abs(tot - orc.value(eps0, p0, m0)) < 1e-6 * max(1, abs(tot)), f"record {r}"
        checked += 1
    assert checked > 50


def test_shrink_fixture(orc):
    # stabtests.c:105-255: 18 lines; status, k, Q, h, D[<k], full G, Gbar, J
    L = orc.read_fixture_lines("shrinkTests.txt.gz")
    nrec = 0
    for r in range(len(L) // 18):
        rec = L[r * 18:(r + 1) * 18]
        n, k, Q, alpha = (int(x) for x in rec[0:4])
        h = orc.parse_list(rec[4])
        D = orc.parse_list(rec[5])
        xi = orc.bits_to_u64(orc.parse_list(rec[6]))
        G, Gb, J = _matrix(rec[7], n), _matrix(rec[8], n), _matrix(rec[9], n)
        st0, k1, Q1 = int(rec[10]), int(rec[11]), int(rec[12])
        h1 = orc.parse_list(rec[13])
        D1 = orc.parse_list(rec[14])
        G1, Gb1, J1 = _matrix(rec[15], n), _matrix(rec[16], n), _matrix(rec[17], n)
        s = orc.state_from_arrays(n, k, Q, h, D, G, Gb, J)
        status = orc.lib().orc_shrink(ctypes.byref(s), xi, alpha, 0)
        out = orc.state_from_arrays(n, k1, Q1, h1, D1, G1, Gb1, J1)
        assert status == st0, f"record {r}"
        assert orc.same_state(s, out, full=True), f"record {r}"
        nrec += 1
    assert nrec == 100


def test_measure_pauli_fixture(orc):
    # stabtests.c:401-580: 19 lines; result (one-sided 1e-4 in the reference;
    # two-sided here), k, Q, h, D[<k], full G, Gbar, J on k x k
    L = orc.read_fixture_lines("measurePauliTests.txt.gz")
    nrec = 0
    for r in range(len(L) // 19):
        rec = L[r * 19:(r + 1) * 19]
        n, k, Q, m = (int(x) for x in rec[0:4])
        h = orc.parse_list(rec[4])
        D = orc.parse_list(rec[5])
        zeta = orc.bits_to_u64(orc.parse_list(rec[6]))
        xi = orc.bits_to_u64(orc.parse_list(rec[7]))
        G, Gb, J = _matrix(rec[8], n), _matrix(rec[9], n), _matrix(rec[10], n)
        res0 = float(rec[11])
        k1, Q1 = int(rec[12]), int(rec[13])
        h1 = orc.parse_list(rec[14])
        D1 = orc.parse_list(rec[15])
        G1, Gb1, J1 = _matrix(rec[16], n), _matrix(rec[17], n), _matrix(rec[18], n)
        s = orc.state_from_arrays(n, k, Q, h, D, G, Gb, J)
        code = orc.lib().orc_measurePauli(ctypes.byref(s), m, zeta, xi)
        res = {0: 0.0, 1: 1.0, 2: 2 ** -0.5}[code]
        out = orc.state_from_arrays(n, k1, Q1, h1, D1, G1, Gb1, J1)
        assert abs(res - res0) < 1e-4, f"record {r}"
        assert orc.same_state(s, out, full=False), f"record {r}"
        nrec += 1
    assert nrec == 30


def test_extend_fixture(orc):
    # extendTests.txt (8 lines, no C consumer in the reference): k, xi, G, Gbar -> k, G, Gbar
    L = orc.read_fixture_lines("extendTests.txt.gz")
    nrec = 0
    for r in range(len(L) // 8):
        rec = L[r * 8:(r + 1) * 8]
        n, k = int(rec[0]), int(rec[1])
        xi = orc.bits_to_u64(orc.parse_list(rec[2]))
        G, Gb = _matrix(rec[3], n), _matrix(rec[4], n)
        k1 = int(rec[5])
        G1, Gb1 = _matrix(rec[6], n), _matrix(rec[7], n)
        s = orc.state_from_arrays(n, k, 0, [0] * n, [], G, Gb, np.zeros((n, n)))
        orc.lib().orc_extend(ctypes.byref(s), xi)
        assert s.k == k1, f"record {r}"
        for i in range(n):
            assert s.G[i] == orc.bits_to_u64(G1[i]), f"record {r} G row {i}"
            assert s.Gb[i] == orc.bits_to_u64(Gb1[i]), f"record {r} Gbar row {i}"
        nrec += 1
    assert nrec == 10


def test_inner_product_fixture(orc):
    # stabtests.c:258-398 on the surviving record: (1, -40, 4)
    L = orc.read_fixture_lines("innerProductTests-backup.txt.gz")
    rec = L[:19]

    def st(off):
        n, k, Q = int(rec[off]), int(rec[off + 1]), int(rec[off + 2])
        h = orc.parse_list(rec[off + 3])
        D = orc.parse_list(rec[off + 4])
        G, Gb, J = _matrix(rec[off + 5], n), _matrix(rec[off + 6], n), _matrix(rec[off + 7], n)
        return orc.state_from_arrays(n, k, Q, h, D, G, Gb, J)

    s1, s2 = st(0), st(8)
    eps0, p0, m0 = int(rec[16]), int(rec[17]), int(rec[18])
    eps, p, m = orc.inner(s1, s2)
    assert (eps, p, m % 8) == (eps0, p0, m0 % 8)


def test_update_qd_fixture(orc):
    # QDtests.json (99 records, updateQD in/out; no consumer in the reference)
    data = orc.load_json_fixture("QDtests.json.gz")
    assert len(data) == 99
    for r, rec in enumerate(data):
        a, b = rec["in"], rec["out"]
        n, k = int(a["n"]), int(a["k"])
        Jin = np.zeros((n, n))
        Jk = np.array(a["J"], dtype=float).reshape(-1)
        kk = int(round(len(Jk) ** 0.5))
        Jin[:kk, :kk] = Jk.reshape(kk, kk)
        s = orc.state_from_arrays(n, k, int(a["Q"]), a["h"], a["D"], a["G"], a["Gbar"], Jin)
        y = orc.bits_to_u64(a["y"])
        orc.lib().orc_updateQD(ctypes.byref(s), y)
        assert s.Q == int(b["Q"]) % 8, f"record {r}"
        Dout = [int(x) % 8 for x in b["D"]]
        assert s.Dlist(len(Dout))[:k] == Dout[:k], f"record {r}"


@pytest.mark.parametrize("fname", ["Unpack.txt.gz", "UnpackRevised.txt.gz"])
def test_unpack_semantics(orc, fname):
    # The brute-force state-vector model used to pin the oracle is itself
    # pinned by the reference's Unpack fixtures (state -> 2^n amplitudes).
    data = orc.load_json_fixture(fname)
    checked = 0
    for rec in data:
        ps = rec["output"]["psi_out"]
        n, k = int(ps["n"]), int(ps["k"])
        Jk = np.array(ps["J"], dtype=float).reshape(-1)
        J = np.zeros((n, n))
        kk = int(round(len(Jk) ** 0.5))
        if kk > 0:
            J[:kk, :kk] = Jk.reshape(kk, kk)
        s = orc.state_from_arrays(n, k, int(ps["Q"]), ps["h"], ps["D"], ps["G"], ps["Gbar"], J)
        psi = orc.unpack(s)
        re = np.array(rec["unpackReals"]["root"], dtype=float).reshape(-1)
        im = np.array(rec["unpackImaginaries"]["root"], dtype=float).reshape(-1)
        assert np.allclose(psi, re + 1j * im, atol=1e-6)
        checked += 1
    assert checked >= 10
