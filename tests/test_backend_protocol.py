"""The drop-in boundary: the backend executable's stdin/stdout contract
(libcirc/probability.c:44-221 <-> libcirc/probability.py:216-312), its
host-side rules and its failure behaviour.  The CPU tests need no GPU: they
cover parsing, the closed forms that never reach the engine, and that any
projector work fails loudly (no CPU fallback) when there is no gfx950 GPU.
The GPU tests run the executable end to end against the oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

from circuitsimulator_amd import protocol
from circuitsimulator_amd.engine import BACKEND_PATH

pytestmark = pytest.mark.skipif(not os.path.exists(BACKEND_PATH), reason="cse_backend not built")

EMPTY = ([], [], [])


def _gpu_present():
    return os.path.exists("/dev/kfd")


def test_empty_projectors_give_norm_squared():
    # exactProjector: Nstabs == 0 -> norm^2 (innerprod.c:150); decompose's
    # exact norm 2^{floor(t/2)/2} (2 cos(pi/8) for odd t) (probability.c:319-323)
    for t in (1, 2, 3, 6, 7):
        payload = protocol.make_payload(EMPTY, EMPTY, {"noapprox": True}, t=t)
        num, den, info = protocol.run_backend(payload)
        norm = 2 ** (math.floor(t / 2) / 2) * (2 * math.cos(math.pi / 8) if t % 2 else 1)
        assert abs(num - norm ** 2) < 1e-12 * norm ** 2
        assert num == den
        assert info[0] == "C backend print mode is on."


def test_clifford_projectors_closed_form():
    # t = 0: (1 + #{+I} - #{-I}) / (1 + Nstabs) (innerprod.c:157-167)
    G = ([0, 2, 2], [[], [], []], [[], [], []])
    H = ([0], [[]], [[]])
    payload = "0\n0\n1\n2000\n1\n0\n0\n1\n1e-05\n0\n0\n0\n0\n" + "3\n0\n0\n2\n2\n" + "1\n0\n0\n"
    num, den, _ = protocol.run_backend(payload)
    assert num == pytest.approx((1 + 1 - 2) / 4)
    assert den == pytest.approx((1 + 1) / 2)


def test_payload_file_argument(tmp_path):
    payload = protocol.make_payload(EMPTY, EMPTY, {"noapprox": True}, t=2)
    f = tmp_path / "payload.txt"
    f.write_text(payload)
    num, den, info = protocol.run_backend("", argv=(str(f),))
    assert num == pytest.approx(2.0)
    assert any("Reading arguments from file" in ln for ln in info)


def test_source_name_from_stdin_when_no_argv():
    # probability.c:54 -- with no argument the first token names the source
    payload = "stdin\n" + protocol.make_payload(EMPTY, EMPTY, {"noapprox": True}, t=2)
    num, den, _ = protocol.run_backend(payload, argv=())
    assert num == pytest.approx(2.0)


@pytest.mark.parametrize("payload", ["", "0 0 1", "0 0 1 2000 1 3 0 1 x", "0 0 1 2000 1 3 0 1 1e-5 0 0 0 0 2 3 0 1"])
def test_malformed_input_fails_loudly(payload):
    p = subprocess.run([BACKEND_PATH, "stdin"], input=payload.encode(), capture_output=True)
    assert p.returncode == 2
    last = p.stdout.decode().splitlines()[-1]
    assert last.startswith("Error:")
    with pytest.raises(RuntimeError):
        protocol.run_backend(payload)


def test_missing_payload_file():
    p = subprocess.run([BACKEND_PATH, "/nonexistent/payload"], capture_output=True)
    assert p.returncode == 2


def test_wide_projector_rejected():
    G = ([0], [[0] * 65], [[1] * 65])
    payload = protocol.make_payload(G, G, {"noapprox": True}, t=65)
    p = subprocess.run([BACKEND_PATH, "stdin"], input=payload.encode(), capture_output=True)
    assert p.returncode == 3


def test_nonzero_mpi_ranks_exit_quietly():
    for var in ("PMI_RANK", "OMPI_COMM_WORLD_RANK"):
        env = dict(os.environ, **{var: "1"})
        p = subprocess.run([BACKEND_PATH, "stdin"], input=b"", capture_output=True, env=env)
        assert p.returncode == 0 and p.stdout == b""


@pytest.mark.skipif(_gpu_present(), reason="checks the no-GPU failure path")
def test_projector_work_without_gpu_fails_loudly():
    G = ([0], [[1, 0]], [[0, 0]])
    payload = protocol.make_payload(G, EMPTY, {"noapprox": True}, t=2)
    p = subprocess.run([BACKEND_PATH, "stdin"], input=payload.encode(), capture_output=True)
    assert p.returncode == 4
    assert "no usable gfx950 GPU" in p.stdout.decode()


def test_decompose_rules():
    from circuitsimulator_amd.engine import decompose
    # auto k = ceil(1 - 2t log2 cos(pi/8) - log2 fidbound); k > t/2 reverts to exact
    t = 10
    k_auto = math.ceil(1 - 2 * t * math.log2(math.cos(math.pi / 8)) - math.log2(1e-5))
    L, norm, exact, k = decompose(t, 1, exact=0, k=0, fidbound=1e-5)
    assert k == k_auto and exact == 1
    # forced k stays, L drawn with k rows of t bits
    L, norm, exact, k = decompose(t, 1, exact=0, k=3)
    assert exact == 0 and k == 3 and len(L) == 3 and all(v < 2 ** t for v in L)
    # k > t is clamped to t
    L, norm, exact, k = decompose(4, 1, exact=0, k=9)
    assert k == 4
    # rank check: resampled until full rank
    L, norm, exact, k = decompose(12, 7, exact=0, k=6, rank=1)
    from circuitsimulator_amd.workloads import gf2_rank
    assert gf2_rank(L) == 6
    # fidelity: norm = sqrt(2^k Z_L), Z_L = sum_x 2^{-floor(|x|/2)} (probability.c:396)
    L, norm, exact, k = decompose(6, 3, exact=0, k=2, fidelity=1)
    zl = 0.0
    for i in range(4):
        x = 0
        for j in range(2):
            if (i >> (1 - j)) & 1:
                x ^= L[j]
        zl += 2.0 ** (-(bin(x).count("1") // 2))
    assert norm == pytest.approx(math.sqrt(4 * zl))


# ------------------------------------------------------------------ GPU
def _bits(v, t):
    return [(v >> j) & 1 for j in range(t)]


def _proj_lists(ph, xs, zs, t):
    return (list(ph), [_bits(x, t) for x in xs], [_bits(z, t) for z in zs])


@pytest.mark.gpu
@pytest.mark.parametrize("t", [3, 6, 8])
def test_backend_exact_mode_end_to_end(t):
    import orc
    from gpu_util import oracle_exact_terms, random_commuting_projector, term_value
    rng = np.random.default_rng(300 + t)
    phG, xG, zG = random_commuting_projector(rng, t, 3)
    phH, xH, zH = phG[:2], xG[:2], zG[:2]
    payload = protocol.make_payload(_proj_lists(phG, xG, zG, t), _proj_lists(phH, xH, zH, t),
                                    {"noapprox": True, "exact": True})
    num, den, info = protocol.run_backend(payload)

    def oracle_value(ph, xs, zs):
        tot = 0
        for (pt, s, i, j) in oracle_exact_terms(t, 0, [], 1, ph, xs, zs):
            if pt:
                v = term_value(pt, s)
                tot += v if i == j else 2 * v.real
        return abs(tot)

    assert num == pytest.approx(oracle_value(phG, xG, zG), rel=1e-14, abs=1e-14)
    assert den == pytest.approx(oracle_value(phH, xH, zH), rel=1e-14, abs=1e-14)


@pytest.mark.gpu
def test_backend_sampled_mode_matches_engine():
    # forceSample keeps sampling on (probability.c:162-174); with a fixed seed
    # the backend's numerator equals the engine's estimator on the same draws
    from circuitsimulator_amd import Engine
    from circuitsimulator_amd.engine import Projector
    from circuitsimulator_amd.workloads import commuting_paulis
    rng = np.random.default_rng(9)
    t, k, r = 12, 4, 5
    ph, xs, zs = commuting_paulis(rng, t, r)
    payload = protocol.make_payload(_proj_lists(ph, xs, zs, t), _proj_lists(ph[:2], xs[:2], zs[:2], t),
                                    {"exact": False, "k": k, "samples": 300, "bins": 1, "forceSample": True})
    env = dict(os.environ, CSE_SEED="12345")
    num, den, info = protocol.run_backend(payload, env=env)
    assert num > 0 and den > 0
    # the L drawn by decompose is printed; rebuild it and recompute
    i0 = info.index("L:")
    rows = [ln.strip().strip("[]") for ln in info[i0 + 1:i0 + 1 + k]]
    L = [sum(int(c) << j for j, c in enumerate(row)) for row in rows]
    with Engine(0) as E:
        res, _ = E.sampled_projector(Projector.make(ph, xs, zs, t), L, k, 0, 12345, 0, 300)
    assert num == pytest.approx(res.mean, rel=1e-15)


def test_decompose_many_equals_per_seed_calls():
    # one cse_decompose_many call per batch of seeds gives each seed's
    # decomposition exactly as its own cse_decompose call
    from circuitsimulator_amd.probability import decompose, decompose_many
    seeds = [3, 2 ** 61 + 5, 977, 12345678901]
    for t, cfg in [(40, {"k": 8}), (10, {}), (30, {"k": 6, "rank": True}), (24, {"k": 5, "fidelity": True}),
                   (16, {"exact": True}), (5, {"k": 9})]:
        cfg = dict(cfg, quiet=True, fidbound=1e-5)
        Ls, norm, ex, k = decompose_many(t, cfg, seeds)
        for i, sd in enumerate(seeds):
            L1, n1, e1, k1 = decompose(t, cfg, sd)
            assert (e1, k1) == (ex, k), (t, cfg)
            if ex:
                assert Ls is None
            else:
                assert [int(v) for v in Ls[i]] == L1, (t, cfg, sd)
        assert norm == decompose(t, cfg, seeds[-1])[1]


def _row_space_weights(L, k):
    xs = np.zeros(1, dtype=np.uint64)
    for row in L[:k]:
        xs = np.concatenate([xs, xs ^ np.uint64(row)])
    return np.bincount(np.bitwise_count(xs), minlength=65)


def test_decompose_fidelity_is_exact():
    # Z_L = sum over the 2^k combinations of 2^-floor(|x|/2), summed exactly
    from fractions import Fraction
    from circuitsimulator_amd.engine import decompose
    for t, k, seed in [(10, 6, 1), (16, 12, 2), (24, 16, 3), (40, 18, 4)]:
        L, norm, exact, kk = decompose(t, seed, exact=0, k=k, fidelity=1)
        assert kk == k and not exact
        hist = _row_space_weights(L, k)
        zl = sum(Fraction(int(c), 2 ** (w // 2)) for w, c in enumerate(hist))
        assert norm == pytest.approx(float((Fraction(2 ** k) * zl)) ** 0.5, rel=1e-15)


@pytest.mark.gpu
def test_lspace_weights_on_engine():
    from circuitsimulator_amd import Engine
    rng = np.random.default_rng(11)
    with Engine(0) as E:
        for k, t in [(0, 5), (1, 5), (7, 12), (8, 30), (13, 48), (18, 64), (19, 40), (22, 48)]:
            L = [int(rng.integers(0, 2 ** 62)) & ((1 << t) - 1) for _ in range(k)]
            assert list(E.lspace_weights(L, k)) == list(_row_space_weights(L, k))



@pytest.mark.gpu
def test_decompose_fidelity_device_path():
    # k >= 20 takes the device histogram inside cse_decompose
    from fractions import Fraction
    from circuitsimulator_amd.engine import decompose
    L, norm, exact, k = decompose(48, 9, exact=0, k=21, fidelity=1)
    hist = _row_space_weights(L, 21)
    zl = sum(Fraction(int(c), 2 ** (w // 2)) for w, c in enumerate(hist))
    assert norm == pytest.approx(float(Fraction(2 ** 21) * zl) ** 0.5, rel=1e-15)
