"""cse_exact_projector_noalloc (the entry point that allocates nothing, with
its buffers in the library's static device array) against
cse_exact_projector, both through the C ABI: identical exact histograms and
values on two calls in a row, the second smaller than the first, for the
|L> and the |H^t> decompositions; Kahan-FP32 shadows to their FP32 bound;
chi above the array's limit rejected on the host."""
import numpy as np
import pytest

import orc
from gpu_util import random_commuting_projector

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from circuitsimulator_amd import Engine
    e = Engine(0)
    yield e
    e.close()


def _case(seed, t, r, k):
    from circuitsimulator_amd.engine import Projector
    rng = np.random.default_rng(seed)
    ph, xs, zs = random_commuting_projector(rng, t, r)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    return Projector.make(ph, xs, zs, t), L


def _same(a, ha, b, hb):
    assert np.array_equal(ha, hb)
    assert (a.re, a.im, a.value, a.terms, a.ips, a.pmin) == (b.re, b.im, b.value, b.terms, b.ips, b.pmin)
    scale = max(1.0, abs(a.value))
    assert abs(a.kahan_re - b.kahan_re) <= 1e-5 * scale
    assert abs(a.kahan_im - b.kahan_im) <= 1e-5 * scale


# (seed, t, r, k, exact): a larger call, then a smaller one on the same array
SEQUENCES = {
    "L": [(1101, 24, 12, 11, 0), (1102, 10, 5, 4, 0)],
    "H": [(1103, 22, 9, 0, 1), (1104, 7, 3, 0, 1)],
    "L_max": [(1105, 30, 10, 13, 0), (1106, 12, 6, 6, 0)],
}


@pytest.mark.parametrize("name", sorted(SEQUENCES))
def test_noalloc_matches_exact_projector_twice(E, name):
    for seed, t, r, k, exact in SEQUENCES[name]:
        print(f"noalloc {name}: t={t} r={r} k={k} exact={exact}", flush=True)
        P, L = _case(seed, t, r, k)
        want, hw = E.exact_projector(P, L, k, exact, want_hist=True)
        got, hg = E.exact_projector_noalloc(P, L, k, exact, want_hist=True)
        _same(got, hg, want, hw)
        chi = (1 << ((t + 1) // 2)) if exact else (1 << k)
        assert got.terms == chi * (chi + 1) // 2
        assert got.im == 0.0 and got.re >= 0.0


def test_noalloc_after_a_larger_call_of_the_other_kind(E):
    # an |H^t> call after an |L> call of larger chi and the reverse: nothing of
    # the previous call's queue, sums or records leaks into the next
    for seed, t, r, k, exact in [(1107, 26, 11, 0, 1), (1108, 16, 7, 9, 0), (1109, 9, 4, 0, 1)]:
        P, L = _case(seed, t, r, k)
        want, hw = E.exact_projector(P, L, k, exact, want_hist=True)
        got, hg = E.exact_projector_noalloc(P, L, k, exact, want_hist=True)
        _same(got, hg, want, hw)


def test_noalloc_rejects_chi_above_its_limit(E):
    from circuitsimulator_amd.engine import EngineError
    P, L = _case(1110, 30, 8, 14)
    with pytest.raises(EngineError, match="no-allocation limit"):
        E.exact_projector_noalloc(P, L, 14, 0)
    P, _ = _case(1111, 28, 8, 0)
    with pytest.raises(EngineError, match="no-allocation limit"):
        E.exact_projector_noalloc(P, [], 0, 1)
    # the context and the array still work afterwards
    P, L = _case(1112, 8, 4, 3)
    want, hw = E.exact_projector(P, L, 3, 0, want_hist=True)
    got, hg = E.exact_projector_noalloc(P, L, 3, 0, want_hist=True)
    _same(got, hg, want, hw)
