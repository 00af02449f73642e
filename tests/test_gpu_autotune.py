"""The autotuner on the device: its choice is one of the candidates, every
candidate was timed, and a launch with the tuned settings gives the same
exact results as the engine's defaults, bit for bit.  The tuner's own
launches run every candidate (claim caps; sweep kernels and item sizes) on
the same inputs, and both sampled-mode sweep kernels are compared term by
term."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from circuitsimulator_amd import Engine
    e = Engine(0)
    yield e
    e.close()


def test_tuned_exact_batch_matches_default(E):
    from circuitsimulator_amd import autotune
    from circuitsimulator_amd.workloads import uneven_instances
    autotune.clear_cache()
    inst, _ = uneven_instances(40, seed=77, kmax=10)
    base, _, _ = E.batch_exact(inst)
    timings = {}
    kern, c = autotune.tune_exact(E, inst, timings=timings)
    assert (kern, c) == (0, 0) or (kern in autotune.KERNELS and c in (0,) + autotune.EXACT_CAPS)
    assert len(timings) == 1 + len(autotune.KERNELS) * (1 + len(autotune.EXACT_CAPS))
    assert all(v > 0 for v in timings.values())
    # a candidate is chosen only when it beats the engine's rule by 2%
    assert (kern, c) == (0, 0) or timings[(kern, c)] * 1.02 < timings[(0, 0)]
    assert E.sweep_kernels == [0, 0]
    tuned, _, _ = E.batch_exact(inst, chunk="auto")
    for a, b in zip(base, tuned):
        assert (a.re, a.im, a.value, a.terms, a.ips, a.pmin) == (b.re, b.im, b.value, b.terms, b.ips, b.pmin)


@pytest.mark.parametrize("chunk", [0, 1024])
def test_exact_kernels_give_identical_terms(E, chunk):
    # both exact-mode sweeps (seven and eight waves per SIMD): the same
    # per-term results and exact totals
    from circuitsimulator_amd.workloads import uneven_instances
    inst, specs = uneven_instances(24, seed=78, kmax=9)
    nterms = sum((1 << k) * ((1 << k) + 1) // 2 for (_, _, k, *_r) in specs)
    out = []
    for kern in (1, 2):
        E.set_sweep_kernel(kern, mode=1)
        try:
            out.append(E.batch_exact(inst, chunk=chunk, want_terms=True, nterms=nterms))
        finally:
            E.set_sweep_kernel(0, mode=1)
    (ra, ta, _), (rb, tb, _) = out
    assert np.array_equal(ta, tb)
    for a, b in zip(ra, rb):
        assert (a.re, a.im, a.value, a.terms, a.ips) == (b.re, b.im, b.value, b.terms, b.ips)


def _sampled_instances():
    from circuitsimulator_amd.engine import sampled_instance_array
    from circuitsimulator_amd.workloads import sampled_config
    A = sampled_instance_array(3)
    for i, (t, r, k, seed, count) in enumerate([(24, 8, 9, 5, 3000), (30, 10, 6, 6, 500), (16, 4, 0, 7, 40)]):
        P, _, L = sampled_config(t, r, k, seed=300 + i)
        A[i]["P"] = np.frombuffer(bytes(P), dtype=A.dtype["P"])[0]
        A[i]["k"] = k
        A[i]["exact"] = 1 if k == 0 else 0
        A[i]["L"][:len(L)] = L
        A[i]["seed"], A[i]["first"], A[i]["count"] = seed, 10 * i, count
    return A


def test_tuned_sampled_batch_matches_default(E):
    from circuitsimulator_amd import autotune
    autotune.clear_cache()
    A = _sampled_instances()
    m0, r0, _, s0 = E.batch_sampled(A, want_samples=True)
    timings = {}
    kern, c = autotune.tune_sampled(E, A, timings=timings)
    assert (kern, c) == (0, 0) or (kern in autotune.KERNELS and c in (0,) + autotune.CANDIDATES)
    assert len(timings) == 1 + len(autotune.KERNELS) * (1 + len(autotune.CANDIDATES))
    assert E.sweep_kernels == [0, 0]
    m1, r1, _, s1 = E.batch_sampled(A, chunk="auto", want_samples=True)
    assert E.sweep_kernels == [0, 0]
    assert np.array_equal(m0, m1)
    assert [(a.ips, a.vanished) for a in r0] == [(b.ips, b.vanished) for b in r1]
    for a, b in zip(s0, s1):
        assert np.array_equal(a["zsum"], b["zsum"])


@pytest.mark.parametrize("t,k,exact", [(24, 9, 0), (48, 12, 0), (20, 0, 1)])
def test_sweep_kernels_give_identical_terms(E, t, k, exact):
    # both sampled-mode sweep kernels: the same per-term results, per-sample
    # sums and exponents, for |L> records (a share of them with three or more
    # non-pivot columns) and |H^t> terms
    from circuitsimulator_amd.workloads import sampled_config
    P, _, L = sampled_config(t, 10, k, seed=900 + t)
    out = []
    for kern in (1, 2):
        E.set_sweep_kernel(kern)
        try:
            res, ex = E.sampled_projector(P, L, k, exact, 55, 0, 64, want_terms=True)
        finally:
            E.set_sweep_kernel(0)
        out.append((res, ex))
    (a, ea), (b, eb) = out
    assert np.array_equal(ea["terms"], eb["terms"])
    assert np.array_equal(ea["zsum"], eb["zsum"]) and np.array_equal(ea["sexp"], eb["sexp"])
    assert a.mean == b.mean and a.ips == b.ips
