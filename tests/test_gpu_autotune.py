"""The item-size autotuner on the device: its choice is one of the candidates,
every candidate was timed, and a launch with the tuned size gives the same
exact results as the engine's default size, bit for bit.  Each of the
engine's launches here is forced with a fixed size, so they also run every
candidate against the default on the same inputs."""
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
    c = autotune.tune_exact_chunk(E, inst, timings=timings)
    assert c in (0,) + autotune.CANDIDATES
    assert sorted(timings) == [0] + sorted(autotune.CANDIDATES) and all(v > 0 for v in timings.values())
    # a fixed size is chosen only when it beats the engine's rule by 3%
    assert c == 0 or timings[c] * 1.03 < timings[0]
    tuned, _, _ = E.batch_exact(inst, chunk="auto")
    for a, b in zip(base, tuned):
        assert (a.re, a.im, a.value, a.terms, a.ips, a.pmin) == (b.re, b.im, b.value, b.terms, b.ips, b.pmin)


def test_tuned_sampled_batch_matches_default(E):
    from circuitsimulator_amd import autotune
    from circuitsimulator_amd.engine import sampled_instance_array
    from circuitsimulator_amd.workloads import sampled_config
    autotune.clear_cache()
    A = sampled_instance_array(3)
    for i, (t, r, k, seed, count) in enumerate([(24, 8, 9, 5, 3000), (30, 10, 6, 6, 500), (16, 4, 0, 7, 40)]):
        P, _, L = sampled_config(t, r, k, seed=300 + i)
        A[i]["P"] = np.frombuffer(bytes(P), dtype=A.dtype["P"])[0]
        A[i]["k"] = k
        A[i]["exact"] = 1 if k == 0 else 0
        A[i]["L"][:len(L)] = L
        A[i]["seed"], A[i]["first"], A[i]["count"] = seed, 10 * i, count
    m0, r0, _, _ = E.batch_sampled(A)
    timings = {}
    c = autotune.tune_sampled_chunk(E, A, timings=timings)
    assert c in (0,) + autotune.CANDIDATES and sorted(timings) == [0] + sorted(autotune.CANDIDATES)
    m1, r1, _, _ = E.batch_sampled(A, chunk="auto")
    assert np.array_equal(m0, m1)
    assert [(a.ips, a.vanished) for a in r0] == [(b.ips, b.vanished) for b in r1]
