"""The item-size autotuner's host logic: item counts against enumeration of
the sweep's work items, the proxy it times and its cache (a stand-in engine
that records what it is asked to run; no GPU)."""
import numpy as np
import pytest

from circuitsimulator_amd import autotune
from circuitsimulator_amd.engine import Projector, instance_array, sampled_instance_array


def enumerate_exact_items(chi, chunk):
    # record i keeps the chunks of j whose last j reaches the diagonal j >= i
    n = 0
    for i in range(chi):
        for c in range(-(-chi // chunk)):
            last = min((c + 1) * chunk, chi) - 1
            if last >= i:
                n += 1
    return n


@pytest.mark.parametrize("chi", [1, 2, 3, 8, 64, 100, 256, 300, 1024])
@pytest.mark.parametrize("chunk", [1, 7, 64, 256, 1000])
def test_exact_items_match_enumeration(chi, chunk):
    assert autotune.exact_items(chi, chunk) == enumerate_exact_items(chi, chunk)


def test_exact_items_at_the_two_point_workload():
    # chi = 2048 in items of 256: 8 chunks a record, one partial per record
    assert autotune.exact_items(2048, 256) == 9216


class FakeEngine:
    """Answers device_info and records each batch it is asked to time; the
    sweep time of a candidate is `cost[chunk]`."""

    def __init__(self, cost, cus=4):
        self.cost = cost
        self.cus = cus
        self.runs = []

    def device_info(self):
        return self.cus, "fake"

    def batch_exact(self, instances, chunk=0):
        self.runs.append(("exact", chunk, instances))
        return [], None, self.cost[chunk]

    def batch_sampled(self, instances, chunk=0):
        self.runs.append(("sampled", chunk, instances))
        return None, [], self.cost[chunk], []


def exact_array(ks, t=20):
    A = instance_array(len(ks))
    for i, k in enumerate(ks):
        A[i]["P"]["nqubits"] = t
        A[i]["P"]["nstabs"] = 1
        A[i]["k"] = k
    return A


COST = {0: 2.0, 128: 3.0, 192: 2.6, 256: 2.0, 384: 1.8, 512: 1.5, 1024: 2.5}


def test_exact_tuner_picks_the_fastest_and_caches():
    autotune.clear_cache()
    E = FakeEngine(COST)
    A = exact_array([2, 3, 4, 5])
    timings = {}
    assert autotune.tune_exact_chunk(E, A, timings=timings) == 512
    assert timings == COST
    # two repetitions per candidate, on the whole (small) workload
    assert [r[1] for r in E.runs] == [0, 0, 128, 128, 192, 192, 256, 256, 384, 384, 512, 512, 1024, 1024]
    assert all(len(r[2]) == 4 for r in E.runs)
    n = len(E.runs)
    assert autotune.tune_exact_chunk(E, A) == 512
    assert len(E.runs) == n
    # a list of Instance takes the same key
    lst = [autotune_instance(k) for k in [2, 3, 4, 5]]
    assert autotune.tune_exact_chunk(E, lst) == 512
    assert len(E.runs) == n


def autotune_instance(k, t=20):
    from circuitsimulator_amd.engine import Instance
    inst = Instance()
    inst.P = Projector.make([0], [1], [0], t)
    inst.k = k
    return inst


def test_exact_proxy_keeps_the_largest_instances():
    autotune.clear_cache()
    E = FakeEngine(COST, cus=1)
    ks = [2] * 50 + [12] + [3] * 50 + [11]
    A = exact_array(ks, t=30)
    autotune.tune_exact_chunk(E, A, items_per_wave=2)
    proxy = E.runs[0][2]
    # target 2 items x 28 waves; the k=12 instance alone has more
    assert 0 < len(proxy) < len(A)
    assert 12 in list(proxy["k"])
    target = 2 * 1 * 28
    got = sum(autotune.exact_items(1 << int(k), 256) for k in proxy["k"])
    assert got >= target


def test_exact_key_separates_shapes():
    autotune.clear_cache()
    E = FakeEngine(COST)
    autotune.tune_exact_chunk(E, exact_array([4] * 3, t=20))
    n = len(E.runs)
    # wider than the packed layout: another key
    autotune.tune_exact_chunk(E, exact_array([4] * 3, t=62))
    assert len(E.runs) == 2 * n
    # the |H^t> decomposition: another key
    A = exact_array([0] * 3, t=8)
    A["exact"] = 1
    autotune.tune_exact_chunk(E, A)
    assert len(E.runs) == 3 * n


def test_sampled_proxy_scales_counts():
    autotune.clear_cache()
    E = FakeEngine(COST, cus=1)
    A = sampled_instance_array(3)
    A["P"]["nqubits"] = 24
    A["P"]["nstabs"] = 1
    A["k"] = [10, 8, 6]
    A["seed"] = [5, 6, 7]
    A["first"] = [0, 100, 200]
    A["count"] = [4096, 1024, 3]
    timings = {}
    assert autotune.tune_sampled_chunk(E, A, items_per_wave=4, timings=timings) == 512
    proxy = E.runs[0][2]
    assert proxy.flags["C_CONTIGUOUS"]
    assert list(proxy["seed"]) == [5, 6, 7] and list(proxy["first"]) == [0, 100, 200]
    total = sum(autotune.sampled_items(1 << int(k), 256, int(c)) for k, c in zip(A["k"], A["count"]))
    target = 4 * 1 * 32
    share = target / total
    assert list(proxy["count"]) == [max(1, int(np.ceil(c * share))) for c in A["count"]]
    assert all(c >= 1 for c in proxy["count"])
    # the caller's array is untouched
    assert list(A["count"]) == [4096, 1024, 3]


def test_the_engine_rule_stays_within_the_tolerance():
    # a fixed size that wins by less than tol is noise: 0 (the engine's rule)
    autotune.clear_cache()
    E = FakeEngine({0: 2.0, 128: 2.1, 192: 2.05, 256: 2.0, 384: 1.96, 512: 1.95, 1024: 2.2})
    assert autotune.tune_exact_chunk(E, exact_array([3, 4])) == 0
    autotune.clear_cache()
    assert autotune.tune_exact_chunk(E, exact_array([3, 4]), tol=0.0) == 512
    autotune.clear_cache()
    E = FakeEngine({0: 2.0, 128: 1.7, 192: 1.8, 256: 2.0, 384: 1.9, 512: 1.9, 1024: 2.2})
    assert autotune.tune_exact_chunk(E, exact_array([3, 4])) == 128


def test_engine_rejects_unknown_chunk_words():
    from circuitsimulator_amd.engine import Engine
    E = Engine.__new__(Engine)
    with pytest.raises(ValueError, match="auto"):
        E.batch_exact(exact_array([2]), chunk="fast")
    with pytest.raises(ValueError, match="auto"):
        E.batch_sampled(sampled_instance_array(1), chunk="fast")
