"""The autotuner's host logic: item counts against enumeration of the
sweep's work items, the candidates it times (claim caps in exact mode;
kernel and item size in sampled mode), the proxy and its cache (a stand-in
engine that records what it is asked to run; no GPU)."""
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
    sweep time of a candidate is `cost[(kernel, chunk)]`."""

    def __init__(self, cost, cus=4):
        self.cost = cost
        self.cus = cus
        self.runs = []
        self.sweep_kernels = [0, 0]

    def device_info(self):
        return self.cus, "fake"

    def set_sweep_kernel(self, kernel, mode=0):
        self.sweep_kernels[mode] = kernel

    def batch_exact(self, instances, chunk=0):
        kc = (self.sweep_kernels[1], chunk)
        self.runs.append(("exact", kc, instances))
        return [], None, self.cost[kc]

    def batch_sampled(self, instances, chunk=0):
        kc = (self.sweep_kernels[0], chunk)
        self.runs.append(("sampled", kc, instances))
        return None, [], self.cost[kc], []


def exact_array(ks, t=20):
    A = instance_array(len(ks))
    for i, k in enumerate(ks):
        A[i]["P"]["nqubits"] = t
        A[i]["P"]["nstabs"] = 1
        A[i]["k"] = k
    return A


def cost_table(sizes, best, base=2.0, slow=None):
    """sweep times of every (kernel, size) and the rule (0, 0): the kernel-2
    candidates 10% slower, the size adding a little, `best` at half, and
    `slow` entries overriding"""
    cost = {(0, 0): base}
    for kern in autotune.KERNELS:
        for c in (0,) + tuple(sizes):
            cost[(kern, c)] = base * (1.1 if kern == 2 else 1.0) + 0.001 * c / 64
    if best is not None:
        cost[best] = 0.5 * base
    cost.update(slow or {})
    return cost


def exact_cost(best=(1, 512), **kw):
    return cost_table(autotune.EXACT_CAPS, best, **kw)


def test_exact_tuner_picks_the_fastest_and_caches():
    autotune.clear_cache()
    cost = exact_cost(best=(2, 1024))
    E = FakeEngine(cost)
    A = exact_array([2, 3, 4, 5])
    timings = {}
    assert autotune.tune_exact(E, A, timings=timings) == (2, 1024)
    assert timings == cost
    # two repetitions per candidate, on the whole (small) workload, the
    # engine's rule first; the context's kernel is left as it was
    kcs = [(0, 0)] + [(k, c) for k in autotune.KERNELS for c in (0,) + autotune.EXACT_CAPS]
    assert [r[1] for r in E.runs] == [kc for kc in kcs for _ in range(2)]
    assert all(len(r[2]) == 4 for r in E.runs)
    assert E.sweep_kernels == [0, 0]
    n = len(E.runs)
    assert autotune.tune_exact(E, A) == (2, 1024)
    assert len(E.runs) == n
    # a list of Instance takes the same key
    lst = [autotune_instance(k) for k in [2, 3, 4, 5]]
    assert autotune.tune_exact(E, lst) == (2, 1024)
    assert len(E.runs) == n


def autotune_instance(k, t=20):
    from circuitsimulator_amd.engine import Instance
    inst = Instance()
    inst.P = Projector.make([0], [1], [0], t)
    inst.k = k
    return inst


def test_exact_proxy_keeps_the_largest_instances():
    autotune.clear_cache()
    E = FakeEngine(exact_cost(), cus=1)
    ks = [2] * 50 + [12] + [3] * 50 + [11]
    A = exact_array(ks, t=30)
    autotune.tune_exact(E, A, items_per_wave=2)
    proxy = E.runs[0][2]
    # target 2 items x 28 waves; the k=12 instance alone has more
    assert 0 < len(proxy) < len(A)
    assert 12 in list(proxy["k"])
    target = 2 * 1 * 28
    got = sum(autotune.exact_items(1 << int(k), 256) for k in proxy["k"])
    assert got >= target


def test_exact_key_separates_shapes():
    autotune.clear_cache()
    E = FakeEngine(exact_cost())
    autotune.tune_exact(E, exact_array([4] * 3, t=20))
    n = len(E.runs)
    # wider than the packed layout: another key
    autotune.tune_exact(E, exact_array([4] * 3, t=62))
    assert len(E.runs) == 2 * n
    # the |H^t> decomposition: another key
    A = exact_array([0] * 3, t=8)
    A["exact"] = 1
    autotune.tune_exact(E, A)
    assert len(E.runs) == 3 * n


def sampled_cost(best=(2, 384), **kw):
    return cost_table(autotune.CANDIDATES, best, **kw)


def test_sampled_proxy_scales_counts():
    autotune.clear_cache()
    E = FakeEngine(sampled_cost(best=(1, 512)), cus=1)
    A = sampled_instance_array(3)
    A["P"]["nqubits"] = 24
    A["P"]["nstabs"] = 1
    A["k"] = [10, 8, 6]
    A["seed"] = [5, 6, 7]
    A["first"] = [0, 100, 200]
    A["count"] = [4096, 1024, 3]
    timings = {}
    assert autotune.tune_sampled(E, A, items_per_wave=4, timings=timings) == (1, 512)
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


def test_sampled_tuner_times_every_kernel_and_size_and_restores_the_kernel():
    autotune.clear_cache()
    E = FakeEngine(sampled_cost(best=(2, 384)))
    E.sweep_kernels[0] = 1
    A = sampled_instance_array(2)
    A["P"]["nqubits"] = 20
    A["P"]["nstabs"] = 1
    A["k"] = [6, 7]
    A["count"] = [10, 20]
    timings = {}
    assert autotune.tune_sampled(E, A, timings=timings) == (2, 384)
    assert sorted(timings) == sorted([(0, 0)] + [(k, c) for k in autotune.KERNELS
                                                 for c in (0,) + autotune.CANDIDATES])
    assert E.sweep_kernels[0] == 1
    # a candidate within the tolerance of the engine's rule leaves the rule
    autotune.clear_cache()
    assert autotune.tune_sampled(FakeEngine(sampled_cost(best=None, slow={(2, 384): 1.97})), A) == (0, 0)


def test_the_engine_rule_stays_within_the_tolerance():
    # a candidate that wins by less than tol is noise: (0, 0), the engine's rule
    autotune.clear_cache()
    cost = exact_cost(best=None, slow={(1, 512): 1.97})
    assert autotune.tune_exact(FakeEngine(cost), exact_array([3, 4])) == (0, 0)
    autotune.clear_cache()
    assert autotune.tune_exact(FakeEngine(cost), exact_array([3, 4]), tol=0.0) == (1, 512)
    autotune.clear_cache()
    cost = exact_cost(best=None, slow={(2, 256): 1.7, (1, 512): 1.8})
    assert autotune.tune_exact(FakeEngine(cost), exact_array([3, 4])) == (2, 256)


def test_engine_rejects_unknown_chunk_words():
    from circuitsimulator_amd.engine import Engine
    E = Engine.__new__(Engine)
    with pytest.raises(ValueError, match="auto"):
        E.batch_exact(exact_array([2]), chunk="fast")
    with pytest.raises(ValueError, match="auto"):
        E.batch_sampled(sampled_instance_array(1), chunk="fast")
