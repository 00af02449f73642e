"""Exact-mode flushes of a wave's partial sums (cse_kernels.hip sweep_body).

A wave keeps its LDS histogram of one instance (32-bit bins) across its
segments and flushes it to the instance's int64 histogram when it moves to
another instance, at its exit (merged per block), and once it has folded
CSE_FLUSH_TERMS terms of the instance (read when a context is created;
2^30 by default, which keeps every bin below 2^31 since a term adds at most 2
to one bin).  Bounds of 64 and 1000 terms force many flushes within one
instance: the histograms, totals and term counts must not change."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(kern, P, L, k, inst):
    from circuitsimulator_amd import Engine
    with Engine(0) as E:
        E.set_sweep_kernel(kern, mode=1)
        r, h = E.exact_projector(P, L, k, 0, want_hist=True)
        rb, _, _ = E.batch_exact(inst)
        return (r.re, r.im, r.ips), h.copy(), [(a.re, a.im, a.value, a.terms, a.ips) for a in rb]


def test_partial_flushes_give_identical_histograms(monkeypatch):
    from circuitsimulator_amd.workloads import sampled_config, uneven_instances
    P, _, L = sampled_config(24, 12, 10, seed=1002)  # one instance, chi = 1024
    inst, _ = uneven_instances(24, seed=79, kmax=9)
    monkeypatch.delenv("CSE_FLUSH_TERMS", raising=False)
    base = {kern: _run(kern, P, L, 10, inst) for kern in (1, 2)}
    assert base[1][0] == base[2][0] and np.array_equal(base[1][1], base[2][1]) and base[1][2] == base[2][2]
    assert base[1][0][2] == 1024 * 1025 // 2
    for bound in ("64", "1000"):
        monkeypatch.setenv("CSE_FLUSH_TERMS", bound)
        for kern in (1, 2):
            tot, h, batch = _run(kern, P, L, 10, inst)
            print(f"CSE_FLUSH_TERMS={bound} kernel {kern}", flush=True)
            assert tot == base[1][0]
            assert np.array_equal(h, base[1][1])
            assert batch == base[1][2]


@pytest.mark.parametrize("bound", ["0", "-5", str((1 << 30) + 1)])
def test_flush_bound_outside_its_range_is_refused(monkeypatch, bound):
    from circuitsimulator_amd import Engine
    from circuitsimulator_amd.engine import EngineError
    monkeypatch.setenv("CSE_FLUSH_TERMS", bound)
    with pytest.raises(EngineError, match="CSE_FLUSH_TERMS"):
        Engine(0)
