"""Batched calls replayed from captured HIP graphs (cse_set_graphs) against
the same calls run eagerly on another context: every replay of a captured
shape gives the eager result exactly (exact sums, per-sample sums,
counts), with eager work and synchronizations on the graph context between
the replays, several shapes alternating (more than one graph cached), and
calls with per-term output (never captured) in between."""
import numpy as np
import pytest

import orc
from gpu_util import random_commuting_projector

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engines():
    from circuitsimulator_amd import Engine
    g, e = Engine(0), Engine(0)
    g.set_graphs(True)
    yield g, e
    g.close()
    e.close()


def _instances(seed, specs):
    from circuitsimulator_amd.engine import Instance, Projector
    rng = np.random.default_rng(seed)
    out = []
    for t, r, k, exact in specs:
        ph, xs, zs = random_commuting_projector(rng, t, r)
        inst = Instance()
        inst.P = Projector.make(ph, xs, zs, t)
        inst.exact = exact
        inst.k = k
        for i in range(k):
            inst.L[i] = orc.rand_bits(rng, t)
        out.append(inst)
    return out


def _same_exact(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert (x.re, x.im, x.value, x.terms, x.ips) == (y.re, y.im, y.value, y.terms, y.ips)
        assert abs(x.kahan_re - y.kahan_re) <= 1e-5 * max(1.0, abs(x.value))


SHAPES = {
    "c1": [(6, 3, 0, 1), (6, 2, 0, 1)],               # G', H' of a C1-size probability
    "mixed": [(12, 4, 5, 0), (9, 3, 0, 1), (20, 6, 7, 0)],
}


def test_exact_batches_replayed_match_eager(engines):
    g, e = engines
    for rep in range(4):
        for name, spec in SHAPES.items():
            insts = _instances(500 + 10 * rep + len(name), spec)
            print(f"graph replay {rep} of shape {name}", flush=True)
            got, _, _ = g.batch_exact(insts)
            want, _, _ = e.batch_exact(insts)
            _same_exact(got, want)
        # eager work on the graph context between replays: a per-term call
        insts = _instances(900 + rep, SHAPES["mixed"])
        nt = sum(((1 << (i.k if not i.exact else (i.P.nqubits + 1) // 2)) *
                  ((1 << (i.k if not i.exact else (i.P.nqubits + 1) // 2)) + 1) // 2) for i in insts)
        got, terms, _ = g.batch_exact(insts, want_terms=True, nterms=nt)
        want, wterms, _ = e.batch_exact(insts, want_terms=True, nterms=nt)
        _same_exact(got, want)
        assert np.array_equal(terms, wterms)


def test_exact_projector_replayed_matches_eager(engines):
    from circuitsimulator_amd.engine import Projector
    g, e = engines
    rng = np.random.default_rng(77)
    for rep in range(5):
        ph, xs, zs = random_commuting_projector(rng, 16, 6)
        L = [orc.rand_bits(rng, 16) for _ in range(6)]
        P = Projector.make(ph, xs, zs, 16)
        a, ha = g.exact_projector(P, L, 6, 0, want_hist=True)
        b, hb = e.exact_projector(P, L, 6, 0, want_hist=True)
        assert np.array_equal(ha, hb)
        assert (a.re, a.im, a.value, a.terms, a.ips) == (b.re, b.im, b.value, b.terms, b.ips)


def test_sampled_batches_replayed_match_eager(engines):
    from circuitsimulator_amd.engine import sampled_instance_array
    from circuitsimulator_amd.workloads import commuting_paulis, random_L
    g, e = engines
    rng = np.random.default_rng(31)
    t, r, k = 20, 6, 6
    ph, xs, zs = commuting_paulis(rng, t, r)
    L = random_L(rng, t, k)
    from circuitsimulator_amd.engine import Projector
    P = Projector.make(ph, xs, zs, t)
    for rep in range(4):
        A = sampled_instance_array(3)
        for i in range(3):
            A[i]["P"] = np.frombuffer(bytes(P), dtype=A.dtype["P"])[0]
            A[i]["k"] = k
            A[i]["L"][:k] = L
            A[i]["seed"], A[i]["first"], A[i]["count"] = 40 + rep, 100 * i, 100
        mg, _, _, sg = g.batch_sampled(A, want_samples=True)
        me, _, _, se = e.batch_sampled(A, want_samples=True)
        assert np.array_equal(mg, me)
        for a, b in zip(sg, se):
            assert np.array_equal(a["zsum"], b["zsum"]) and np.array_equal(a["sexp"], b["sexp"])
        # a single sampled call of another shape in between
        ra, _ = g.sampled_projector(P, L, k, 0, 7 + rep, 0, 64)
        rb, _ = e.sampled_projector(P, L, k, 0, 7 + rep, 0, 64)
        assert ra.mean == rb.mean and ra.ips == rb.ips


def test_batch_exact_values_read_in_place(engines):
    # batch_exact_values reads every instance's value from the results array
    # in place: the same values as batch_exact's result objects, replayed and
    # eager
    from circuitsimulator_amd.workloads import uneven_instances
    g, e = engines
    inst, _ = uneven_instances(40, seed=91, kmax=9)
    for E in (g, e, g):
        res, _, _ = E.batch_exact(inst)
        vals, ms = E.batch_exact_values(inst)
        assert vals.dtype == np.float64 and vals.shape == (len(inst),) and ms > 0
        assert np.array_equal(vals, np.array([r.value for r in res]))
