"""Host logic of bench.py's parity leg, on the CPU: engine-shaped outputs
built from the oracle pass it with zero mismatches, and a planted wrong term,
sum or projection exponent is counted.  (On the GPU the same leg compares the
engine's outputs for timed step 0's first draws.)"""
import math

import numpy as np
import pytest

import orc
from gpu_util import random_commuting_projector


def _fake_engine_outputs(t, k, L, Po, seed, count):
    """what E.sampled_projector(..., want_terms=True) returns for samples
    [0, count), built from the oracle"""
    thr = orc.thresholds(t)
    chi = 1 << k
    terms = np.zeros((count, chi), dtype=np.int32)
    zsum = np.zeros((count, 4), dtype=np.int64)
    sexp = np.full(count, -1, dtype=np.int32)
    kahan = np.zeros((count, 2), dtype=np.float64)
    for s in range(count):
        th = orc.random_state(t, seed, s, thr)
        se = orc.project(th, Po)
        if se < 0:
            continue
        cs, tm, _ = orc.sampled_sweep([th], t, L, k, terms=True)
        terms[s], zsum[s], sexp[s] = tm[0], cs[0], se
        re = im = 0.0
        for v in tm[0].tolist():
            if v:
                v -= 1
                p, m = -(v >> 3), v & 7
                mag = 2.0 ** (p / 2)
                re += mag * math.cos(math.pi * m / 4)
                im += mag * math.sin(math.pi * m / 4)
        kahan[s] = (re, im)
    return {"terms": terms, "zsum": zsum, "sexp": sexp, "kahan": kahan}


@pytest.fixture(scope="module")
def setup():
    import bench
    rng = np.random.default_rng(31)
    t, r, k = 12, 6, 4
    ph, xs, zs = random_commuting_projector(rng, t, r)
    L = [orc.rand_bits(rng, t) for _ in range(k)]
    seed, threads = 77, 3
    idx, _, _, ncheck = bench.baseline_thetas((ph, xs, zs), t, seed, threads)
    Po = orc.make_projector(ph, xs, zs, t)
    gpu = _fake_engine_outputs(t, k, L, Po, seed, ncheck)
    return bench, t, k, L, (ph, xs, zs), seed, threads, gpu


def _leg(setup, gpu):
    bench, t, k, L, Pt, seed, threads, _ = setup
    detail = {}
    cb = bench.cpu_baseline(Pt, L, t, k, seed, budget_s=0.05, threads=threads, detail=detail)
    assert cb["cores"] >= 1 and cb["value"] > 0
    return bench.parity_leg(t, k, L, detail, gpu, threads), detail


def test_parity_leg_clean(setup):
    par, detail = _leg(setup, setup[-1])
    assert par["mismatches"] == 0, par
    assert par["oracle_terms_checked"] == par["thetas"] * (1 << setup[2])
    if detail.get("kind") == "reference":
        assert par["reference_terms_checked"] == par["thetas"] * detail["nphi"]
    est = par["estimator"]
    assert est["abs_err_int64"] == 0.0
    assert est["abs_err_kahan_fp32"] <= 1e-9 * max(1.0, est["fp64_oracle"])


@pytest.mark.parametrize("plant", ["term", "sum", "sexp"])
def test_parity_leg_counts_planted_mistakes(setup, plant):
    gpu = {key: v.copy() for key, v in setup[-1].items()}
    s = int(np.nonzero(gpu["sexp"] >= 0)[0][0])
    if plant == "term":
        j = int(np.nonzero(gpu["terms"][s])[0][0])
        gpu["terms"][s, j] += 8   # p one step lower
    elif plant == "sum":
        gpu["zsum"][s, 0] += 1
    else:
        gpu["sexp"][s] += 1
    par, _ = _leg(setup, gpu)
    assert par["mismatches"] > 0
