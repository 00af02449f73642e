#!/usr/bin/env python3
"""Benchmark of the stabilizer-rank inner-product engine on one MI355X.

Metric: inner products/s at t=48 (batched persistent engine vs the plain
one-launch-per-pair kernel), with |error| of the sampled estimator against the
FP64 oracle.  A step is one full norm-estimation pass (innerprod.c:45-144):
M random stabilizer states theta drawn on the device, each projected by the
r=24 Pauli group G, then <phi_i|theta'> for all chi=2^12 decomposition states
phi_i of |L> -- M*chi = 2^28 inner products per step at the default M=2^16.

Printed: one JSON line (rank 0).  `value` is the whole-call throughput
(theta generation + projection + sweep + exact estimator), inputs resident.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic work per inner product of the elimination algorithm the sweep
# runs (fixation of the pivot coordinates outside supp(x_j), restrictions on
# non-pivot columns, A/B/C elimination steps), in 64-bit word operations of a
# dense-row formulation: the mean over 64 projected random states x 512
# seeded j of this workload, from the register-level model of the kernel
# (tests/test_sweep_model.py, algorithmic_work), which checks it.  Two int32
# lane operations per word operation.
ALG_WORDOPS_PER_IP = 982.0
ALG_LANE_OPS_PER_IP = 2 * ALG_WORDOPS_PER_IP
# The canonical dense-row algorithm's count (the reference's shrink /
# basis-alignment / exponential-sum sequence, oracle/cse_oracle.c
# orc_ip_wordops), mean over the same workload
CANONICAL_WORDOPS_PER_IP = 14743

# Per inner product of the sweep kernel on this workload, from rocprofv3
# --pmc over one dispatch of 2^28 inner products
# (profiles/k_sweep_norm_estimation_t48_r24_k12_M65536.md):
VALU_INSTS_PER_IP = 163.8    # SQ_INSTS_VALU / inner products (wave64 instructions)
SALU_INSTS_PER_IP = 159.1    # SQ_INSTS_SALU / inner products
VALU_LANE_OPS_PER_IP = VALU_INSTS_PER_IP * 64
# HBM bytes per inner product of the sweep: (FETCH_SIZE x 2.000 + WRITE_SIZE x
# 1.000) KiB, factors from a 1 GiB copy kernel, / inner products
SWEEP_BYTES_PER_IP = 2.80
CLOCK_GHZ = 2.38             # GRBM_GUI_ACTIVE / kernel time / 8 XCDs
PEAK_INT_TOPS = 39.3  # 256 CU x 4 SIMD x 16 lanes x 2.4 GHz, int32 lane-ops/s


def exact_mean(t, sexp, zsum):
    """FP64 oracle of the estimator: the exact value sum_s 2^{t-s-64}(A + B sqrt2) / M,
    from the exact per-sample Z[omega] sums, rounded once."""
    from decimal import Decimal, getcontext
    getcontext().prec = 80
    r2 = Decimal(2).sqrt()
    tot = Decimal(0)
    for s, c in zip(sexp.tolist(), zsum.tolist()):
        if s < 0:
            continue
        c0, c1, c2, c3 = c
        A = c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3
        B = c0 * c1 - c0 * c3 + c1 * c2 + c2 * c3
        tot += (Decimal(A) + Decimal(B) * r2) * (Decimal(2) ** (t - s - 64))
    return float(tot / len(sexp))


def baseline_thetas(P_tuple, t, seed, count):
    """The first `count` non-vanishing projected thetas of the Philox stream
    keyed by seed (the draws of timed step 0): (sample indices, oracle
    states, projection exponents, the first index past them)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    ph, xs, zs = P_tuple
    Po = orc.make_projector(ph, xs, zs, t)
    thr = orc.thresholds(t)
    idx, states, sexp = [], [], []
    st = 0
    while len(states) < count:
        th = orc.random_state(t, seed, st, thr)
        se = orc.project(th, Po)
        if se >= 0:
            idx.append(st)
            states.append(th)
            sexp.append(se)
        st += 1
    return idx, states, sexp, st


def cpu_baseline(P_tuple, L, t, k, seed, budget_s=15.0, lib="libref.so", threads=0, detail=None):
    """The reference's own C (oracle/_ref/libref.so, compiled from the reference
    sources at -O2; libref_O0.so at -O0) on a bounded sample of the same
    workload: the first nphi decomposition states for `threads` projected
    thetas (default: the host's cores, at most 16), one POSIX thread per theta.
    Falls back to the oracle (a port) when the library is absent.  detail (a
    dict) receives the sample: theta indices and states, nphi, the reference's
    per-term results and per-theta partial sums."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    idx, thetas, sexp, _ = baseline_thetas(P_tuple, t, seed, threads)
    if detail is not None:
        detail.update({"index": idx, "thetas": thetas, "sexp": sexp})
    if orc.ref_available(lib):
        t0 = time.perf_counter()
        orc.sampled_sweep(thetas[:1], t, L, k, nphi=2, threads=1, lib_ref=lib)
        per_ip = (time.perf_counter() - t0) / 2
        nphi = int(max(2, min(1 << k, budget_s / per_ip)))
        t0 = time.perf_counter()
        cs, tm, n = orc.sampled_sweep(thetas, t, L, k, nphi=nphi, threads=threads, terms=True, lib_ref=lib)
        dt = time.perf_counter() - t0
        if detail is not None:
            detail.update({"nphi": nphi, "ref_csum": cs, "ref_terms": tm, "kind": "reference"})
        return {"value": n / dt, "unit": "inner products/s", "cores": threads, "kind": "reference",
                "sample": f"{threads} projected thetas x first {nphi} of {1 << k} |L> states at t={t} "
                          f"({n} inner products, reference stabilizer.c/stateprep.c built "
                          f"{'-O0' if lib == 'libref_O0.so' else '-O2'}, "
                          f"{threads} threads, {dt:.1f} s)"}
    # oracle port, single thread
    phis = [orc.prepL(i, t, L) for i in range(64)]
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < budget_s:
        for p in phis:
            orc.inner(thetas[n % len(thetas)], p)
            n += 1
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "inner products/s", "cores": 1, "kind": "port",
            "sample": f"{n} inner products theta' x prepL at t={t} by the oracle, 1 thread, {dt:.1f} s"}


def zsum_partial(terms, nphi):
    """Z[w] sums (scale 2^32) of the first nphi packed terms of each row, as
    the engine and the oracle accumulate them"""
    import numpy as np
    out = np.zeros((terms.shape[0], 4), dtype=np.int64)
    for r in range(terms.shape[0]):
        c = [0, 0, 0, 0]
        for v in terms[r, :nphi].tolist():
            if not v:
                continue
            v -= 1
            p, m = -(v >> 3), v & 7
            e = 32 + (p >> 1)
            val = 1 << e
            parts = [(m, 1)] if not (p & 1) else [((m + 1) & 7, 1), ((m + 3) & 7, -1)]
            for q, sg in parts:
                if q >= 4:
                    q, sg = q - 4, -sg
                c[q] += sg * val
        out[r] = c
    return out


def subsample_estimator(t, zsums, sexps):
    """mean over the given samples of 2^t |pf sum_i <phi_i|theta'>|^2 from exact
    Z[w] sums (pf = 2^{-sexp/2}), the exact value rounded once"""
    from decimal import Decimal, getcontext
    getcontext().prec = 80
    r2 = Decimal(2).sqrt()
    tot = Decimal(0)
    for s, c in zip(sexps, zsums):
        if s < 0:
            continue
        c0, c1, c2, c3 = (int(v) for v in c)
        A = c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3
        B = c0 * c1 - c0 * c3 + c1 * c2 + c2 * c3
        tot += (Decimal(A) + Decimal(B) * r2) * (Decimal(2) ** (t - s - 64))
    return float(tot / len(sexps))


def parity_leg(t, k, L, detail, gpu, threads):
    """Untimed check of the engine against the reference C and the oracle on
    the cpu_baseline's thetas (the first non-vanishing draws of timed step
    0): the reference's per-term (eps, p, m mod 8) and partial sums over
    j < nphi, the oracle's full per-theta sums over all chi terms, the
    projection exponents, and the estimator over these thetas from the
    engine's int64 and Kahan-FP32 paths against the oracle's FP64 value."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    idx = detail["index"]
    chi = 1 << k
    g_terms = gpu["terms"][idx]
    g_sexp = gpu["sexp"]
    g_zsum = gpu["zsum"][idx]
    g_kahan = gpu["kahan"][idx]
    out = {"thetas": len(idx), "theta_indices": idx}
    # every draw before the last checked one: the engine's projection
    # vanishes exactly where the oracle's does, with the same exponent
    want_sexp = np.full(len(g_sexp), -1, dtype=np.int32)
    want_sexp[idx] = detail["sexp"]
    out["projection_exponent_mismatches"] = int(np.count_nonzero(g_sexp != want_sexp))
    out["vanished_terms_nonzero"] = int(np.count_nonzero(gpu["terms"][want_sexp < 0]))
    if detail.get("kind") == "reference":
        nphi = detail["nphi"]
        out["reference"] = "reference stabilizer.c/stateprep.c compiled from source (oracle/_ref)"
        out["reference_terms_checked"] = int(len(idx) * nphi)
        out["reference_term_mismatches"] = int(np.count_nonzero(g_terms[:, :nphi] != detail["ref_terms"]))
        out["reference_partial_sum_mismatches"] = int(
            np.count_nonzero((zsum_partial(g_terms, nphi) != detail["ref_csum"]).any(axis=1)))
    t0 = time.perf_counter()
    cs, tm, n = orc.sampled_sweep(detail["thetas"], t, L, k, nphi=chi, threads=threads, terms=True)
    out["oracle_terms_checked"] = int(n)
    out["oracle_term_mismatches"] = int(np.count_nonzero(g_terms != tm))
    out["oracle_sum_mismatches"] = int(np.count_nonzero((g_zsum != cs).any(axis=1)))
    out["oracle_s"] = round(time.perf_counter() - t0, 2)
    se = detail["sexp"]
    fp64 = subsample_estimator(t, cs, se)
    i64 = subsample_estimator(t, g_zsum, g_sexp[idx])
    kv = [2.0 ** (t - s) * (kr * kr + ki * ki) for s, (kr, ki) in zip(g_sexp[idx].tolist(), g_kahan.tolist())]
    kah = sum(kv) / len(kv)
    out["estimator"] = {"fp64_oracle": fp64, "int64_exact": i64, "kahan_fp32": kah,
                        "abs_err_int64": abs(i64 - fp64), "abs_err_kahan_fp32": abs(kah - fp64)}
    bad = [key for key in ("projection_exponent_mismatches", "vanished_terms_nonzero", "reference_term_mismatches",
                           "reference_partial_sum_mismatches", "oracle_term_mismatches", "oracle_sum_mismatches")
           if out.get(key, 0)]
    out["mismatches"] = int(sum(out.get(key, 0) for key in bad))
    return out


def chunk_arg(args, default):
    """--chunk: an item size in j, 0 for the engine's rule, or "auto" for the
    autotuner's choice; `default` where it is not given"""
    v = args.chunk if args.chunk is not None else default
    return v if v == "auto" else int(v)


def tuned_exact_chunk(E, instances, arg):
    """(claim cap, tuning record): the autotuner runs before the warmup,
    outside the timed region, on a proxy of the same instances, and leaves
    its kernel choice set on the context"""
    if arg != "auto":
        return arg, {"size": arg}
    from circuitsimulator_amd.autotune import tune_exact
    tm = {}
    t0 = time.perf_counter()
    kern, c = tune_exact(E, instances, timings=tm)
    E.set_sweep_kernel(kern, mode=1)
    return c, {"kernel": kern, "size": c, "tune_s": time.perf_counter() - t0,
               "proxy_sweep_ms": {str(key): round(v, 4) for key, v in tm.items()}}


def exact_pairs_config(args):
    """C2, reading B (the reference-native loop, innerprod.c:148-261): exact
    mode at t=24, r=12, k=11, chi=2048, so 2,098,176 pair terms, each one
    inner product <phi_j|Pi phi_i>, as one cse_batch_exact instance (what
    cse_exact_projector runs) with the autotuner's item size.  value = pair
    terms/s of the whole call."""
    from circuitsimulator_amd import Engine
    from circuitsimulator_amd.engine import Instance
    from circuitsimulator_amd.workloads import sampled_config
    t, r, k = 24, 12, 11
    P, P_tuple, L = sampled_config(t, r, k, seed=1002)
    chi = 1 << k
    nterms = chi * (chi + 1) // 2
    inst = Instance()
    inst.P = P
    inst.k = k
    for i, v in enumerate(L):
        inst.L[i] = v
    with Engine(0) as E:
        if args.kernel:
            E.set_sweep_kernel(args.kernel, mode=1)
        chunk, tuning = tuned_exact_chunk(E, [inst], chunk_arg(args, "auto"))
        for _ in range(args.warmup):
            E.batch_exact([inst], chunk=chunk)
        t0 = time.perf_counter()
        kernel_ms = 0.0
        for _ in range(args.steps):
            rl, _, kms = E.batch_exact([inst], chunk=chunk)
            res = rl[0]
            kernel_ms += kms
        dt = time.perf_counter() - t0
    return {
        "metric": "exact projector pair terms/s (one inner product each), t=24 r=12 k=11",
        "value": nterms * args.steps / dt, "unit": "inner products/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1000.0 / args.steps, "higher_is_better": True,
        "dtype": "int64", "data": "synthetic (seeded commuting Pauli group and L)",
        "config": {"workload": f"C2 reading B: exact projector t={t} r={r} k={k}, {nterms} pair terms",
                   "t": t, "r": r, "k": k, "chi": chi, "seed": 1002, "chunk": chunk},
        "tuning": tuning,
        "kernel_ip_per_s": nterms * args.steps / (kernel_ms / 1000.0),
        "result": {"value_int64_exact": res.value, "re": res.re, "im": res.im,
                   "kahan_fp32": [res.kahan_re, res.kahan_im],
                   "abs_err_kahan_fp32": abs(complex(res.kahan_re, res.kahan_im) - complex(res.re, res.im))},
    }


def sampled_h_config(args):
    """Sampled mode with the |H^t> decomposition (probability()'s default
    decomposition; innerprod.c:88-144 with prepH, stateprep.c:36-81): t=40,
    r=12, 512 random states x chi = 2^20 terms <prepH(j)|theta'> each.
    value = inner products/s of the whole call."""
    from circuitsimulator_amd import Engine
    from circuitsimulator_amd.workloads import sampled_config
    t, r, M = 40, 12, 512
    P, _, _ = sampled_config(t, r, 1, seed=1040)
    chi = 1 << ((t + 1) // 2)
    with Engine(0) as E:
        if args.kernel:
            E.set_sweep_kernel(args.kernel, mode=0)
        for w in range(args.warmup):
            E.sampled_projector(P, [], 0, 1, 7, 0, 64, want_samples=False)
        t0 = time.perf_counter()
        kernel_ms = 0.0
        ips = 0
        for _ in range(args.steps):
            res, _ = E.sampled_projector(P, [], 0, 1, 7, 0, M, want_samples=False)
            kernel_ms += res.kernel_ms
            ips += res.ips
        dt = time.perf_counter() - t0
    return {
        "metric": "sampled-mode inner products/s with the |H^t> decomposition, t=40 r=12",
        "value": ips / dt, "unit": "inner products/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1000.0 / args.steps, "higher_is_better": True,
        "dtype": "int64", "data": "synthetic (seeded Philox random stabilizer states; seeded commuting Pauli group)",
        "config": {"workload": f"sampled |H^t>: t={t} r={r}, {M} random states x chi=2^{(t + 1) // 2}",
                   "t": t, "r": r, "samples": M, "chi": chi, "seed": 1040},
        "kernel_ip_per_s": ips / (kernel_ms / 1000.0),
        "result": {"mean_int64_exact": res.mean, "mean_kahan_fp32": res.mean_kahan, "vanished": res.vanished},
    }


def exact_h_config(args):
    """The exact projector with the |H^t> decomposition (probability()'s
    default, innerprod.c:148-261 with prepH, stateprep.c:36-81): t=30, r=10,
    chi = 2^15, so 536,887,296 pair terms, each one inner product
    <prepH(j)|Pi prepH(i)>.  value = pair terms/s of the whole call."""
    from circuitsimulator_amd import Engine
    from circuitsimulator_amd.workloads import sampled_config
    t, r = 30, 10
    P, _, _ = sampled_config(t, r, 1, seed=1030)
    chi = 1 << ((t + 1) // 2)
    nterms = chi * (chi + 1) // 2
    with Engine(0) as E:
        if args.kernel:
            E.set_sweep_kernel(args.kernel, mode=1)
        for _ in range(args.warmup):
            E.exact_projector(P, [], 0, 1)
        t0 = time.perf_counter()
        kernel_ms = 0.0
        for _ in range(args.steps):
            res, _ = E.exact_projector(P, [], 0, 1)
            kernel_ms += res.kernel_ms
        dt = time.perf_counter() - t0
    return {
        "metric": "exact projector pair terms/s with the |H^t> decomposition, t=30 r=10",
        "value": nterms * args.steps / dt, "unit": "inner products/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1000.0 / args.steps, "higher_is_better": True,
        "dtype": "int64", "data": "synthetic (seeded commuting Pauli group)",
        "config": {"workload": f"exact projector, |H^t> decomposition, t={t} r={r}, {nterms} pair terms",
                   "t": t, "r": r, "chi": chi, "seed": 1030},
        "kernel_ip_per_s": nterms * args.steps / (kernel_ms / 1000.0),
        "result": {"value_int64_exact": res.value, "re": res.re, "im": res.im},
    }


def batch_config(args):
    """C3: 4096 uneven exact-mode instances (t ~ U{8..48}, r ~ U{1..t/2},
    k ~ U{2..14}), every pair term of every instance in one persistent
    launch.  value = pair terms/s of the whole call."""
    from circuitsimulator_amd import Engine
    from circuitsimulator_amd.workloads import uneven_instances
    count = args.instances
    inst, specs = uneven_instances(count, seed=1003)
    nterms = sum((1 << k) * ((1 << k) + 1) // 2 for (_, _, k, *_rest) in specs)
    with Engine(0) as E:
        if args.kernel:
            E.set_sweep_kernel(args.kernel, mode=1)
        chunk, tuning = tuned_exact_chunk(E, inst, chunk_arg(args, "auto"))
        for _ in range(args.warmup):
            E.batch_exact(inst, chunk=chunk)
        t0 = time.perf_counter()
        kernel_ms = 0.0
        for _ in range(args.steps):
            res, _, ms = E.batch_exact(inst, chunk=chunk)
            kernel_ms += ms
        dt = time.perf_counter() - t0
    ims = [abs(r.im) for r in res]
    return {
        "metric": "batched uneven exact instances: pair terms/s (one inner product each)",
        "value": nterms * args.steps / dt, "unit": "inner products/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1000.0 / args.steps, "higher_is_better": True,
        "dtype": "int64", "data": "synthetic (seeded per-instance commuting Pauli groups and L)",
        "config": {"workload": f"C3: {count} instances t~U{{8..48}} r~U{{1..t/2}} k~U{{2..14}}, "
                               f"{nterms} pair terms", "instances": count, "seed": 1003, "chunk": chunk},
        "tuning": tuning,
        "kernel_ip_per_s": nterms * args.steps / (kernel_ms / 1000.0),
        "checks": {"max_abs_im": max(ims), "min_value": min(r.re for r in res),
                   "terms_reported": int(sum(r.terms for r in res)), "terms_expected": nterms},
    }


def circuit_config(args):
    """C1: probability of {0:0, 1:0} for the seeded 8-qubit circuit of 40
    {H, S, CX} gates and 6 T gates, exact decomposition and exact pair sums
    (noapprox), through the whole front end (gadgetize, truncate) and the
    engine.  value = probabilities/s."""
    from circuitsimulator_amd.probability import probability
    from circuitsimulator_amd.workloads import c1_circuit
    circ, y = c1_circuit()
    cfg = {"noapprox": True, "exact": True, "quiet": True, "x": "", "y": y, "seed": 1}
    for _ in range(args.warmup):
        probability(circ, {0: 0, 1: 0}, config=dict(cfg))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p = probability(circ, {0: 0, 1: 0}, config=dict(cfg))
    dt = time.perf_counter() - t0
    return {
        "metric": "probability evaluations/s (front end + engine), C1 circuit, exact",
        "value": args.steps / dt, "unit": "probabilities/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1000.0 / args.steps, "higher_is_better": True,
        "dtype": "int64", "data": "synthetic (seeded circuit)",
        "config": {"workload": "C1: n=8, 40 {H,S,CX} + 6 T, measure {0:0, 1:0}, t=6, chi=8, "
                               "36 pair terms per projector", "seed": 1001},
        "probability": p,
    }


def exact_projector_oracle(rec):
    """exactProjector (innerprod.c:148-199) of one cse_instance record by the
    oracle: |sum_{i<=j} <phi_j|Pi phi_i>| (2 Re off the diagonal), summed
    exactly in Q(sqrt2) and rounded once to FP64."""
    import ctypes
    from decimal import Decimal, getcontext
    from fractions import Fraction
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc
    t, g = int(rec["P"]["nqubits"]), int(rec["P"]["nstabs"])
    ex, k = int(rec["exact"]), int(rec["k"])
    Po = orc.make_projector([int(v) for v in rec["P"]["phase"][:g]], [int(v) for v in rec["P"]["xs"][:g]],
                            [int(v) for v in rec["P"]["zs"][:g]], t)
    Larr = (orc.U64 * max(1, k))(*([int(v) for v in rec["L"][:k]] if k else [0]))
    chi = (1 << ((t + 1) // 2)) if ex else (1 << k)
    # real part: sum of c 2^{P/2} cos(pi M/4) = A + B sqrt2 (exact-mode totals are real)
    A, B = Fraction(0), Fraction(0)
    i_, j_, ip, se = orc.U64(), orc.U64(), orc.IpRes(), ctypes.c_int()
    for l in range(chi * (chi + 1) // 2):
        if orc.lib().orc_exact_term(l, ctypes.byref(Po), Larr, k, ex, ctypes.byref(i_), ctypes.byref(j_),
                                    ctypes.byref(ip), ctypes.byref(se)) or not ip.eps:
            continue
        P, M = ip.p - se.value, ip.m % 8
        w = 1 if i_.value == j_.value else 2      # 2 Re(z) off the diagonal
        cos2 = {0: (1, 0), 1: (0, Fraction(1, 2)), 2: (0, 0), 3: (0, Fraction(-1, 2)), 4: (-1, 0),
                5: (0, Fraction(-1, 2)), 6: (0, 0), 7: (0, Fraction(1, 2))}[M]  # cos = a + b sqrt2
        scale = Fraction(2) ** (P // 2)
        a, b = cos2
        if P % 2:   # times sqrt2
            a, b = 2 * b, a
        A += w * scale * a
        B += w * scale * b
    getcontext().prec = 60
    v = Decimal(A.numerator) / Decimal(A.denominator) + Decimal(B.numerator) / Decimal(B.denominator) * \
        Decimal(2).sqrt()
    return float(abs(v))


def c5_error_leg(circ, nq, shots=8, k=5, seed=1006):
    """Per-shot |error| of the sampling workload: an untimed run of the same
    circuit at k = 5 (chi = 32: the oracle's exact pair sums take seconds)
    for a seeded subsample of shots, every engine value checked against the
    oracle's exact projector sum (rounded once), and each shot's
    probabilities and FP64 conditional-probability ratios recomputed from
    the oracle's values along the same draws (sample.py:47,70)."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor

    from circuitsimulator_amd.sample import sampleShots
    trace = []
    sampleShots(circ, {}, list(range(nq)), shots,
                config={"exact": False, "k": k, "quiet": True, "rng": np.random.default_rng(seed), "trace": trace})
    batches = [e for e in trace if "instances" in e]
    recs = [rec for b in batches for rec in b["instances"]]
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as pool:
        ovals = list(pool.map(exact_projector_oracle, recs))
    engine_vals = [v for b in batches for v in b["values"]]
    val_err = max((abs(v - o) / max(abs(o), 1e-300) for v, o in zip(engine_vals, ovals)), default=0.0)
    # P per round from the oracle's values, then the chain of ratios
    it = iter(ovals)
    perr = np.zeros(shots)
    p0err = np.zeros(shots)
    pending = None
    psofar = np.ones(shots)
    for e in trace:
        if "instances" in e:
            ov = np.array([next(it) for _ in range(len(e["instances"]))])
            Por = e["P"].copy()
            todo, inv, nnum = e["todo"], e["inv"], e["nnum"]
            num = ov[:nnum][inv]
            den = np.full(len(todo), ov[nnum]) if e["exact"] and len(ov) > nnum else (
                ov[nnum:] if len(ov) > nnum else np.full(len(todo), e["norm"] ** 2))
            Por[todo] = np.where(num == 0, 0.0, e["base"] * num / np.where(num == 0, 1.0, den))
            pending = Por
            continue
        Por = pending if pending is not None else e["P"]
        pending = None
        P0or = Por / psofar
        perr = np.maximum(perr, np.abs(e["P"] - Por))
        p0err = np.maximum(p0err, np.abs(e["P0"] - P0or))
        psofar = np.where(e["draw"] == 0, P0or, 1 - P0or)
    return {"shots": shots, "k": k, "seed": seed, "instances_checked": len(recs),
            "max_rel_err_projector_values": val_err,
            "per_shot_abs_err_P": [float(x) for x in perr],
            "per_shot_abs_err_P0": [float(x) for x in p0err],
            "max_abs_err_P0": float(p0err.max())}


def shots_config(args):
    """C5: qubit-by-qubit sampling of 8 qubits of a seeded 40-qubit circuit
    of 400 {H, S, CX} gates and 40 T gates (t = 40), 1024 shots, |L> with
    k = 8 (chi = 256; 2000 samples per bin exceed chi, so the reference's
    rule evaluates each norm exactly), all shots advanced together.  value =
    shots/s."""
    import numpy as np

    from circuitsimulator_amd.sample import sampleShots
    from circuitsimulator_amd.workloads import c1_circuit
    circ, _ = c1_circuit(seed=1005, n=40, gates=400, ts=40)
    shots = args.shots
    cfg = {"exact": False, "k": 8, "quiet": True}
    if args.kernel:
        from circuitsimulator_amd.engine import shared_engine
        shared_engine(0).set_sweep_kernel(args.kernel, mode=1)
    if args.warmup:
        sampleShots(circ, {}, list(range(8)), 8, config=dict(cfg, rng=np.random.default_rng(7)))
    # each timed step is a full run of `shots` shots with its own draws
    freq = {}
    t0 = time.perf_counter()
    for s in range(args.steps):
        out = sampleShots(circ, {}, list(range(8)), shots, config=dict(cfg, rng=np.random.default_rng(1005 + s)))
        if s == 0:
            for o in out:
                freq[o] = freq.get(o, 0) + 1
    dt = time.perf_counter() - t0
    check = None if args.no_cpu_baseline else c5_error_leg(circ, 8)
    return {
        "metric": "sampled shots/s, 8 qubits of a 40-qubit Clifford+T circuit (t=40)",
        "value": args.steps * shots / dt, "unit": "shots/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1000.0 / args.steps, "higher_is_better": True, "dtype": "int64",
        "data": "synthetic (seeded circuit)",
        "config": {"workload": "C5: n=40, 400 {H,S,CX} + 40 T, sample qubits 0..7, |L> k=8, "
                               f"{shots} shots", "seed": 1005, "shots": shots, "k": 8},
        "inner_products": shots * 8 * 2 * (256 * 257 // 2),
        "distinct_outcomes": len(freq),  # of the first timed step
        "error_vs_fp64_oracle": check,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 3; c1: 1,000 probability evaluations of about 0.13 ms each; c4: 60; "
                         "c5: runs of --shots shots)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 1; c1: 20)")
    ap.add_argument("--samples", type=int, default=65536, help="random thetas per step (M)")
    ap.add_argument("--chunk", default=None,
                    help="item size in j (exact mode: the claim cap), 0 for the engine's rule, or auto for the "
                         "autotuner's choice (c4: kernel and item size) (default: auto for c2 and c3, 0 otherwise)")
    ap.add_argument("--plain-pairs", type=int, default=2048)
    ap.add_argument("--general-pairs", type=int, default=65536,
                    help="pairs of the general kernel's throughput leg (one launch)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--config", choices=["c4", "c1", "c2", "c3", "c5", "h30", "hs40"], default="c4",
                    help="c4: the headline norm estimation (default); c1, c2, c3, c5: the other configurations; "
                         "h30: the exact projector with the |H^t> decomposition at t=30; hs40: sampled mode "
                         "with the |H^t> decomposition at t=40")
    ap.add_argument("--shots", type=int, default=1024, help="c5: number of shots")
    ap.add_argument("--instances", type=int, default=4096, help="c3: number of instances")
    ap.add_argument("--cpu-baseline-only", action="store_true",
                    help="c4: time only the CPU baseline (no GPU) and print it as one JSON line")
    ap.add_argument("--ref-opt", choices=["O2", "O0"], default="O2",
                    help="optimisation level of the reference C build the CPU baseline times")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: cores, at most 16)")
    ap.add_argument("--kernel", type=int, default=0, choices=[0, 1, 2],
                    help="sweep kernel of the configuration's mode (cse_set_sweep_kernel; 0: the engine's choice)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"c1": 1000, "c4": 60}.get(args.config, 3)
    if args.warmup is None:
        args.warmup = 20 if args.config == "c1" else 1
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: the engine runs on one GPU; --gpus above 1 is not supported (a call's work is not "
              "split across GPUs).", file=sys.stderr)
        sys.exit(2)

    if args.config != "c4":
        out = {"c1": circuit_config, "c2": exact_pairs_config, "c3": batch_config,
               "c5": shots_config, "h30": exact_h_config, "hs40": sampled_h_config}[args.config](args)
        print(json.dumps(out), flush=True)
        return

    import numpy as np

    from circuitsimulator_amd import Engine
    from circuitsimulator_amd.workloads import sampled_config

    t, r, k = 48, 24, 12
    M = args.samples
    P, P_tuple, L = sampled_config(t, r, k, seed=1004)
    reflib = "libref_O0.so" if args.ref_opt == "O0" else "libref.so"
    if args.cpu_baseline_only:
        print(json.dumps({"cpu_baseline": cpu_baseline(P_tuple, L, t, k, 1004, lib=reflib,
                                                       threads=args.cpu_threads)}), flush=True)
        return
    chi = 1 << k
    check = not args.no_cpu_baseline
    cpu_threads = args.cpu_threads or max(1, min(16, os.cpu_count() or 1))
    if check:
        # the thetas of the CPU baseline and the parity leg: the first
        # non-vanishing draws of timed step 0 (seed 1004)
        _, _, _, ncheck = baseline_thetas(P_tuple, t, 1004, cpu_threads)
    with Engine(0) as E:
        if args.kernel:
            E.set_sweep_kernel(args.kernel, mode=0)
        cus, devname = E.device_info()
        chunk = chunk_arg(args, 0)
        tuning = {"size": chunk}
        if chunk == "auto":
            # the same draw as one step, as a one-instance sampled batch
            from circuitsimulator_amd.autotune import tune_sampled
            from circuitsimulator_amd.engine import sampled_instance_array
            A = sampled_instance_array(1)
            A[0]["P"] = np.frombuffer(bytes(P), dtype=A.dtype["P"])[0]
            A[0]["k"] = k
            A[0]["L"][:k] = L
            A[0]["seed"], A[0]["count"] = 1004, M
            tm = {}
            t1 = time.perf_counter()
            kern, chunk = tune_sampled(E, A, timings=tm)
            E.set_sweep_kernel(kern)
            tuning = {"kernel": kern, "size": chunk, "tune_s": time.perf_counter() - t1,
                      "proxy_sweep_ms": {str(key): round(v, 4) for key, v in tm.items()}}
        for w in range(args.warmup):
            E.sampled_projector(P, L, k, 0, 1004 + 100 + w, 0, M, chunk=chunk, want_samples=False)
        ips = 0
        kernel_ms = prep_ms = 0.0
        t0 = time.perf_counter()
        for s in range(args.steps):
            last = s == args.steps - 1
            res, ex = E.sampled_projector(P, L, k, 0, 1004 + s, 0, M, chunk=chunk, want_samples=last)
            ips += res.ips
            kernel_ms += res.kernel_ms
            prep_ms += res.prep_ms
        dt = time.perf_counter() - t0
        value = ips / dt
        kernel_ips = ips / (kernel_ms / 1000.0)

        # the exact value of the last step's estimator from its exact
        # per-sample sums, rounded once: the int64 path's conversion and the
        # Kahan-FP32 path against it (the sums themselves are checked against
        # the reference and the oracle in the parity leg)
        exact_last = exact_mean(t, ex["sexp"], ex["zsum"])
        err_int64 = abs(res.mean - exact_last)
        err_kahan = abs(res.mean_kahan - exact_last)

        # plain: the general per-pair kernel, one launch per pair
        npairs = args.plain_pairs
        th, sexp = E.random_states(t, 7, 0, 64, P)
        ok = np.nonzero(sexp >= 0)[0]
        rng = np.random.default_rng(5)
        A = th[ok[rng.integers(0, len(ok), size=npairs)]]
        B = E.prep_states(0, t, L, rng.integers(0, chi, size=npairs))
        E.inner_products(t, A[:64], B[:64], one_launch_per_pair=True)  # warm
        _, plain_ms = E.inner_products(t, A, B, one_launch_per_pair=True)
        _, general_batched_ms = E.inner_products(t, A, B, one_launch_per_pair=False)
        plain_ips = npairs / (plain_ms / 1000.0)
        # the same general kernel on enough pairs to fill the device (the
        # plain leg's pairs repeated): its throughput, not one wave's latency
        gen_pairs = max(npairs, args.general_pairs)
        reps = -(-gen_pairs // npairs)
        Ag, Bg = np.concatenate([A] * reps)[:gen_pairs], np.concatenate([B] * reps)[:gen_pairs]
        E.inner_products(t, Ag[:npairs], Bg[:npairs], one_launch_per_pair=False)  # warm
        _, general_full_ms = E.inner_products(t, Ag, Bg, one_launch_per_pair=False)

        # untimed: every term of timed step 0's first draws, for the parity leg
        gpu_check = None
        if check:
            _, gpu_check = E.sampled_projector(P, L, k, 0, 1004, 0, ncheck, chunk=chunk, want_terms=True)

    out = {
        "metric": "inner products/s, batched vs plain, t=48 on 1 GPU; |err| vs FP64 oracle",
        "value": value,
        "unit": "inner products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1000.0 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded Philox random stabilizer states; seeded commuting Pauli group and L)",
        "config": {"workload": f"norm estimation t={t} r={r} k={k} (chi={chi}) M={M} random thetas, "
                               f"{M * chi} inner products per step",
                   "t": t, "r": r, "k": k, "chi": chi, "samples": M, "seed": 1004},
        "device": devname,
        "kernel_ip_per_s": kernel_ips,
        "sweep_ms_per_step": kernel_ms / args.steps,
        "prep_ms_per_step": prep_ms / args.steps,
        "plain_ip_per_s": plain_ips,
        "plain_pairs": npairs,
        "general_batched_ip_per_s": npairs / (general_batched_ms / 1000.0),
        "general_full_ip_per_s": gen_pairs / (general_full_ms / 1000.0),
        "general_full_pairs": gen_pairs,
        "batched_vs_plain": value / plain_ips,
        "estimator": {"int64_exact": res.mean, "kahan_fp32": res.mean_kahan,
                      "exact_value_rounded_once": exact_last,
                      "abs_err_int64": err_int64, "abs_err_kahan_fp32": err_kahan},
    }
    if "tune_s" in tuning:
        out["tuning"] = tuning
    # integer-VALU roofline of the sweep kernel: the algorithm's own work
    # (int32 lane-operations per inner product of the elimination algorithm
    # it runs) at the kernel's rate, against the VALU peak
    achieved = kernel_ips * ALG_LANE_OPS_PER_IP / 1e12
    out["roofline"] = {"bound": "valu", "achieved": achieved, "peak": PEAK_INT_TOPS, "unit": "Tops/s",
                       "frac": achieved / PEAK_INT_TOPS,
                       "traffic": SWEEP_BYTES_PER_IP * M * chi,
                       "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE of one sweep dispatch, calibrated",
                       "kernel": "k_sweep_packed",
                       "per_unit": f"{ALG_LANE_OPS_PER_IP:.0f} int32 lane-ops ({ALG_WORDOPS_PER_IP:.0f} 64-bit "
                                   f"word-ops) per inner product of the fixation/elimination algorithm",
                       "canonical_wordops_per_ip": CANONICAL_WORDOPS_PER_IP}
    # the binding limit: instruction issue (one SALU per CU per cycle, one
    # wave64 VALU per SIMD per 4 cycles), from the executed instruction counts
    out["issue"] = {
        "salu_frac": kernel_ips * SALU_INSTS_PER_IP / (cus * CLOCK_GHZ * 1e9),
        "valu_frac": kernel_ips * VALU_INSTS_PER_IP / (cus * 4 * CLOCK_GHZ * 1e9 / 4),
        "salu_per_ip": SALU_INSTS_PER_IP, "valu_per_ip": VALU_INSTS_PER_IP,
        "executed_valu_lane_ops_per_ip": VALU_LANE_OPS_PER_IP,
        "hbm_GBps": kernel_ips * SWEEP_BYTES_PER_IP / 1e9,
    }
    if check:
        detail = {}
        out["cpu_baseline"] = cpu_baseline(P_tuple, L, t, k, 1004, lib=reflib, threads=cpu_threads, detail=detail)
        par = parity_leg(t, k, L, detail, gpu_check, cpu_threads)
        # the estimator over the checked thetas against the oracle's FP64 value
        out["estimator"]["vs_fp64_oracle"] = dict(par.pop("estimator"), samples=len(detail["index"]))
        out["parity"] = par
    print(json.dumps(out), flush=True)
    if check and out["parity"]["mismatches"]:
        print(f"bench.py: parity leg found {out['parity']['mismatches']} mismatches", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
