"""Autotuner for the engine's persistent sweep: the knobs that change how the
work is cut and which kernel runs it, never the results.

Exact mode (cse_batch_exact): the sweep's occupancy and the claim cap.
Kernel 1 runs the exact sweep at seven waves per SIMD, kernel 2 at eight
(more latency hidden, more values spilled): long launches gain a few percent
at eight, short ones (a few slices per wave) lose a few, and the engine's
rule (kernel 0) takes eight from 16 slices per resident wave.  The sweep hands out slices of 64 j, a static first claim per wave
and then guided self-scheduling from eight counters; `chunk` caps a claim.
Larger caps cut the per-segment cost (the record load, the lane tables) on
launches with many slices per wave; the guided claims shrink at the end of
the queue whatever the cap.

Sampled mode (cse_batch_sampled): the sweep kernel and the item size.  Kernel
1 is the eight-wave sweep whose hand-scheduled path takes records of at most
two non-pivot columns (others run compiled code), kernel 2 the eight-wave
sweep with the general restriction loop (cse_set_sweep_kernel); the item size
trades per-item cost (record load, tables, flush of the per-record sums)
against balance.

Every setting gives the same terms and the same exact results (integer sums,
independent of order), so the choice is a timing question only.
tune_exact / tune_sampled time each candidate on the device (the
sweep's HIP-event time, best of `reps`) on a proxy of the workload, together
with the engine's own rule, and return the fastest, cached per (device, kind,
shape key) for the process.  The engine's rule stays unless a candidate
beats it by more than `tol`.

Proxy.  A workload with many items per resident wave is throughput-bound, and
the candidates rank on a part of it as on the whole: exact batches keep their
largest instances (they carry most of the terms) plus a stride of the rest,
sampled batches keep a proportional share of every instance's samples, either
way down to about `items_per_wave` items per resident wave.  A workload
smaller than that is timed whole.
"""
import math
import threading

import numpy as np

# sampled mode: item sizes in j (terms); 64 terms run per asm block, so
# multiples of 64
CANDIDATES = (128, 192, 256, 384, 512, 1024)
# the sweep kernels of either mode (cse_set_sweep_kernel)
KERNELS = (1, 2)
# exact mode: claim caps in j (multiples of the 64-j slice)
EXACT_CAPS = (256, 512, 1024, 2048, 4096)
# resident waves per CU of the two packed sweep kernels (4 waves per block)
_WAVES_PER_CU = {"exact": 7 * 4, "sampled": 8 * 4}
# packed layout (D in the J rows) up to this width
_PACKED_MAX_T = 60

_cache = {}
_lock = threading.Lock()


def exact_items(chi, chunk):
    """items of an exact-mode instance: the chunks of j in [0, chi) of record
    i that reach the diagonal j >= i (the engine's upper-triangle walk)."""
    nch = -(-chi // chunk)
    return chunk * ((nch - 1) * (nch + 2) // 2) + (chi - (nch - 1) * chunk)


def sampled_items(chi, chunk, count):
    """items of a sampled-mode instance: every chunk of j for every sample"""
    return count * (-(-chi // chunk))


def _chi(t, k, exact):
    return 1 << ((t + 1) // 2) if exact else 1 << k


def _exact_fields(instances):
    """(t, k, exact) per instance of a list of Instance or an instance_array"""
    if isinstance(instances, np.ndarray):
        return [(int(t), int(k), int(e)) for t, k, e in
                zip(instances["P"]["nqubits"], instances["k"], instances["exact"])]
    return [(int(i.P.nqubits), int(i.k), int(i.exact)) for i in instances]


def _key(E, kind, ts, hdec, chis, items):
    cus, name = E.device_info()
    packed = max(ts) <= _PACKED_MAX_T
    waves = cus * _WAVES_PER_CU[kind]
    # items per wave in octaves, saturating where the launch is throughput-bound
    load = min(int(math.log2(max(1.0, items / waves))), 6)
    mean_chi = sum(chis) / len(chis)
    return (name, cus, kind, packed, bool(hdec), load, int(round(math.log2(mean_chi))))


def _best(run, candidates, reps, tol, rule=0):
    """the fastest of candidates, or `rule` (the engine's own choice, timed
    first) unless a candidate beats it by more than tol"""
    timings = {c: min(run(c) for _ in range(reps)) for c in (rule,) + tuple(candidates)}
    best = min(timings, key=timings.get)
    if timings[rule] <= timings[best] * (1.0 + tol):
        best = rule
    return best, timings


def tune_exact(E, instances, kernels=KERNELS, candidates=EXACT_CAPS, reps=2, items_per_wave=32, tol=0.02,
               timings=None):
    """(kernel, claim cap in j) for cse_batch_exact on these instances (a
    list of Instance or an instance_array): (0, 0) for the engine's rule,
    else a kernel of `kernels` with a cap of (0,) + `candidates`.
    `timings`, if a dict, receives the sweep milliseconds of each (kernel,
    cap) on the proxy."""
    from .engine import INSTANCE_DTYPE, Instance
    fields = _exact_fields(instances)
    if not fields:
        raise ValueError("no instances")
    chis = [_chi(t, k, e) for t, k, e in fields]
    per = [exact_items(c, 256) for c in chis]
    total = sum(per)
    key = _key(E, "exact", [f[0] for f in fields], any(f[2] for f in fields), chis, total)
    with _lock:
        if key in _cache:
            return _cache[key]
    cus, _ = E.device_info()
    target = items_per_wave * cus * _WAVES_PER_CU["exact"]
    if total <= target:
        pick = list(range(len(fields)))
    else:
        order = sorted(range(len(fields)), key=lambda i: -per[i])
        stride = max(1, total // target)
        pick, got = [], 0
        for pos in range(0, len(order), stride):
            pick.append(order[pos])
            got += per[order[pos]]
            if got >= target:
                break
    if isinstance(instances, np.ndarray):
        proxy = np.ascontiguousarray(instances[pick])
        assert proxy.dtype == INSTANCE_DTYPE
    else:
        proxy = [instances[i] for i in pick]
        assert all(isinstance(p, Instance) for p in proxy)
    keep = getattr(E, "sweep_kernels", [0, 0])[1]

    def run(kc):
        kernel, chunk = kc
        E.set_sweep_kernel(kernel, mode=1)
        try:
            return E.batch_exact(proxy, chunk=chunk)[2]
        finally:
            E.set_sweep_kernel(keep, mode=1)

    cands = [(kern, c) for kern in kernels for c in (0,) + tuple(candidates)]
    best, t = _best(run, cands, reps, tol, rule=(0, 0))
    if timings is not None:
        timings.update(t)
    with _lock:
        _cache[key] = best
    return best


def tune_sampled(E, instances, kernels=KERNELS, candidates=CANDIDATES, reps=2, items_per_wave=32, tol=0.02,
                 timings=None):
    """(kernel, item size) for cse_batch_sampled on these instances (a
    sampled_instance_array): (0, 0) for the engine's rule, else a kernel of
    `kernels` with an item size of (0,) + `candidates`.  The proxy keeps
    each instance's seed and first sample and a proportional share of its
    count.  `timings`, if a dict, receives the sweep milliseconds of each
    (kernel, size) on the proxy."""
    if len(instances) == 0:
        raise ValueError("no instances")
    ts = [int(t) for t in instances["P"]["nqubits"]]
    chis = [_chi(t, int(k), int(e)) for t, k, e in zip(ts, instances["k"], instances["exact"])]
    counts = [int(c) for c in instances["count"]]
    total = sum(sampled_items(c, 256, n) for c, n in zip(chis, counts))
    key = _key(E, "sampled", ts, any(int(e) for e in instances["exact"]), chis, total)
    with _lock:
        if key in _cache:
            return _cache[key]
    cus, _ = E.device_info()
    target = items_per_wave * cus * _WAVES_PER_CU["sampled"]
    proxy = np.ascontiguousarray(instances.copy())
    if total > target:
        share = target / total
        proxy["count"] = [max(1, int(math.ceil(n * share))) for n in counts]
    keep = getattr(E, "sweep_kernels", [0, 0])[0]

    def run(kc):
        kernel, chunk = kc
        E.set_sweep_kernel(kernel, mode=0)
        try:
            return E.batch_sampled(proxy, chunk=chunk)[2]
        finally:
            E.set_sweep_kernel(keep, mode=0)

    cands = [(kern, c) for kern in kernels for c in (0,) + tuple(candidates)]
    best, t = _best(run, cands, reps, tol, rule=(0, 0))
    if timings is not None:
        timings.update(t)
    with _lock:
        _cache[key] = best
    return best


def clear_cache():
    with _lock:
        _cache.clear()
