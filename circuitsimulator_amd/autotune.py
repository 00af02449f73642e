"""Item-size autotuner for the engine's persistent sweep.

The sweep's device work queue hands out (record, chunk of j) items.  The
chunk size trades per-item cost (the record load, the lane tables, the flush
of the histogram or of the per-record sums) against balance: a launch ends
when the last resident wave finishes its last item.  Every chunk size gives
the same terms and the same exact results (integer sums, independent of
order), so the choice is a timing question only.

tune_exact_chunk / tune_sampled_chunk time each candidate on the device (the
sweep's HIP-event time, best of `reps`) on a proxy of the workload, together
with the engine's own size rule (chunk 0), and return the fastest, cached per
(device, kind, shape key) for the process.  The engine's rule stays (0 is
returned) unless a fixed size beats it by more than `tol`: on large launches
every size lands within about 1-2% and the proxy ranks them within about 2%
of the whole launch, while small launches (a few items per resident wave)
differ by 5-50% (measured on MI355X: 2,048 records of one projector run 5%
faster in items of 192 than of 256, and 70% slower in items of 1024).

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

# item sizes in j (terms); 64 terms run per asm block, so multiples of 64
CANDIDATES = (128, 192, 256, 384, 512, 1024)
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


def _best(run, candidates, reps, tol):
    timings = {c: min(run(c) for _ in range(reps)) for c in (0,) + tuple(candidates)}
    best = min(timings, key=timings.get)
    if timings[0] <= timings[best] * (1.0 + tol):
        best = 0
    return best, timings


def tune_exact_chunk(E, instances, candidates=CANDIDATES, reps=2, items_per_wave=32, tol=0.03, timings=None):
    """The item size for cse_batch_exact on these instances (a list of
    Instance or an instance_array): a candidate, or 0 for the engine's rule.
    `timings`, if a dict, receives the sweep milliseconds of each candidate
    (0 included) on the proxy."""
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

    def run(c):
        return E.batch_exact(proxy, chunk=c)[2]

    best, t = _best(run, candidates, reps, tol)
    if timings is not None:
        timings.update(t)
    with _lock:
        _cache[key] = best
    return best


def tune_sampled_chunk(E, instances, candidates=CANDIDATES, reps=2, items_per_wave=32, tol=0.03,
                       timings=None):
    """The item size for cse_batch_sampled on these instances (a
    sampled_instance_array), or 0 for the engine's rule; the proxy keeps each
    instance's seed and first sample and a proportional share of its
    count."""
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

    def run(c):
        return E.batch_sampled(proxy, chunk=c)[2]

    best, t = _best(run, candidates, reps, tol)
    if timings is not None:
        timings.update(t)
    with _lock:
        _cache[key] = best
    return best


def clear_cache():
    with _lock:
        _cache.clear()
