"""The exact-mode work queue of the sweep (cse_kernels.hip sweep_body,
cse_internal.h tri_item), restated on the CPU: the triangle enumeration of
(record, slice) items is a bijection onto the slices that reach the
diagonal, and the claims -- a static first claim per wave, then ranges of
slice indices taken with one atomic add from one of eight counters by
guided self-scheduling, each cut into one segment per record -- cover every
pair term j >= i of every instance exactly once, whatever the order in
which the waves claim, and every wave leaves."""
import math
import random

import pytest


def tri_before(ch, nch, g):
    return ch * (g * (2 * nch - g + 1) // 2)


def tri_items(chi, ch, nch):
    return tri_before(ch, nch, nch - 1) + (chi - (nch - 1) * ch)


def tri_item(l, ch, nch):
    """the kernel's inverse: block from the root of the quadratic with at
    most a step of correction, then the record and slice within the block"""
    a = 2.0 * nch + 1.0
    g = int((a - math.sqrt(a * a - 8.0 * l / ch)) * 0.5)
    g = max(0, min(nch - 1, g))
    while g < nch - 1 and tri_before(ch, nch, g + 1) <= l:
        g += 1
    while g > 0 and tri_before(ch, nch, g) > l:
        g -= 1
    off, per = l - tri_before(ch, nch, g), nch - g
    return g * ch + off // per, g + off % per


SLICE = 64  # CSE_SLICE_J


def slice_layout(chi, width=SLICE):
    ch = min(width, chi)
    return ch, (chi + ch - 1) // ch


@pytest.mark.parametrize("width", [16, 64])
@pytest.mark.parametrize("chi", [1, 2, 8, 63, 64, 65, 200, 2048, 4096])
def test_triangle_enumeration_is_a_bijection(chi, width):
    ch, nch = slice_layout(chi, width)
    n = tri_items(chi, ch, nch)
    seen = set()
    for l in range(n):
        r, c = tri_item(l, ch, nch)
        assert 0 <= r < chi and c >= r // ch and c < nch
        seen.add((r, c))
    assert len(seen) == n == sum(nch - r // ch for r in range(chi))
    # a record's slices are consecutive, from the one holding j = r on
    rows = {}
    for l in range(n):
        r, c = tri_item(l, ch, nch)
        rows.setdefault(r, []).append((l, c))
    for r, lc in rows.items():
        ls = [l for l, _ in lc]
        assert ls == list(range(ls[0], ls[0] + len(ls)))
        assert [c for _, c in lc] == list(range(r // ch, nch))


NQ = 8  # counters of the dynamic range (CSE_NQUEUES)
GSS_DIV = 2  # CSE_GSS_DIV
STATIC_DIV = 2  # CSE_STATIC_DIV


def run_queue(chis, waves, take_max, seed, width=SLICE):
    """The kernel's claim loop for `waves` waves (blocks of 4) claiming in a
    random interleaving: a static first claim per wave (slices [g f0,
    g f0 + f0) of [0, static_end)), then the dynamic range cut into NQ parts
    with a counter each; a wave starts at the part of its block (block % NQ)
    and moves to the next part when its claim lands past the part's end or
    it reads the counter past it.  A claim is sized from the counter as read
    just before it; another wave may claim between that read and the add.
    Returns every (instance, i, j) term its segments cover and the claims."""
    rng = random.Random(seed)
    lay = [slice_layout(c, width) for c in chis]
    off = [0]
    for c, (ch, nch) in zip(chis, lay):
        off.append(off[-1] + tri_items(c, ch, nch))
    nitems = off[-1]
    f0 = max(1, min(take_max, nitems // (STATIC_DIV * waves)))
    static_end = min(f0 * waves, nitems)
    part = -(-(nitems - static_end) // NQ)
    wpp = -(-waves // NQ)
    counters = [0] * NQ

    def bounds(q):
        q0 = static_end + q * part
        return q0, min(q0 + part, nitems)

    state = []
    for g in range(waves):
        qx = (g // 4) % NQ
        q0, q1 = bounds(qx)
        state.append({"cur": g * f0, "end": min(g * f0 + f0, static_end), "qx": qx, "tried": 0,
                      "q0": q0, "q1": q1, "run": True, "read": None})
    terms, claims = [], []
    while any(w["run"] for w in state):
        w = rng.choice([w for w in state if w["run"]])
        if w["cur"] >= w["end"]:
            if w["tried"] == NQ:
                w["run"] = False
                continue
            if w["read"] is None:
                # first step of a claim: read the counter (other waves may
                # claim before this wave adds)
                w["read"] = w["q0"] + counters[w["qx"]]
                continue
            pos, w["read"] = w["read"], None
            base = w["q1"]
            if pos < w["q1"]:
                take = max(1, min(take_max, (w["q1"] - pos) // (wpp * GSS_DIV)))
                base = w["q0"] + counters[w["qx"]]
                counters[w["qx"]] += take
            if base < w["q1"]:
                w["cur"], w["end"] = base, min(base + take, w["q1"])
                claims.append(w["end"] - w["cur"])
            else:
                w["tried"] += 1
                w["qx"] = (w["qx"] + 1) % NQ
                w["q0"], w["q1"] = bounds(w["qx"])
            continue
        item = w["cur"]
        inst = max(i for i in range(len(chis)) if off[i] <= item)
        ch, nch = lay[inst]
        r, c = tri_item(item - off[inst], ch, nch)
        nsl = min(w["end"] - w["cur"], nch - c)
        w["cur"] += nsl
        j0, j1 = max(c * ch, r), min((c + nsl) * ch, chis[inst])
        terms += [(inst, r, j) for j in range(j0, j1)]
    return terms, claims


@pytest.mark.parametrize("width", [16, 64])
@pytest.mark.parametrize("chis,waves,take_max", [([2048], 300, 4), ([1024], 3000, 8), ([8, 8], 2, 4),
                                                  ([1, 3, 64, 65, 130], 17, 3), ([512, 1024], 64, 1),
                                                  ([8], 64, 4), ([2048], 100, 8)])
def test_guided_claims_cover_every_term_once(chis, waves, take_max, width):
    for seed in range(2):
        terms, _ = run_queue(chis, waves, take_max, seed, width)
        want = [(q, i, j) for q, c in enumerate(chis) for i in range(c) for j in range(i, c)]
        assert sorted(terms) == want


def test_claims_shrink_to_one_slice_as_the_parts_drain():
    # one chi = 2048 instance (C2's shape) on 64 waves: the dynamic claims
    # run from about remaining / (waves per part x 2) down to single slices
    _, claims = run_queue([2048], 64, 64, seed=3)
    n = len(claims) // 10
    assert sum(claims[:n]) > 8 * sum(claims[-n:])
    assert claims.count(1) >= 8
