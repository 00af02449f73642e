"""Seeded synthetic workloads of the engine's benchmark configurations.

C2/C3/C4 projector groups: r independent commuting Hermitian Paulis on t
qubits, each drawn uniformly among the Paulis that commute with the earlier
ones (the distribution rejection sampling would give, drawn directly from the
symplectic complement), rejecting only draws in the span of the earlier
generators; phase m = (xi . zeta mod 2) + 2b makes i^m Z(zeta) X(xi) Hermitian.
L: k x t uniform rows with rank k.
"""
import numpy as np

from .engine import Instance, Projector


def _bits(rng, n):
    v = (int(rng.integers(0, 2 ** 32)) << 32) | int(rng.integers(0, 2 ** 32))
    return v & ((1 << n) - 1)


def commuting_paulis(rng, t, r):
    xs, zs, ph = [], [], []
    full = (1 << t) - 1
    tries = 0
    while len(xs) < r:
        tries += 1
        if tries > 100000:
            raise RuntimeError(f"could not draw {r} independent commuting Paulis on {t} qubits")
        ech = []  # fully reduced constraint rows (pivot bit, row)
        for x2, z2 in zip(xs, zs):
            c = (z2 << t) | x2  # parity(c & ((x << t) | z)) = x.z2 + z.x2
            for pb, e in ech:
                if (c >> pb) & 1:
                    c ^= e
            if c:
                pb = c.bit_length() - 1
                ech = [(q, e ^ c) if (e >> pb) & 1 else (q, e) for q, e in ech]
                ech.append((pb, c))
        v = (_bits(rng, t) << t) | _bits(rng, t)
        for pb, e in ech:
            if bin(e & v).count("1") % 2:
                v ^= 1 << pb
        xi, ze = v >> t, v & full
        if xi == 0 and ze == 0:
            continue
        red = []
        for w in [(x2 << t) | z2 for x2, z2 in zip(xs, zs)] + [(xi << t) | ze]:
            for b in red:
                w = min(w, w ^ b)
            if w:
                red.append(w)
        if len(red) < len(xs) + 1:
            continue
        par = bin(xi & ze).count("1") % 2
        ph.append((par + 2 * int(rng.integers(0, 2))) % 4)
        xs.append(xi)
        zs.append(ze)
    return ph, xs, zs


def gf2_rank(rows):
    red = []
    for w in rows:
        for b in red:
            w = min(w, w ^ b)
        if w:
            red.append(w)
    return len(red)


def random_L(rng, t, k):
    if not 0 <= k <= t:
        raise ValueError(f"a {k} x {t} L cannot have rank {k}")
    while True:
        L = [_bits(rng, t) for _ in range(k)]
        if gf2_rank(L) == k:
            return L


def sampled_config(t=48, r=24, k=12, seed=1004):
    """C4: norm estimation with a |L> decomposition (chi = 2^k)."""
    rng = np.random.default_rng(seed)
    ph, xs, zs = commuting_paulis(rng, t, r)
    L = random_L(rng, t, k)
    return Projector.make(ph, xs, zs, t), (ph, xs, zs), L


def uneven_instances(count, seed=1003, tmin=8, tmax=48, kmin=2, kmax=14):
    """C3: independent exact-mode instances, t ~ U{tmin..tmax},
    r ~ U{1..floor(t/2)}, k ~ U{kmin..min(kmax, t)} (L has rank k, so k <= t,
    as decompose clamps k to t), each with its own group and L."""
    rng = np.random.default_rng(seed)
    out, specs = [], []
    for _ in range(count):
        t = int(rng.integers(tmin, tmax + 1))
        r = int(rng.integers(1, t // 2 + 1))
        k = int(rng.integers(kmin, min(kmax, t) + 1))
        ph, xs, zs = commuting_paulis(rng, t, r)
        L = random_L(rng, t, k)
        inst = Instance()
        inst.P = Projector.make(ph, xs, zs, t)
        inst.exact = 0
        inst.k = k
        for i, v in enumerate(L):
            inst.L[i] = v
        out.append(inst)
        specs.append((t, r, k, ph, xs, zs, L))
    return out, specs


def c1_circuit(seed=1001, n=8, gates=40, ts=6):
    """C1: a compiled n-qubit circuit of `gates` gates drawn uniformly from
    {H, S, CX} (CX on an ordered pair of distinct qubits, uniform), then `ts`
    T gates inserted at uniform slots on uniform qubits, redrawn until no
    qubit has two T gates in a row; and the postselection y, fixed by the
    seed.  Returns (circuit text, y string)."""
    from .compile.gadgetize import countY
    rng = np.random.default_rng(seed)
    while True:
        lines = []
        for _ in range(gates):
            g = int(rng.integers(0, 3))
            row = ["_"] * n
            if g == 2:
                c, tq = rng.choice(n, size=2, replace=False)
                row[c], row[tq] = "C", "X"
            else:
                row[int(rng.integers(0, n))] = "HS"[g]
            lines.append("".join(row))
        for _ in range(ts):
            at = int(rng.integers(0, len(lines) + 1))
            row = ["_"] * n
            row[int(rng.integers(0, n))] = "T"
            lines.insert(at, "".join(row))
        circ = "\n".join(lines) + "\n"
        try:
            countY(circ)
        except SyntaxError:
            continue
        y = "".join(str(int(b)) for b in rng.integers(0, 2, size=ts))
        return circ, y

