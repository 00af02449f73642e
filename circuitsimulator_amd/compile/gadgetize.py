"""Gadgetization of a compiled circuit into one stabilizer projector
(libcirc/compile/gadgetize.py).

Every T gate is replaced by a gadget that consumes one magic-state ancilla:
a CX from the qubit to the ancilla, an S correction on the qubit when the
ancilla's postselected outcome y is 1, and H S^3 on the ancilla.  The
projector onto the requested outcomes (Z generators on measured qubits and
on the T ancillas) is conjugated backwards through the circuit, gate by
gate, with Clifford lookup tables on (phase, x, z) where a generator is
i^phase Z(z) X(x).  T ancillas are numbered from the end of the circuit: the
last T uses qubit n, the one before it n + 1, and so on.
"""
import numpy as np

from .compilecirc import standardGate

import functools

# The conjugation tables of gadgetize.py:119-123, (x, z) -> (phase increment,
# x', z'), as updates of bitplanes over all generators at once: H swaps x and
# z and adds 2 where both are set; X adds 2 where z is set; S adds x to the
# phase and z ^= x; CX (gadgetize.py:146-156) takes x_t ^= x_c, z_c ^= z_t.
_SINGLE = ("H", "X", "S")


def countY(circuit):
    """Number of T gates not under a measurement, the measured qubits (in
    order of first measurement), and the measured qubits that condition a T
    gate (gadgetize.py:14-61).  Rejects operations on measured qubits and
    T T on one qubit with nothing in between."""
    ts, measured, conditional = _count(circuit)
    return ts, list(measured), list(conditional)


@functools.lru_cache(maxsize=16)
def _count(circuit):
    ts = 0
    pending_t = []  # qubits whose last gate was a T
    measured = []
    conditional = []
    for number, line in enumerate(circuit.splitlines(), start=1):
        groups = standardGate(line, lineMode=True)
        if "M" in groups:
            q = groups.index("M")
            if q not in measured:
                measured.append(q)
        for q in measured:
            if groups[q] not in ("M", "_"):
                raise SyntaxError("Line %d: Cannot perform operations to measured qubit %d" % (number, q))
        if "T" in groups and "M" not in groups:
            ts += 1
            q = groups.index("T")
            if q in pending_t:
                raise SyntaxError("Line %d: Consecutive Ts on qubit %d" % (number, q))
            pending_t.append(q)
        # a qubit touched by this line is no longer right after a T; the
        # list is filtered while it is walked, as the reference does
        for q in pending_t:
            if groups[q] != "_":
                pending_t.remove(q)
        if "T" in groups and "M" in groups:
            conditional.append(groups.index("M"))
    return ts, tuple(measured), tuple(conditional)


@functools.lru_cache(maxsize=16)
def _steps(circuit):
    """The non-empty lines of circuit from last to first, each as (the index
    of its M or -1, gate, qubits)."""
    out = []
    for line in reversed(circuit.splitlines()):
        if line == "":
            continue
        groups = standardGate(line, lineMode=True)
        gate = standardGate(line)
        out.append((groups.index("M") if "M" in groups else -1, gate,
                    tuple(groups.index(letter) for letter in gate)))
    return tuple(out)


class _Planes:
    """Generators i^phase Z(z) X(x) as bitplanes: bit g of X[q] (Z[q]) is
    generator g's x (z) on qubit q, bit g of P0 + 2 P1 its phase."""

    def __init__(self, size):
        self.X = [0] * size
        self.Z = [0] * size
        self.P0 = self.P1 = 0

    def add_phase(self, m):
        self.P1 ^= self.P0 & m
        self.P0 ^= m

    def single(self, gate, q):
        X, Z = self.X, self.Z
        if gate == "H":
            X[q], Z[q] = Z[q], X[q]
            self.P1 ^= X[q] & Z[q]
        elif gate == "X":
            self.P1 ^= Z[q]
        elif gate == "S":
            self.add_phase(X[q])
            Z[q] ^= X[q]
        else:
            raise KeyError(gate)

    def cx(self, c, t):
        self.X[t] ^= self.X[c]
        self.Z[c] ^= self.Z[t]


def _bit_rows(cols, ng):
    """(ng, len(cols)) 0/1 int64 matrix, column q the bits of cols[q]."""
    nb = max(1, (ng + 7) // 8)
    raw = b"".join(c.to_bytes(nb, "little") for c in cols)
    bits = np.unpackbits(np.frombuffer(raw, dtype=np.uint8).reshape(len(cols), nb), axis=1,
                         bitorder="little")[:, :ng]
    return bits.T.astype(np.int64)


def gadgetize(circuit, Mdict, y):
    """Projector (phases, xs, zs) of the outcomes Mdict on the circuit's
    qubits and y on the T ancillas, in the qubit order of the circuit
    followed by the ancillas (gadgetize.py:66-181).  Every generator is
    conjugated backwards through the circuit, all generators at once."""
    pl, ng, size = _conjugate(circuit, Mdict, y)
    phases = [((pl.P0 >> g) & 1) + 2 * ((pl.P1 >> g) & 1) for g in range(ng)]
    xr, zr = _bit_rows(pl.X, ng), _bit_rows(pl.Z, ng)
    return phases, [xr[g] for g in range(ng)], [zr[g] for g in range(ng)]


def gadgetize_masks(circuit, Mdict, y):
    """gadgetize as (phases, xs, zs, size) with xs[g] (zs[g]) an int whose
    bit q is generator g's x (z) on qubit q."""
    pl, ng, size = _conjugate(circuit, Mdict, y)
    xs, zs = _mask_rows(pl, ng, size)
    phases = [((pl.P0 >> g) & 1) + 2 * ((pl.P1 >> g) & 1) for g in range(ng)]
    return phases, xs, zs, size


def gadgetize_masks_pair(circuit, MG, MH, y):
    """(gadgetize_masks(circuit, MG, y), gadgetize_masks(circuit, MH, y))
    from one backward pass, when MG and MH agree on every measurement that
    controls a gate (else two passes).  The pass conjugates Z_q once for
    each qubit q of MG or MH and each ancilla; an outcome only flips the
    generator's phase bit P1, which no gate reads, so each projector takes
    its rows with its own outcomes applied at the end."""
    if any(MG.get(m) != MH.get(m) for m in _controls(circuit)):
        return gadgetize_masks(circuit, MG, y), gadgetize_masks(circuit, MH, y)
    qs = list(MG.keys()) + [q for q in MH.keys() if q not in MG]
    pl, ng, size = _conjugate(circuit, MG, y, qubits=qs)
    xs, zs = _mask_rows(pl, ng, size)
    slot = {q: i for i, q in enumerate(qs)}
    anc = list(range(len(qs), ng))
    out = []
    for M in (MG, MH):
        rows = [slot[q] for q in M.keys()] + anc
        P1 = pl.P1 ^ sum(1 << slot[q] for q in M.keys() if M[q] == 1)
        phases = [((pl.P0 >> g) & 1) + 2 * ((P1 >> g) & 1) for g in rows]
        out.append((phases, [xs[g] for g in rows], [zs[g] for g in rows], size))
    return out[0], out[1]


@functools.lru_cache(maxsize=16)
def _controls(circuit):
    """The measurements that control a gate of circuit."""
    return frozenset(m for m, _, _ in _steps(circuit) if m >= 0)


def _mask_rows(pl, ng, size):
    """(xs, zs): generator g's x (z) row as an int, bit q for qubit q."""
    xs, zs = [0] * ng, [0] * ng
    for q in range(size):
        for planes, acc in ((pl.X, xs), (pl.Z, zs)):
            m = planes[q]
            while m:
                g = (m & -m).bit_length() - 1
                acc[g] |= 1 << q
                m &= m - 1
    return xs, zs


def _conjugate(circuit, Mdict, y, qubits=None):
    """The bitplanes of gadgetize's generators after the backward pass, their
    number and the qubit count n + len(y).  The generators are Z_q with
    outcome Mdict[q] for q in Mdict (or Z_q with outcome 0 for q in
    `qubits`, when given), then Z on each ancilla with outcome y; Mdict
    decides the classically controlled gates either way."""
    n = len(standardGate(circuit.split("\n", 1)[0], lineMode=True))
    size = n + len(y)
    first = [(q, Mdict[q]) for q in Mdict.keys()] if qubits is None else [(q, 0) for q in qubits]
    gens = first + [(q, y[q - n]) for q in range(n, size)]
    ng = len(gens)
    pl = _Planes(size)
    for g, (q, outcome) in enumerate(gens):
        pl.Z[q] |= 1 << g
        if outcome == 1:
            pl.P1 |= 1 << g

    # _Planes.single and .cx inlined on local planes (this loop is most of
    # a probability evaluation's host time)
    X, Z = pl.X, pl.Z
    P0, P1 = pl.P0, pl.P1
    tpos = n
    for m, gate, qubits in _steps(circuit):
        if m >= 0 and Mdict[m] == 0:
            continue  # the gate runs only on outcome 1
        if ng:
            if gate == "T":
                q, a = qubits[0], tpos
                X[a] ^= X[q]                  # CX q -> ancilla
                Z[q] ^= Z[a]
                if y[a - n] == 1:             # S on q
                    xq = X[q]
                    P1 ^= P0 & xq
                    P0 ^= xq
                    Z[q] ^= xq
                X[a], Z[a] = Z[a], X[a]       # H on the ancilla
                P1 ^= X[a] & Z[a]
                # S three times on the ancilla: X stays, Z and P0 flip by
                # X once, and P1 gains X & ~P0 (P0 before the three)
                xa = X[a]
                P1 ^= xa & ~P0
                P0 ^= xa
                Z[a] ^= xa
            elif gate == "H":
                q = qubits[0]
                X[q], Z[q] = Z[q], X[q]
                P1 ^= X[q] & Z[q]
            elif gate == "X":
                P1 ^= Z[qubits[0]]
            elif gate == "S":
                q = qubits[0]
                xq = X[q]
                P1 ^= P0 & xq
                P0 ^= xq
                Z[q] ^= xq
            elif len(gate) == 1:
                raise KeyError(gate)
            else:
                c, tq = qubits
                X[tq] ^= X[c]
                Z[c] ^= Z[tq]
        if gate == "T":
            tpos += 1
    pl.P0, pl.P1 = P0, P1
    return pl, ng, size


def _product_phase(x1, z1, x2, z2):
    """2 for every qubit where Z of the left factor passes X of the right
    (gadgetize.py:193-201).  The left factor is a running sum that is not
    reduced mod 2, and the reference's table only matches entries 0 and 1,
    so a qubit where it has reached 2 contributes nothing."""
    return 2 * sum(1 for a, b, c, d in zip(x1, z1, x2, z2) if a in (0, 1) and b == 1 and c == 1 and d in (0, 1))


def expandGenerators(proj):
    """All 2^r - 1 non-identity elements of the group of commuting
    generators, element i being the product of the generators j with bit j
    of i set, most significant first (gadgetize.py:186-217)."""
    phases, xs, zs = proj
    r = len(phases)
    if r == 0:
        return [], [], []
    n = len(xs[0])
    out_ph, out_x, out_z = [], [], []
    for i in range(1, 2 ** r):
        ph, x, z = 0, np.zeros(n), np.zeros(n)
        for j in range(r):
            if (i >> (r - 1 - j)) & 1:
                ph = ph + phases[j] + _product_phase(x, z, xs[j], zs[j])
                x = x + xs[j]
                z = z + zs[j]
        out_ph.append(ph % 4)
        out_x.append(x % 2)
        out_z.append(z % 2)
    return out_ph, out_x, out_z
