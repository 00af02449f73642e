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

# conjugation tables (x, z) -> (phase increment, x', z') (gadgetize.py:119-123)
_TABLE = {
    "H": {(0, 0): (0, 0, 0), (1, 0): (0, 0, 1), (0, 1): (0, 1, 0), (1, 1): (2, 1, 1)},
    "X": {(0, 0): (0, 0, 0), (1, 0): (0, 1, 0), (0, 1): (2, 0, 1), (1, 1): (2, 1, 1)},
    "S": {(0, 0): (0, 0, 0), (1, 0): (1, 1, 1), (0, 1): (0, 0, 1), (1, 1): (1, 1, 0)},
}


def countY(circuit):
    """Number of T gates not under a measurement, the measured qubits (in
    order of first measurement), and the measured qubits that condition a T
    gate (gadgetize.py:14-61).  Rejects operations on measured qubits and
    T T on one qubit with nothing in between."""
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
    return ts, measured, conditional


def _conjugate(gate, qubits, phase, x, z):
    """Conjugate one generator (phase, x, z) by a gate (gadgetize.py:125-156).
    x and z are lists of 0/1, changed in place."""
    if len(gate) == 1:
        q = qubits[0]
        dph, x[q], z[q] = _TABLE[gate][(int(x[q]), int(z[q]))]
        return (phase + dph) % 4
    c, tq = qubits
    # CX: X_c -> X_c X_t, Z_t -> Z_c Z_t, no phase in the ZX order
    xc, zc, xt, zt = int(x[c]), int(z[c]), int(x[tq]), int(z[tq])
    x[c], z[c] = xc, zc ^ zt
    x[tq], z[tq] = xc ^ xt, zt
    return phase


def gadgetize(circuit, Mdict, y):
    """Projector (phases, xs, zs) of the outcomes Mdict on the circuit's
    qubits and y on the T ancillas, in the qubit order of the circuit
    followed by the ancillas (gadgetize.py:66-181)."""
    n = len(standardGate(circuit.split("\n", 1)[0], lineMode=True))
    size = n + len(y)
    phases, xs, zs = [], [], []

    def z_generator(q, outcome):
        xs.append([0] * size)
        z = [0] * size
        z[q] = 1
        zs.append(z)
        phases.append(2 if outcome == 1 else 0)

    for q in Mdict.keys():
        z_generator(q, Mdict[q])
    for q in range(n, size):
        z_generator(q, y[q - n])

    tpos = n
    for line in reversed(circuit.splitlines()):
        if line == "":
            continue
        groups = standardGate(line, lineMode=True)
        if "M" in groups and Mdict[groups.index("M")] == 0:
            continue  # the gate runs only on outcome 1
        gate = standardGate(line)
        qubits = [groups.index(letter) for letter in gate]
        for g in range(len(phases)):
            if gate == "T":
                ph = _conjugate("CX", [qubits[0], tpos], phases[g], xs[g], zs[g])
                if y[tpos - n] == 1:
                    ph = _conjugate("S", qubits, ph, xs[g], zs[g])
                for h in ("H", "S", "S", "S"):
                    ph = _conjugate(h, [tpos], ph, xs[g], zs[g])
                phases[g] = ph
            else:
                phases[g] = _conjugate(gate, qubits, phases[g], xs[g], zs[g])
        if gate == "T":
            tpos += 1
    return phases, [np.array(v) for v in xs], [np.array(v) for v in zs]


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
