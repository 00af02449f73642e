"""Projectors G and H of a compiled circuit and their truncation to the
magic-state qubits (libcirc/compile/projectors.py).

G projects onto the requested outcomes together with the circuit's own
measurements x and the T-gadget outcomes y; H onto x and y only.  Truncating
projects the first n (circuit) qubits onto |0^n>: the generator products
whose X part vanishes there survive, restricted to the t ancillas.
"""
import numpy as np

from . import compilecirc, gadgetize


def _bits_of(s, length, name):
    if len(s) != length:
        raise ValueError("%s needs to have length %d" % (name, length))
    out = []
    for ch in s:
        if ch not in "01":
            raise ValueError("Only 0s and 1s allowed in %s string." % name)
        out.append(int(ch))
    return np.array(out, dtype=int)


def projectors(circ, measure, verbose=False, x=None, y=None, rng=None):
    """(G, H, n, t) of a compiled circuit (projectors.py:19-83).  measure maps
    qubit indexes to 0/1 and receives the circuit's measurement outcomes;
    x (circuit measurements) and y (T ancillas) are 0/1 strings, drawn at
    random when not given (from rng, default numpy's global generator)."""
    rand = rng if rng is not None else np.random
    n = len(compilecirc.standardGate(circ.splitlines()[0], lineMode=True))
    t, Ms, MTs = gadgetize.countY(circ)

    if x is None:
        Mselect = rand.randint(2, size=len(Ms)) if rng is None else rng.integers(0, 2, size=len(Ms))
    else:
        Mselect = _bits_of(x, len(Ms), "x")

    # a T under a measurement counts when that measurement gave 1
    for M in MTs:
        if Mselect[Ms.index(M)] == 1:
            t += 1

    Mdict = {}
    for i, M in enumerate(Ms):
        Mdict[M] = Mselect[i]
        if M not in measure:
            measure[M] = Mselect[i]

    if y is None:
        y = rand.randint(2, size=t) if rng is None else rng.integers(0, 2, size=t)
    else:
        y = _bits_of(y, t, "y")

    if verbose:
        print("n = %d, t = %d" % (n, t))
        print("Measurement info:")
        print("x: " + "".join(str(measure[i]) if i in measure else "_" for i in range(n)))
        print("y: " + "".join(str(int(v)) for v in y))

    G = gadgetize.gadgetize(circ, measure, y)
    H = gadgetize.gadgetize(circ, Mdict, y)
    return G, H, n, t


def _null_space(A):
    """Rows spanning {v : v A = 0} over GF(2), built column by column as the
    reference does (projectors.py:96-122): rows already orthogonal to the
    column stay, the others are summed pairwise (row j with row j - 1)."""
    r = len(A)
    X = np.eye(r, dtype=int)
    for col in range(len(A[0])):
        y = X.dot(A[:, col]) % 2
        good = [X[i] for i in range(len(y)) if y[i] == 0]
        bad = [X[i] for i in range(len(y)) if y[i] == 1]
        fixed = [(bad[j] + bad[j - 1]) % 2 for j in range(1, len(bad))]
        X = np.array(good + fixed, dtype=int).reshape((len(good) + len(fixed), r))
    return X


def _multiply(a, b):
    """Product of two generators (phase, x, z) (projectors.py:124-138)."""
    ph1, x1, z1 = a
    ph2, x2, z2 = b
    ph = 2 * int(np.sum((np.asarray(z1) == 1) & (np.asarray(x2) == 1)))
    return (ph1 + ph2 + ph) % 4, (x1 + x2) % 2, (z1 + z2) % 2


def truncate(n, P):
    """(P', u): P projected onto |0^n> on its first n qubits, with u the
    number of generators lost (projectors.py:87-162).  Products equal to
    +identity on the t magic qubits are dropped; -identity is kept."""
    phases, xs, zs = P
    if len(phases) == 0:
        return ([], [], []), 0
    size = len(xs[0])
    t = size - n
    A = np.array(xs, dtype=int)[:, :n]
    X = _null_space(A)
    out_ph, out_x, out_z = [], [], []
    for row in X:
        gen = (0, np.zeros(size, dtype=int), np.zeros(size, dtype=int))
        for i in range(len(row)):
            if row[i] == 1:
                gen = _multiply(gen, (phases[i], np.asarray(xs[i], dtype=int), np.asarray(zs[i], dtype=int)))
        ph, x, z = gen
        if ph == 0 and not x[n:].any() and not z[n:].any():
            continue
        out_ph.append(ph)
        out_x.append(x[n:])
        out_z.append(z[n:])
    return (out_ph, out_x, out_z), len(phases) - len(X)
