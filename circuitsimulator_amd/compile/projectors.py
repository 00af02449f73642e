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
    Mdict, y, n, t = _outcomes(circ, measure, verbose, x, y, rng)
    G = gadgetize.gadgetize(circ, measure, y)
    H = gadgetize.gadgetize(circ, Mdict, y)
    return G, H, n, t


def projectors_masks(circ, measure, verbose=False, x=None, y=None, rng=None):
    """projectors() with G and H as (phases, xs, zs, n + t), each x and z row
    an int (bit q for qubit q); the same draws and messages."""
    Mdict, y, n, t = _outcomes(circ, measure, verbose, x, y, rng)
    G, H = gadgetize.gadgetize_masks_pair(circ, measure, Mdict, y)
    return G, H, n, t


def _outcomes(circ, measure, verbose, x, y, rng):
    """(Mdict, y, n, t): the circuit's measurement outcomes (filled into
    measure as well) and the T-ancilla outcomes, drawn where not given."""
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
    return Mdict, y, n, t


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


def masks(rows, width):
    """Bit j of mask i is rows[i][j], for 0/1 rows of `width` entries."""
    if len(rows) == 0:
        return []
    a = np.asarray(rows, dtype=np.uint8).reshape(len(rows), width)
    packed = np.packbits(a, axis=1, bitorder="little")
    return [int.from_bytes(row.tobytes(), "little") for row in packed]


def _bit_rows(ms, width):
    """0/1 int64 rows of `width` entries, row i the bits of ms[i]."""
    nb = max(1, (width + 7) // 8)
    raw = b"".join(m.to_bytes(nb, "little") for m in ms)
    bits = np.unpackbits(np.frombuffer(raw, dtype=np.uint8).reshape(len(ms), nb), axis=1, bitorder="little")
    return bits[:, :width].astype(np.int64)


def _null_space_masks(cols, r):
    """_null_space on bitmasks: rows are r-bit masks of generators, cols[c]
    the generators with an X on circuit qubit c; the same column-by-column
    construction and row order."""
    X = [1 << i for i in range(r)]
    for cm in cols:
        good, bad = [], []
        for v in X:
            (bad if (v & cm).bit_count() & 1 else good).append(v)
        X = good + [bad[j] ^ bad[j - 1] for j in range(1, len(bad))]
    return X


def truncate(n, P):
    """(P', u): P projected onto |0^n> on its first n qubits, with u the
    number of generators lost (projectors.py:87-162).  Products equal to
    +identity on the t magic qubits are dropped; -identity is kept."""
    phases, xs, zs = P
    if len(phases) == 0:
        return ([], [], []), 0
    size = len(xs[0])
    (ph, xm, zm, t), u = truncate_masks(n, (phases, masks(xs, size), masks(zs, size), size))
    return to_arrays((ph, xm, zm, t)), u


def truncate_masks(n, P):
    """truncate on bitmask rows: P = (phases, xs, zs, size) with int rows;
    returns ((phases, xs, zs, t), u) with rows over the t = size - n magic
    qubits.  The null space and the products run as _null_space and
    _multiply take them (the order of rows and of factors)."""
    phases, xm, zm, size = P
    r = len(phases)
    t = size - n
    if r == 0:
        return ([], [], [], t), 0
    prods = null_products(n, phases, xm, zm)
    out_ph, out_x, out_z = [], [], []
    for _, ph, hx, hz in prods:
        if ph == 0 and not hx and not hz:
            continue
        out_ph.append(ph)
        out_x.append(hx)
        out_z.append(hz)
    return (out_ph, out_x, out_z, t), r - len(prods)


def null_products(n, phases, xm, zm):
    """For each row v of the null space of the generators' X parts on the
    first n qubits (_null_space's rows, as r-bit masks): (v, phase, x, z) of
    the product of the generators in v, in index order as _multiply takes
    them, with x and z shifted down to the qubits from n on."""
    r = len(phases)
    # cols[c]: the generators with an X on circuit qubit c
    cols = [0] * n
    low = (1 << n) - 1
    for i in range(r):
        m = xm[i] & low
        while m:
            c = (m & -m).bit_length() - 1
            cols[c] |= 1 << i
            m &= m - 1
    out = []
    for v in _null_space_masks(cols, r):
        ph, x, z = 0, 0, 0
        i, w = 0, v
        while w:
            if w & 1:
                ph = (ph + int(phases[i]) + 2 * (z & xm[i]).bit_count()) % 4
                x ^= xm[i]
                z ^= zm[i]
            w >>= 1
            i += 1
        out.append((v, ph, x >> n, z >> n))
    return out


def to_arrays(P):
    """(phases, xs, zs) with 0/1 int64 rows of a bitmask projector
    (phases, xs, zs, width)."""
    ph, xm, zm, width = P
    if not ph:
        return [], [], []
    xr, zr = _bit_rows(xm, width), _bit_rows(zm, width)
    return list(ph), list(xr), list(zr)
