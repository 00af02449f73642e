"""sampleQubits(circ, measure, sample, config): draw the outcomes of the
qubits in sample, one qubit at a time from conditional probabilities
(libcirc/sample.py:16-86), each probability computed by probability() on
the gfx950 engine.

The draw order and the arithmetic are the reference's: with constraints in
measure, their probability Psofar is computed first (exact decomposition);
then for each sampled qubit in turn P(qubit = 0, all so far) is computed,
P0 = P / Psofar, the qubit is drawn against a uniform number, and Psofar
becomes P0 or 1 - P0 -- the conditional probability of the outcome just
drawn, as in the reference (sample.py:54,84).  From the third sampled qubit
on, that divides a joint probability by a conditional one, so the draws do
not follow the circuit's distribution (a GHZ state yields 001).  The
config key chain = "joint" keeps Psofar the joint probability of all
outcomes so far instead (P, or Psofar - P after a 1), which samples the
true distribution; the default, "reference", reproduces the reference.
exact defaults to False here, and the file option is disabled.  config may
carry rng (a numpy Generator) for the uniform draws; numpy's global
generator is used otherwise, as in the reference.
"""
import numpy as np

from .compile import gadgetize as _gadgetize
from .compile import projectors as _projectors
from .probability import decompose, probability


def _next_psofar(chain, P, P0, Psofar, bit):
    """Psofar after drawing bit (elementwise for arrays)."""
    if chain == "joint":
        return np.where(bit == 0, P, Psofar - P)
    if chain != "reference":
        raise ValueError("chain must be 'reference' or 'joint'")
    return np.where(bit == 0, P0, 1 - P0)


def sampleQubits(circ, measure, sample, config={}):
    verbose = config.get("verbose")
    if config.get("exact") is None:
        config["exact"] = False
    config["file"] = None
    rng = config.get("rng")

    def uniform():
        return rng.random() if rng is not None else np.random.random()

    def prob(measure, exact=None):
        if exact is not None:
            old = config["exact"]
            config["exact"] = exact
        P = probability(circ, measure, config=config)
        if exact is not None:
            config["exact"] = old
        return P

    chain = config.get("chain", "reference")

    def draw(qubit, measure, Psofar):
        measure[qubit] = 0
        P = prob(measure)
        P0 = P / Psofar
        if verbose:
            print("Probability with qubit %d: %f" % (qubit, P))
            print("Conditional probability of qubit %d: P(0) = %f" % (qubit, P0))
        bit = 0 if uniform() < P0 else 1
        if verbose:
            print("-> Sampled " + str(bit) + "\n")
        measure[qubit] = bit
        return bit, _next_psofar(chain, P, P0, Psofar, bit)

    Psofar = 1
    if len(measure) != 0:
        Psofar = prob(measure, exact=True)
        if verbose:
            print("Constraints present. Probability of satifying: %f\n" % Psofar)
        if Psofar == 0:
            return ""
        if Psofar < 1e-5:
            print("Warning: probability of satisfying qubit constraints very low (%f). Could be impossible."
                  % Psofar)
    out = ""
    for q in sample:
        bit, Psofar = draw(q, measure, Psofar)
        out += str(bit)
    return out


# ---------------------------------------------------------------- batched
def _products(n, rows):
    """The reference's truncation (projectors.py:87-162) split into what is
    common to every shot and what is not: for each null-space row, the
    product of its generators with every outcome 0 -- phase, x and z masks on
    the magic qubits -- and the row itself, so that a shot's phase is that
    phase plus 2 for every generator in the row whose outcome is 1."""
    phases, xs, zs = rows
    size = len(xs[0])
    r = len(phases)
    prods = _projectors.null_products(n, phases, _projectors.masks(xs, size), _projectors.masks(zs, size))
    ph = [p for _, p, _, _ in prods]
    xm = [x for _, _, x, _ in prods]
    zm = [z for _, _, _, z in prods]
    X = np.array([[(v >> i) & 1 for i in range(r)] for v, _, _, _ in prods], dtype=int).reshape(len(prods), r)
    return np.array(ph, dtype=int), xm, zm, X, r - len(prods)


def _shot_projector(prods, values):
    """(phases, x masks, z masks) of one shot's G', or None when it contains
    -identity (the outcome is impossible); +identity products are dropped."""
    ph0, xm, zm, X, _ = prods
    ph = (ph0 + 2 * X.dot(values)) % 4
    keep_ph, keep_x, keep_z = [], [], []
    for j in range(len(ph)):
        if xm[j] == 0 and zm[j] == 0:
            if ph[j] == 2:
                return None
            if ph[j] == 0:
                continue
        keep_ph.append(int(ph[j]))
        keep_x.append(xm[j])
        keep_z.append(zm[j])
    return keep_ph, keep_x, keep_z


def _round_phases(prods, V):
    """Every shot's G' at once: _shot_projector for the outcome rows of V (one
    per shot, padded with 0 for the T-gadget rows).  The products' x and z
    masks are shared, so a shot's G' differs only in its phases.  Returns the
    phases of the kept generators (shots x g), their x and z masks, and which
    shots' G' contain -identity (the outcome is impossible)."""
    ph0, xm, zm, X, _ = prods
    V = np.asarray(V, dtype=np.int64)
    PH = (np.asarray(ph0, dtype=np.int64)[None, :] + 2 * (V @ X.T.astype(np.int64))) % 4
    ident = np.array([xm[j] == 0 and zm[j] == 0 for j in range(len(xm))], dtype=bool)
    if (PH[:, ident] % 2).any():
        raise AssertionError("a product of commuting Hermitian generators has phase +-i on the identity")
    impossible = (PH[:, ident] == 2).any(axis=1)
    keep = np.nonzero(~ident)[0]
    return PH[:, keep], [xm[j] for j in keep], [zm[j] for j in keep], impossible


def _bits(P, t):
    """0/1 arrays of a mask-form projector (for probability()'s helpers)."""
    ph, xm, zm = P
    return (ph, [np.array([(x >> j) & 1 for j in range(t)]) for x in xm],
            [np.array([(z >> j) & 1 for j in range(t)]) for z in zm])


# samples per engine call in sampled mode, as cse_projector_norm batches them
_CALL_SAMPLES = 1 << 20
# the denominator's seed is the numerator's XOR this (projector_norms)
_H_SEED = 0x9E3779B97F4A7C15


def _median(vals):
    """Median of each row as multiSampledProjector takes it
    (innerprod.c:30-38): the middle value, or the mean of the middle two."""
    s = np.sort(vals, axis=1)
    b = s.shape[1]
    return s[:, (b - 1) // 2] if b % 2 else (s[:, b // 2] + s[:, b // 2 - 1]) / 2


def _sampled_medians(engine, phases, xs, zs, ex, k, Ls, seeds, samples, bins, fill, seed_xor):
    """multiSampledProjector (innerprod.c:23-41) for every seed in seeds:
    the projector (xs, zs) with phases one row per seed or one row for all,
    the decomposition exact (ex) or row i of Ls.  Bin b of seed sd draws the
    samples [b samples, (b + 1) samples) keyed by sd ^ seed_xor, as
    cse_projector_norm does, so each median equals projector_norm's; every
    bin of every seed is one instance of cse_batch_sampled, as many per call
    as _CALL_SAMPLES allows."""
    from .engine import sampled_instance_array
    ns = len(seeds)
    A = sampled_instance_array(ns * bins)
    phases = np.asarray(phases)
    rows = np.repeat(phases, bins, axis=0) if len(phases) == ns else phases
    fill(A, slice(None), rows, xs, zs, ex, k, None if Ls is None else np.repeat(Ls, bins, axis=0))
    A["seed"] = np.repeat(np.array(seeds, dtype=np.uint64) ^ np.uint64(seed_xor), bins)
    A["first"] = np.tile(np.arange(bins, dtype=np.int64) * samples, ns)
    A["count"] = samples
    means = np.empty(ns * bins)
    per = _CALL_SAMPLES // samples
    for i0 in range(0, len(A), per):
        means[i0:i0 + per] = engine.batch_sampled(A[i0:i0 + per])[0]
    return _median(means.reshape(ns, bins))


def sampleShots(circ, measure, sample, shots, config={}):
    """shots independent sampleQubits(circ, measure, sample, config) draws
    (libcirc/sample.py:16-86), advanced together one sampled qubit per
    round.

    Circuit and gadgets are conjugated once: each generator's conjugation
    through the circuit is independent of the others, so G for any set of
    outcomes is a subset of the rows of one gadgetization, with the phase
    +2 on outcome 1.  Per round the null space of the truncation and the
    generator masks are shared by all shots (they depend on the X and Z
    parts only) and a shot's G' differs in its phases, which one matrix
    product gives for all shots.  Each call's decomposition is drawn
    afresh, as probability() does (one seed per call), and all numerators
    and denominators of a round are evaluated by one batched engine launch
    (cse_batch_exact, its instances written as one array) whenever the
    reference's rule evaluates them exactly (probability.c:162-174; with the
    default 2000 samples this is every |L> of k <= 11 and every exact |H> of
    t <= 22); with the exact decomposition every distinct G' and H' is
    evaluated once.  Otherwise the norms are sampled, every bin of every
    shot's numerator and denominator an instance of one batched launch
    (cse_batch_sampled, up to 2^20 samples per launch), each median equal
    to what probability() computes with the same seed.

    Differences from calling sampleQubits shots times, none of which changes
    a probability: one T-ancilla postselection y serves the whole run (any y
    gives the same exact probabilities; draw-for-draw identity with the
    reference is impossible anyway, its generators being unseeded), and the
    uniform draws are taken round by round.  A circuit with its own
    measurements is sampled shot by shot with sampleQubits, and a qubit both
    constrained and sampled is rejected.  Returns one string per shot (""
    when the constraints cannot be met).
    """
    from .compile.compilecirc import standardGate
    from .engine import instance_array, sampling_kept, shared_engine
    from .probability import _pick_sampling, decompose_many, projector_norms

    t, Ms, _ = _gadgetize.countY(circ)
    rng = config.get("rng")
    if Ms:
        return [sampleQubits(circ, dict(measure), list(sample), dict(config)) for _ in range(shots)]
    if any(q in measure for q in sample):
        raise ValueError("a sampled qubit is also constrained; use sampleQubits")
    cfg = dict(config)
    if cfg.get("exact") is None:
        cfg["exact"] = False
    if cfg.get("fidbound") is None:
        cfg["fidbound"] = 1e-5
    if not cfg.get("noapprox"):
        _pick_sampling(cfg, t, True, False)

    def rand_int(hi, size):
        return rng.integers(0, hi, size=size) if rng is not None else np.random.randint(0, hi, size=size)

    y = cfg.get("y")
    ybits = [int(c) for c in y] if y is not None else [int(b) for b in rand_int(2, t)]
    n = len(standardGate(circ.splitlines()[0], lineMode=True))
    order = list(measure.keys()) + list(sample)
    nq = len(order)
    G = _gadgetize.gadgetize(circ, {q: 0 for q in order}, ybits)
    Hp, v = _projectors.truncate(n, _gadgetize.gadgetize(circ, {}, ybits))
    Hm = ([int(p) for p in Hp[0]], [sum(int(b) << j for j, b in enumerate(x)) for x in Hp[1]],
          [sum(int(b) << j for j, b in enumerate(z)) for z in Hp[2]])
    device = int(cfg.get("device") or 0)
    engine = None
    # config["trace"] (a list): every exact batch of the run is appended to
    # it (its instance array, the engine's values, how they combine into
    # each shot's P), then each round's P, P0 and draws, for checks
    trace = cfg.get("trace")

    def fill(A, rows, phases, xs, zs, ex, k, Ls):
        """Records rows of A: projector (phases: one row per record or one
        shared row), decomposition (Ls: one row of L per record, or None)."""
        g = len(xs)
        A["P"]["nstabs"][rows] = g
        A["P"]["nqubits"][rows] = t
        if g:
            A["P"]["phase"][rows, :g] = np.asarray(phases, dtype=np.int64) & 3
            A["P"]["xs"][rows, :g] = np.array(xs, dtype=np.uint64)
            A["P"]["zs"][rows, :g] = np.array(zs, dtype=np.uint64)
        A["exact"][rows] = 1 if ex else 0
        A["k"][rows] = k
        if Ls is not None and k:
            A["L"][rows, :k] = Ls

    def evaluate(PH, xk, zk, impossible, uu, exact):
        """P of every shot of a round: PH, xk, zk, impossible as
        _round_phases returns them."""
        nonlocal engine
        shots_r = PH.shape[0]
        out = np.zeros(shots_r)
        base = 2.0 ** (v - uu)
        g = len(xk)
        # probability.py:140-150: the first len(G') generators of H' agree
        if g <= len(Hm[0]) and xk == Hm[1][:g] and zk == Hm[2][:g]:
            same = (PH == np.asarray(Hm[0][:g], dtype=np.int64)[None, :]).all(axis=1)
        else:
            same = np.zeros(shots_r, dtype=bool)
        out[same & ~impossible] = base
        todo = np.nonzero(~same & ~impossible)[0]
        if len(todo) == 0:
            return out
        # G' is not empty here (an empty G' agrees with H'); an empty H'
        # needs the exact decomposition's norm
        callcfg = dict(cfg)
        callcfg["exact"] = bool(exact) or len(Hm[0]) == 0
        seeds = [int(sd) for sd in rand_int(2 ** 62, len(todo))]
        Ls, norm, ex, k = decompose_many(t, callcfg, seeds)
        samples, bins = int(callcfg.get("samples") or 0), int(callcfg.get("bins") or 0)
        if t == 0 or sampling_kept(callcfg.get("noapprox"), callcfg.get("forceSample"), ex, t, k, samples, bins):
            if t == 0 or samples <= 0 or bins <= 0 or samples > _CALL_SAMPLES // 2:
                # the t = 0 closed form, the degenerate settings and bins too
                # large to share a call: one call each
                for c, sd in zip(todo, seeds):
                    callcfg["seed"] = sd
                    num, den = projector_norms(_bits((list(PH[c]), xk, zk), t), _bits(Hm, t), t, callcfg)
                    out[c] = 0.0 if num == 0 else base * num / den
                return out
            if engine is None:
                engine = shared_engine(device)
            num = _sampled_medians(engine, PH[todo], xk, zk, ex, k, Ls, seeds, samples, bins, fill, 0)
            if Hm[0]:
                den = _sampled_medians(engine, np.asarray([Hm[0]]), Hm[1], Hm[2], ex, k, Ls, seeds, samples,
                                       bins, fill, _H_SEED)
            else:
                den = np.full(len(todo), norm * norm)
            out[todo] = np.where(num == 0, 0.0, base * num / np.where(num == 0, 1.0, den))
            return out
        if ex:
            # no randomness: one record per distinct G', one for H'
            uniq, inv = np.unique(PH[todo], axis=0, return_inverse=True)
            inv = inv.reshape(-1)
            nnum = len(uniq)
            A = instance_array(nnum + (1 if Hm[0] else 0))
            fill(A, slice(0, nnum), uniq, xk, zk, True, k, None)
            if Hm[0]:
                fill(A, slice(nnum, nnum + 1), [Hm[0]], Hm[1], Hm[2], True, k, None)
        else:
            nnum = len(todo)
            inv = np.arange(nnum)
            A = instance_array(nnum * (2 if Hm[0] else 1))
            fill(A, slice(0, nnum), PH[todo], xk, zk, False, k, Ls)
            if Hm[0]:
                fill(A, slice(nnum, 2 * nnum), [Hm[0]], Hm[1], Hm[2], False, k, Ls)
        if engine is None:
            engine = shared_engine(device)
        vals, _ = engine.batch_exact_values(A)
        num = vals[:nnum][inv]
        if not Hm[0]:
            den = np.full(len(todo), norm * norm)
        elif ex:
            den = np.full(len(todo), vals[nnum])
        else:
            den = vals[nnum:]
        out[todo] = np.where(num == 0, 0.0, base * num / np.where(num == 0, 1.0, den))
        if trace is not None:
            trace.append({"instances": A.copy(), "values": vals, "todo": todo, "inv": inv, "nnum": nnum,
                          "base": base, "exact": bool(ex), "norm": norm, "P": out.copy()})
        return out

    def round_projectors(m, bits):
        """G' of every shot when the first m generators of order are set;
        bits (shots x m) holds the shots' outcomes for them."""
        rows = (list(G[0][:m]) + list(G[0][nq:]), list(G[1][:m]) + list(G[1][nq:]),
                list(G[2][:m]) + list(G[2][nq:]))
        prods = _products(n, rows)
        V = np.zeros((bits.shape[0], len(rows[0])), dtype=np.int64)
        V[:, :m] = bits
        return _round_phases(prods, V), prods[4]

    ncon = len(measure)
    con_bits = np.array([int(measure[q]) for q in order[:ncon]], dtype=np.int64)
    Psofar = np.ones(shots)
    if ncon:
        (PH, xk, zk, imp), uu = round_projectors(ncon, con_bits[None, :])
        P = evaluate(PH, xk, zk, imp, uu, True)[0]
        if P == 0:
            return [""] * shots
        if P < 1e-5:
            print("Warning: probability of satisfying qubit constraints very low (%f). Could be impossible." % P)
        Psofar[:] = P
    bits = np.zeros((shots, ncon + len(sample)), dtype=np.int64)
    bits[:, :ncon] = con_bits
    for r in range(len(sample)):
        m = ncon + r + 1
        (PH, xk, zk, imp), uu = round_projectors(m, bits[:, :m])
        P = evaluate(PH, xk, zk, imp, uu, cfg["exact"])
        P0 = P / Psofar
        u01 = rng.random(shots) if rng is not None else np.random.random(shots)
        draw = np.where(u01 < P0, 0, 1)
        bits[:, m - 1] = draw
        if trace is not None:
            trace.append({"round": r, "P": P.copy(), "Psofar": np.array(Psofar, dtype=float), "P0": P0.copy(),
                          "draw": draw.copy()})
        Psofar = _next_psofar(cfg.get("chain", "reference"), P, P0, Psofar, draw)
    return ["".join(str(int(b)) for b in row[ncon:]) for row in bits]
