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


def _next_psofar(chain, P, P0, Psofar, bit):
    """Psofar after drawing bit (elementwise for arrays)."""
    if chain == "joint":
        return np.where(bit == 0, P, Psofar - P)
    if chain != "reference":
        raise ValueError("chain must be 'reference' or 'joint'")
    return np.where(bit == 0, P0, 1 - P0)
import numpy as np

from .compile import gadgetize as _gadgetize
from .compile import projectors as _projectors
from .probability import decompose, probability


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
    A = np.array(xs, dtype=int)[:, :n]
    X = _projectors._null_space(A)
    ph, xm, zm = [], [], []
    for row in X:
        gen = (0, np.zeros(size, dtype=int), np.zeros(size, dtype=int))
        for i in range(len(row)):
            if row[i] == 1:
                gen = _projectors._multiply(gen, (phases[i], np.asarray(xs[i], dtype=int),
                                                  np.asarray(zs[i], dtype=int)))
        ph.append(gen[0])
        xm.append(sum(int(b) << j for j, b in enumerate(gen[1][n:])))
        zm.append(sum(int(b) << j for j, b in enumerate(gen[2][n:])))
    X = np.array(X, dtype=int).reshape(len(X), len(phases))
    return np.array(ph, dtype=int), xm, zm, X, len(phases) - len(X)


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


def _bits(P, t):
    """0/1 arrays of a mask-form projector (for probability()'s helpers)."""
    ph, xm, zm = P
    return (ph, [np.array([(x >> j) & 1 for j in range(t)]) for x in xm],
            [np.array([(z >> j) & 1 for j in range(t)]) for z in zm])


def sampleShots(circ, measure, sample, shots, config={}):
    """shots independent sampleQubits(circ, measure, sample, config) draws
    (libcirc/sample.py:16-86), advanced together one sampled qubit per
    round.

    Circuit and gadgets are conjugated once: each generator's conjugation
    through the circuit is independent of the others, so G for any set of
    outcomes is a subset of the rows of one gadgetization, with the phase
    +2 on outcome 1.  Per round the null space of the truncation and the
    generator masks are shared by all shots (they depend on the X and Z
    parts only) and a shot's G' differs in its phases.  Each call's
    decomposition is drawn afresh, as probability() does, and all numerators
    and denominators of a round are evaluated by one batched engine launch
    (cse_batch_exact) whenever the reference's rule evaluates them exactly
    (probability.c:162-174; with the default 2000 samples this is every |L>
    of k <= 11 and every exact |H> of t <= 22); otherwise each norm is
    sampled on its own.

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
    from .engine import NMAX, Engine, Instance, Projector, sampling_kept
    from .probability import _pick_sampling, projector_norms

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

    def same_as_H(P):
        # probability.py:140-150: the first len(G') generators of H' agree
        return len(P[0]) <= len(Hm[0]) and all(
            P[0][i] == Hm[0][i] and P[1][i] == Hm[1][i] and P[2][i] == Hm[2][i] for i in range(len(P[0])))

    def evaluate(calls, uu, exact):
        """P of each call; calls hold one G' (mask form) or None each."""
        nonlocal engine
        out = [0.0] * len(calls)
        inst, pending = [], []
        for c, P in enumerate(calls):
            if P is None:
                continue  # -identity: P = 0
            if same_as_H(P):
                out[c] = 2.0 ** (v - uu)
                continue
            callcfg = dict(cfg)
            # an empty projector needs the exact decomposition's norm
            callcfg["exact"] = bool(exact) or len(P[0]) == 0 or len(Hm[0]) == 0
            seed = int(rand_int(2 ** 62, 1)[0])
            L, norm, ex, k = decompose(t, callcfg, seed)
            if t == 0 or sampling_kept(callcfg.get("noapprox"), callcfg.get("forceSample"), ex, t, k,
                                       int(callcfg.get("samples") or 0), int(callcfg.get("bins") or 0)):
                # sampled norms, or the t = 0 closed form: one call each
                callcfg["seed"] = seed
                num, den = projector_norms(_bits(P, t), _bits(Hm, t), t, callcfg)
                out[c] = 0.0 if num == 0 else 2.0 ** (v - uu) * num / den
                continue
            slots = []
            for Q in (P, Hm):
                if len(Q[0]) == 0:
                    slots.append(None)
                    continue
                ins = Instance()
                ins.P = Projector.make(Q[0], Q[1], Q[2], t)
                ins.exact = 1 if ex else 0
                ins.k = k
                for i, r in enumerate(L[:NMAX]):
                    ins.L[i] = r
                slots.append(len(inst))
                inst.append(ins)
            pending.append((c, norm, slots))
        if inst:
            if engine is None:
                engine = Engine(device)
            res, _, _ = engine.batch_exact(inst)
            vals = [r.value for r in res]
            for c, norm, (snum, sden) in pending:
                num = norm * norm if snum is None else vals[snum]
                den = norm * norm if sden is None else vals[sden]
                out[c] = 0.0 if num == 0 else 2.0 ** (v - uu) * num / den
        return out

    def round_projectors(m, bits):
        """G' of every shot when the first m generators of order are set;
        bits[s] holds shot s's outcomes for them."""
        rows = (list(G[0][:m]) + list(G[0][nq:]), list(G[1][:m]) + list(G[1][nq:]),
                list(G[2][:m]) + list(G[2][nq:]))
        prods = _products(n, rows)
        pad = len(rows[0]) - m
        return [_shot_projector(prods, np.array(list(b) + [0] * pad, dtype=int)) for b in bits], prods[4]

    try:
        ncon = len(measure)
        con_bits = [int(measure[q]) for q in order[:ncon]]
        Psofar = np.ones(shots)
        if ncon:
            calls, uu = round_projectors(ncon, [con_bits])
            P = evaluate(calls, uu, True)[0]
            if P == 0:
                return [""] * shots
            if P < 1e-5:
                print("Warning: probability of satisfying qubit constraints very low (%f). Could be impossible." % P)
            Psofar[:] = P
        bits = [list(con_bits) for _ in range(shots)]
        out = [""] * shots
        for r in range(len(sample)):
            calls, uu = round_projectors(ncon + r + 1, [b + [0] for b in bits])
            P = np.array(evaluate(calls, uu, cfg["exact"]))
            P0 = P / Psofar
            u01 = rng.random(shots) if rng is not None else np.random.random(shots)
            draw = np.where(u01 < P0, 0, 1)
            for s in range(shots):
                bits[s].append(int(draw[s]))
                out[s] += str(int(draw[s]))
            Psofar = _next_psofar(cfg.get("chain", "reference"), P, P0, Psofar, draw)
        return out
    finally:
        if engine is not None:
            engine.close()
