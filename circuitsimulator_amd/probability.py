"""probability(circ, measure, config): the probability of measurement outcomes
of a compiled circuit (libcirc/probability.py:67-328), computed with the
gfx950 engine in this process.

The circuit is gadgetized into projectors G (requested outcomes) and H
(circuit measurements only), both truncated to the t magic qubits, and

    P = 2^(v-u) ||G'|H^t>||^2 / ||H'|H^t>||^2

with u, v the generators lost in truncation.  The norms are computed as the
reference's backend computes them (probability.c:153-216): decompose picks
|H^t> exactly (chi = 2^ceil(t/2)) or an |L> of rank k, the sampling-disable
rule may switch to exact evaluation, and each norm is exact or a median of
means of sampled estimates -- all through cse_decompose,
cse_sampling_kept and cse_projector_norms of libcse.

config keys are the reference's (see libcirc/probability.py:14-65):
verbose, silenceprojectors, quiet, direct, noapprox, samples, bins, error,
failprob, exact, k, fidbound, fidelity, rank, file, x, y, forceL,
forceSample.  python, cpath, mpirun and procs select the reference's
backends and are ignored: the engine is the backend.  Added keys: seed (the
Philox key of L and of the random states; default: drawn from numpy's global
generator, like x and y), rng (a numpy Generator for x and y), device.

Deviations, each to avoid a crash or a silent None of the reference: a
Clifford circuit (t = 0) is evaluated with the closed form the C backend
uses (innerprod.c:157-167) instead of returning None; an empty G' passes its
width from t instead of raising IndexError (probability.py:259).
"""
import numpy as np

from .compile import projectors as _projectors
from .engine import Projector, lib, projector_norm, projector_norms_batch, sampling_kept

_SEED_MASK = (1 << 64) - 1


def printProjector(projector):
    """Print generators as +/-(i) Pauli strings (probability.py:431-454)."""
    phases, xs, zs = projector
    for g in range(len(phases)):
        phase = phases[g]
        text = ""
        for i in range(len(xs[g])):
            if xs[g][i] == 1 and zs[g][i] == 1:
                phase -= 1  # Y = i^-1 Z X
                text += "Y"
            elif xs[g][i] == 1:
                text += "X"
            elif zs[g][i] == 1:
                text += "Z"
            else:
                text += "_"
        print({0: " +", 1: " i", 2: " -", 3: "-i"}[phase % 4] + text)
    if len(phases) == 0:
        print("Projector is empty.")


def _engine_projector(P, t):
    """Projector of (phases, xs, zs) with 0/1 rows, or of the bitmask form
    (phases, xs, zs, t) with int rows"""
    if len(P) == 4:
        ph, xs, zs, _ = P
        return Projector.make([int(p) for p in ph], list(xs), list(zs), t)
    ph, xs, zs = P
    return Projector.make([int(p) for p in ph], _projectors.masks(xs, t), _projectors.masks(zs, t), t)


def decompose(t, config, seed=0):
    """(L rows as ints, norm, exact, k) as the backend's decompose
    (probability.c:307-415, probability.py:332-428): the same k auto-pick,
    exact fallback, k <= t clamp, rank and fidelity options."""
    import ctypes
    from .engine import NMAX, U64
    L = (U64 * NMAX)()
    norm = ctypes.c_double()
    exact = ctypes.c_int(1 if config.get("exact") else 0)
    k = ctypes.c_int(int(config.get("k") or 0))
    rc = lib().cse_decompose(int(t), seed & _SEED_MASK, L, ctypes.byref(norm), ctypes.byref(exact),
                             ctypes.byref(k), float(config.get("fidbound") or 1e-5),
                             1 if config.get("rank") else 0, 1 if config.get("fidelity") else 0,
                             1 if config.get("forceL") else 0, 1 if config.get("verbose") else 0,
                             1 if config.get("quiet") else 0)
    if rc != 0:
        raise ValueError("decomposition of t=%d is outside the engine's range" % t)
    rows = [int(L[i]) for i in range(k.value)] if not exact.value else []
    return rows, norm.value, bool(exact.value), k.value


def decompose_many(t, config, seeds):
    """decompose once per seed with the same config, in one call
    (cse_decompose_many): (L rows as a uint64 array of shape (len(seeds), k),
    or None for the exact decomposition, norm, exact, k).  exact and k depend
    on t and config only; norm is the last seed's (it differs between seeds
    only with the fidelity option)."""
    import ctypes
    from .engine import NMAX, U64
    n = len(seeds)
    sd = np.array([int(s) & _SEED_MASK for s in seeds], dtype=np.uint64)
    L = np.zeros((max(n, 1), NMAX), dtype=np.uint64)
    norms = np.zeros(max(n, 1))
    exact = ctypes.c_int(1 if config.get("exact") else 0)
    k = ctypes.c_int(int(config.get("k") or 0))
    rc = lib().cse_decompose_many(int(t), sd.ctypes.data_as(ctypes.POINTER(U64)), n,
                                  L.ctypes.data_as(ctypes.POINTER(U64)), NMAX,
                                  norms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(exact),
                                  ctypes.byref(k), float(config.get("fidbound") or 1e-5),
                                  1 if config.get("rank") else 0, 1 if config.get("fidelity") else 0,
                                  1 if config.get("forceL") else 0, 1 if config.get("verbose") else 0,
                                  1 if config.get("quiet") else 0)
    if rc != 0:
        raise ValueError("decomposition of t=%d is outside the engine's range" % t)
    if exact.value:
        return None, float(norms[n - 1]), True, k.value
    return np.ascontiguousarray(L[:n, :k.value]), float(norms[n - 1]), False, k.value


def _same(Gp, Hp):
    # probability.py:140-150, on bitmask rows: every generator of G' equals
    # H''s at the same index (G' no longer than H')
    same = True
    for i in range(len(Gp[0])):
        if i > len(Hp[0]) - 1:
            same = False
            continue
        same = same and Gp[0][i] == Hp[0][i] and Gp[1][i] == Hp[1][i] and Gp[2][i] == Hp[2][i]
    return same


def _pick_sampling(config, t, quiet, verbose):
    # probability.py:155-193
    if not config.get("samples"):
        config["samples"] = 2000
    if not config.get("bins"):
        config["bins"] = 1
    if not config.get("error"):
        config["error"] = 0.2
    if not config.get("failprob"):
        config["failprob"] = 0.05
    if config["error"] != 0.2 or config["failprob"] != 0.05:
        chi = (2 ** t - 1) / (2 ** t + 1)
        if config["failprob"] < 0.0076:  # median of means is worth it
            config["samples"] = int(np.ceil(6 * chi * config["error"] ** (-2)))
            config["bins"] = int(np.ceil(4.5 * np.log(1 / config["failprob"])))
        else:
            config["samples"] = int(np.ceil(chi * config["error"] ** (-2) * config["failprob"] ** (-1)))
            config["bins"] = 1
        if not quiet:
            print("Autopicking median of %d bins with %d samples per bin." % (config["bins"], config["samples"]))
    if verbose:
        print("Evaluating median of %d bins with %d samples per bin." % (config["bins"], config["samples"]))


def projector_norms(Gprime, Hprime, t, config):
    """(||G'|H^t>||^2, ||H'|H^t>||^2) as the backend computes them
    (probability.c:153-205)."""
    verbose = config.get("verbose")
    seed = config.get("seed")
    if seed is None:
        seed = int(np.random.randint(0, 2 ** 62))
    L, norm, exact, k = decompose(t, config, seed)
    if verbose:
        if exact:
            print("Using exact decomposition of |H^t>: 2^%d" % ((t + 1) // 2))
        else:
            print("Stabilizer rank of |L>: 2^%d" % k)
    samples = int(config.get("samples") or 0)
    bins = int(config.get("bins") or 0)
    sampled = sampling_kept(config.get("noapprox"), config.get("forceSample"), exact, t, k, samples, bins)
    if not config.get("noapprox") and not sampled and verbose:
        print("More samples than terms in exact calculation. Disabling sampling.")
    device = int(config.get("device") or 0)
    # both norms in one call (one batched launch for exact pair sums)
    num, den = projector_norms_batch([_engine_projector(Gprime, t), _engine_projector(Hprime, t)],
                                     [seed, seed ^ 0x9E3779B97F4A7C15], L, k, exact, norm, sampled, samples,
                                     bins, device)
    return num, den


def probability(circ, measure, config={}):
    """Probability of the outcomes in measure (qubit index -> 0/1) for the
    compiled circuit circ (probability.py:67-328)."""
    verbose = config.get("verbose")
    if verbose:
        config["quiet"] = False
    quiet = config.get("quiet")
    if config.get("exact") is None:
        config["exact"] = True
    if config.get("fidbound") is None:
        config["fidbound"] = 1e-5
    # with k given, fidelity is only displayed; quiet switches it off
    if not config.get("exact") and config.get("k") is not None:
        if config.get("quiet") and config.get("fidelity"):
            config["fidelity"] = False

    # the projectors as bitmask rows (arrays only for printing and payloads)
    G, H, n, t = _projectors.projectors_masks(circ, measure, verbose=verbose, x=config.get("x"),
                                              y=config.get("y"), rng=config.get("rng"))
    Gprime, u = _projectors.truncate_masks(n, G)
    Hprime, v = _projectors.truncate_masks(n, H)

    if verbose and not config.get("silenceprojectors"):
        print("Gadgetize circuit:")
        print("G:")
        printProjector(_projectors.to_arrays(G))
        print("H:")
        printProjector(_projectors.to_arrays(H))
        print("Truncate projectors to magic state space:")
        print("Gprime: (truncated by %d)" % u)
        printProjector(_projectors.to_arrays(Gprime))
        print("Hprime: (truncated by %d)" % v)
        printProjector(_projectors.to_arrays(Hprime))

    # -I in the numerator: the outcome is impossible
    for i in range(len(Gprime[0])):
        if Gprime[0][i] == 2 and not Gprime[1][i] and not Gprime[2][i]:
            if verbose:
                print("Found negative identity. Answer is 0.")
            return 0

    if _same(Gprime, Hprime):
        if verbose:
            print("Projectors are identical.")
        return 2 ** (v - u)

    # an empty projector needs the exact decomposition's norm (the
    # reference's precedence: G' empty, or H' empty without exact)
    if len(Gprime[0]) == 0 or (len(Hprime[0]) == 0 and not config["exact"]):
        if not quiet:
            print("Empty projectors found. Using exact decomposition to compute norm.")
        config["exact"] = True

    if not config.get("noapprox"):
        _pick_sampling(config, t, quiet, verbose)

    if config.get("file"):
        from .protocol import make_payload
        with open(config["file"], "w") as f:
            f.write(make_payload(_projectors.to_arrays(Gprime), _projectors.to_arrays(Hprime), config, t=t))
        if not quiet:
            print("Wrote instruction file to %s." % config["file"])
        return None

    numerator, denominator = projector_norms(Gprime, Hprime, t, config)

    if verbose:
        print("|| Gprime |H^t> ||^2 ~= " + str(numerator))
        print("|| Hprime |H^t> ||^2 ~= " + str(denominator))
    if config.get("direct"):
        return (2 ** (v - u) * numerator, denominator)
    if numerator == 0:
        return 0
    return 2 ** (v - u) * (numerator / denominator)
