"""sampleQubits(circ, measure, sample, config): draw the outcomes of the
qubits in sample, one qubit at a time from conditional probabilities
(libcirc/sample.py:16-86), each probability computed by probability() on
the gfx950 engine.

The draw order and the arithmetic are the reference's: with constraints in
measure, their probability Psofar is computed first (exact decomposition);
then for each sampled qubit in turn P(qubit = 0, all so far) is computed,
P0 = P / Psofar, the qubit is drawn against a uniform number, and Psofar
becomes P0 or 1 - P0 -- the conditional probability of the outcome just
drawn, as in the reference.  exact defaults to False here, and the file
option is disabled.  config may carry rng (a numpy Generator) for the
uniform draws; numpy's global generator is used otherwise, as in the
reference.
"""
import numpy as np

from .probability import probability


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
        return bit, (P0 if bit == 0 else 1 - P0)

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
