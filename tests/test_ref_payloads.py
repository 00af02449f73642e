"""The front end and the drop-in's payload against the reference's own
Python (tests/golden/ref_payloads.json, written by
tests/golden/make_payloads.py from libcirc in the build container): for 16
seeded circuits and configurations -- the C1 circuit, random Clifford+T
circuits, and circuit-language sources with measured ancillas, a
conditional T gate and a reused ancilla --

  - compileCircuit gives the reference's compiled text (language cases);
  - projectors() gives the reference's G, H, n, t for the same x, y
    (libcirc/compile/projectors.py:19-84);
  - truncate() gives the reference's G', u and H', v (projectors.py:87-162);
  - probability() with config["file"] writes the reference's stdin payload
    byte for byte (probability.py:216-280), or returns the reference's value
    where it stops before writing one (-I in G', G' = H').

On the GPU every payload goes through the backend executable, whose last
two lines are checked against the oracle's exact projector sums."""
import json
import os

import numpy as np
import pytest

from circuitsimulator_amd.compile import compilecirc, projectors

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "ref_payloads.json")))["cases"]
IDS = [c["name"] for c in CASES]


def _lists(P):
    ph, xs, zs = P
    return [[int(v) for v in ph], [[int(b) for b in x] for x in xs], [[int(b) for b in z] for z in zs]]


def _measure(c):
    return {int(q): b for q, b in c["measure"].items()}


def test_fixture_covers_the_cases():
    assert len(CASES) == 16
    assert any(c["source"] for c in CASES) and any(c["name"] == "random-1001" for c in CASES)
    assert any(len(c["Hprime"][0]) == 0 for c in CASES)          # an empty H'
    assert any(c["config"].get("noapprox") for c in CASES)
    assert any(c["config"].get("forceSample") for c in CASES)


@pytest.mark.parametrize("c", CASES, ids=IDS)
def test_compiled_circuit_matches_reference(c):
    if c["source"] is None:
        pytest.skip("given compiled")
    assert compilecirc.compileCircuit(raw=c["source"]) == c["circuit"]


@pytest.mark.parametrize("c", CASES, ids=IDS)
def test_projectors_and_truncation_match_reference(c):
    G, H, n, t = projectors.projectors(c["circuit"], _measure(c), x=c["x"], y=c["y"])
    assert (n, t) == (c["n"], c["t"])
    assert _lists(G) == c["G"] and _lists(H) == c["H"]
    Gp, u = projectors.truncate(n, G)
    Hp, v = projectors.truncate(n, H)
    assert _lists(Gp) == c["Gprime"] and u == c["u"]
    assert _lists(Hp) == c["Hprime"] and v == c["v"]


@pytest.mark.parametrize("c", CASES, ids=IDS)
def test_payload_matches_reference_byte_for_byte(c, tmp_path):
    # or, where the reference returns before writing one (G' holds -I, or
    # G' = H'), the same value and no file
    from circuitsimulator_amd.probability import probability
    path = tmp_path / "payload.txt"
    cfg = dict(c["config"], file=str(path), x=c["x"], y=c["y"])
    cfg.setdefault("quiet", True)
    got = probability(c["circuit"], _measure(c), config=cfg)
    if c["payload"] is None:
        assert not path.exists()
        assert got == c["result"]
    else:
        assert got is None
        assert path.read_text() == c["payload"]


def _fields(payload):
    f = payload.split()
    keys = ["quiet", "verbose", "noapprox", "samples", "bins", "t", "k", "exact", "fidbound", "fidelity", "rank",
            "forceL", "forceSample"]
    return {key: (float(v) if key == "fidbound" else int(v)) for key, v in zip(keys, f)}


def _evaluated_exactly(c):
    """payloads with the exact |H^t> decomposition whose norms the backend
    sums exactly: noapprox, or the sampling rule of probability.c:162-174
    switching sampling off"""
    if c["payload"] is None:
        return False
    f = _fields(c["payload"])
    if not f["exact"]:
        return False
    return bool(f["noapprox"]) or (not f["forceSample"] and
                                   f["samples"] * f["bins"] * 2 > 2 ** ((f["t"] + 1) // 2) - 1)


def _oracle_norm(P, t, norm):
    """exactProjector (innerprod.c:148-199) by the oracle's exact pair sums
    over the |H^t> decomposition: |sum|, or norm^2 for an empty P"""
    from gpu_util import oracle_exact_terms, term_value
    ph, xs, zs = P
    if len(ph) == 0:
        return norm * norm
    xi = [sum(b << j for j, b in enumerate(x)) for x in xs]
    zi = [sum(b << j for j, b in enumerate(z)) for z in zs]
    tot = 0
    for (pt, s, i, j) in oracle_exact_terms(t, 0, [], 1, ph, xi, zi):
        if pt:
            v = term_value(pt, s)
            tot += v if i == j else 2 * v.real
    return abs(tot)


EXACT = [c for c in CASES if _evaluated_exactly(c)]


def test_exact_payload_cases_present():
    assert len(EXACT) >= 4


@pytest.mark.gpu
@pytest.mark.parametrize("c", EXACT, ids=[c["name"] for c in EXACT])
def test_backend_on_reference_payloads(c):
    # the reference writer's payload through the backend executable: its
    # last two lines equal the oracle's exact sums of the same G', H'
    from circuitsimulator_amd.protocol import run_backend
    t = c["t"]
    num, den, _ = run_backend(c["payload"])
    norm = 2.0 ** ((t // 2) / 2.0) * (2 * np.cos(np.pi / 8) if t % 2 else 1.0)
    assert num == pytest.approx(_oracle_norm(c["Gprime"], t, norm), rel=1e-13, abs=1e-13)
    assert den == pytest.approx(_oracle_norm(c["Hprime"], t, norm), rel=1e-13, abs=1e-13)
