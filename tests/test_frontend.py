"""The circuit front end (compilecirc, gadgetize, projectors, probability,
sampleQubits), pinned by brute-force state vectors.

CPU: compiled circuits act as their definitions say (unitaries checked on
every basis state); the gadget identity
    P(outcomes | x) = 2^(v-u) ||G'|H^t>||^2 / ||H'|H^t>||^2
holds for every postselection y and x, with the norms evaluated densely;
language errors are rejected.  GPU: probability() through the engine agrees
with the brute-force probability, in exact and sampled modes, and the
H (T H)^k known answers.  The circuits here are written for these tests.
"""
import itertools
import os

import numpy as np
import pytest

import circuit_sim as sim
from circuitsimulator_amd.compile import compilecirc, gadgetize, projectors

LIB = """\
// a small gate library
CZ:
_H
CX
_H

Z: // H X H
H
X
H

Ti: // T^-1 = S^3 T
S
S
S
T

CS: // controlled S = T_c T_t CX T^-1_t CX
T_
_T
CX
_Ti
CX

Q: // identity through a measured ancilla: CX, H, measure, conditional Z
CX
_H
ZM

W: // borrows an ancilla and hands it back in |0>
CX
CX
_R

V: // T on the argument when the ancilla (|+>, measured) gives 1
_H
TM
"""


@pytest.fixture
def libdir(tmp_path):
    (tmp_path / "lib.circ").write_text(LIB)
    return tmp_path


def compile_main(libdir, body, imports="import lib.circ\n"):
    f = libdir / "main.circ"
    f.write_text(imports + "main:\n" + body)
    return compilecirc.compileCircuit(fname=str(f))


def unitary_of(circuit, nargs):
    """Columns U|b> on the first nargs qubits (ancillas start and end in 0)."""
    lines = [ln for ln in circuit.splitlines() if ln]
    n = len(sim.groups(lines[0]))
    cols = []
    for b in range(2 ** nargs):
        prep = "".join("X" if (b >> q) & 1 else "_" for q in range(nargs)) + "_" * (n - nargs)
        pre = "".join(["\n".join(("_" * q + "X" + "_" * (n - q - 1)) for q in range(nargs) if (b >> q) & 1)])
        psi, _ = sim.run((pre + "\n" if pre else "") + circuit)
        cols.append(psi)
    return np.array(cols).T, n


def test_standard_gate_forms():
    # compilecirc.py:238-241
    assert compilecirc.standardGate("MTi_C") == "CTi"
    assert compilecirc.standardGate("MTi_C", listMode=True) == ["C", "Ti"]
    assert compilecirc.standardGate("MTi_C", lineMode=True) == ["M", "Ti", "_", "C"]


def test_compiled_controlled_gates_act_as_defined(libdir):
    # CZ and CS on two qubits, checked on every basis state
    for body, diag in (("CZ\n", [1, 1, 1, -1]), ("CS\n", [1, 1, 1, 1j])):
        circ = compile_main(libdir, body)
        assert set(ch for ch in circ if ch not in "\n") <= set("HXSTC_")
        U, n = unitary_of(circ, 2)
        assert n == 2
        # qubit 0 = C, qubit 1 = target: the phase sits on |11>
        assert np.allclose(U, np.diag(diag))


def test_argument_order_follows_sorted_letters(libdir):
    # "SC" calls CS with C on qubit 1 and S on qubit 0
    U, _ = unitary_of(compile_main(libdir, "SC\n"), 2)
    assert np.allclose(U, np.diag([1, 1, 1, 1j]))


def test_measured_ancilla_and_conditional_symbol(libdir):
    circ = compile_main(libdir, "H\nQ\nT\nH\n")
    lines = circ.splitlines()
    assert all(len(ln) == 2 for ln in lines)  # qubit 0 and Q's ancilla
    t, Ms, MTs = gadgetize.countY(circ)
    assert (t, Ms, MTs) == (1, [1], [])
    # the conditional Z expands to H X H, each marked with the measurement
    assert [ln for ln in lines if "M" in ln] == ["HM", "XM", "HM"]
    # identity gadget: P(q0 = 0) = |<0|H T H|0>|^2 whatever the outcome
    for x in (0, 1):
        p = sim.conditional_probability(circ, {0: 0}, {1: x})
        assert p == pytest.approx(np.cos(np.pi / 8) ** 2, abs=1e-12)


def test_reused_ancilla_is_shared(libdir):
    circ = compile_main(libdir, "W\nH\nW\n")
    assert len(circ.splitlines()[0]) == 2  # one ancilla for both calls
    circ2 = compile_main(libdir, "Q\nQ\n")
    assert len(circ2.splitlines()[0]) == 3  # fresh ancilla per call


def test_conditional_t_counts_on_outcome_1(libdir):
    circ = compile_main(libdir, "H\nV\nH\n")
    t, Ms, MTs = gadgetize.countY(circ)
    assert (t, Ms, MTs) == (0, [1], [1])
    for x, tt in (("0", 0), ("1", 1)):
        G, H, n, tcount = projectors.projectors(circ, {0: 0}, x=x, y="1" * tt)
        assert tcount == tt


def test_imports_and_comments(libdir):
    (libdir / "sub.circ").write_text("from lib.circ import CZ Ti\n\nY: // unused\nX\n")
    f = libdir / "m2.circ"
    f.write_text("import sub.circ\n/* block\ncomment */ \nmain:\n H _\nCZ // spaces are dropped\n")
    circ = compilecirc.compileCircuit(fname=str(f))
    assert circ.splitlines()[0] == "H_"
    with pytest.raises(SyntaxError):
        compilecirc.compileCircuit(raw="main:\nH\n/* a */\n")  # '*/' after '/*' on one line
    with pytest.raises(SyntaxError):
        compilecirc.compileCircuit(raw="main:\nH\nHH\n")  # width changes
    with pytest.raises(SyntaxError):
        compilecirc.compileCircuit(raw="main:\nHQ\n")  # undefined symbol
    with pytest.raises(SyntaxError):
        compilecirc.compileCircuit(raw="AB:\nBA\nmain:\nAB\n")  # calls itself
    with pytest.raises(SyntaxError):
        compilecirc.compileCircuit(raw="main:\nHM\n")  # measures a non-ancilla
    (libdir / "a.circ").write_text("import b.circ\n")
    (libdir / "b.circ").write_text("import a.circ\n")
    with pytest.raises(SyntaxError):
        compilecirc.compileCircuit(fname=str(libdir / "a.circ"))  # import loop


def test_countY_rejects_gates_on_measured_qubits():
    with pytest.raises(SyntaxError):
        gadgetize.countY("_H\nCX\nHM\nH_\n_H\n")


def random_clifford_t(rng, n, ngates, nt):
    gates = []
    for _ in range(ngates):
        kind = rng.integers(0, 3)
        if kind == 2 and n > 1:
            c, t = rng.choice(n, size=2, replace=False)
            ln = ["_"] * n
            ln[c], ln[t] = "C", "X"
        else:
            ln = ["_"] * n
            ln[rng.integers(0, n)] = "HS"[min(kind, 1)]
        gates.append("".join(ln))
    for _ in range(nt):
        at = int(rng.integers(0, len(gates) + 1))
        ln = ["_"] * n
        ln[rng.integers(0, n)] = "T"
        gates.insert(at, "".join(ln))
    return "\n".join(gates) + "\n"


def valid_ts(circ):
    try:
        gadgetize.countY(circ)
        return True
    except SyntaxError:
        return False


def gadget_probability(circ, measure, x, y):
    m = dict(measure)
    G, H, n, t = projectors.projectors(circ, m, x=x, y=y)
    Gp, u = projectors.truncate(n, G)
    Hp, v = projectors.truncate(n, H)
    return 2.0 ** (v - u) * sim.magic_norm2(Gp, t) / sim.magic_norm2(Hp, t)


@pytest.mark.parametrize("seed", range(6))
def test_gadget_identity_random_circuits(seed):
    rng = np.random.default_rng(40 + seed)
    n = int(rng.integers(2, 5))
    while True:
        circ = random_clifford_t(rng, n, int(rng.integers(6, 16)), int(rng.integers(1, 5)))
        if valid_ts(circ):
            break
    t = gadgetize.countY(circ)[0]
    psi, _ = sim.run(circ)
    for meas in ({0: 0}, {0: 1}, {0: 0, n - 1: 1}):
        if len(meas) > 1 and n < 2:
            continue
        want = sim.conditional_probability(circ, meas, {})
        for ybits in itertools.islice(itertools.product("01", repeat=t), 4):
            got = gadget_probability(circ, meas, "", "".join(ybits))
            assert got == pytest.approx(want, abs=1e-12)


def test_gadget_identity_with_measurements(libdir):
    circ = compile_main(libdir, "H_\nCS\nV_\n_H\n_Q\nCZ\nH_\n")
    t, Ms, MTs = gadgetize.countY(circ)
    assert Ms and MTs
    for xbits in itertools.product("01", repeat=len(Ms)):
        x = "".join(xbits)
        xd = {M: int(b) for M, b in zip(Ms, xbits)}
        tt = t + sum(1 for M in MTs if xd[M] == 1)
        for meas in ({0: 0}, {1: 1}, {0: 1, 1: 0}):
            want = sim.conditional_probability(circ, meas, xd)
            for ybits in itertools.islice(itertools.product("01", repeat=tt), 3):
                got = gadget_probability(circ, meas, x, "".join(ybits))
                assert got == pytest.approx(want, abs=1e-12)


def test_truncate_keeps_negative_identity():
    # Z_0 Z_1 and -Z_0 Z_1 on 1 magic qubit: the product is -I there
    P = ([0, 2], [np.array([0, 0]), np.array([0, 0])], [np.array([1, 1]), np.array([1, 1])])
    Pp, u = projectors.truncate(1, P)
    assert 2 in Pp[0]


def test_ht_stack_known_answers_dense():
    # libcirc main.py:177-190: P(0) of H (T H)^k
    H = sim.GATES["H"]
    T = sim.GATES["T"]
    for k, quoted in [(1, 0.8536), (2, 0.75), (3, 0.5732), (4, 0.9786)]:
        circ = "H\n" + "T\nH\n" * k
        out = H
        for _ in range(k):
            out = H @ T @ out
        exact = abs(out[0, 0]) ** 2
        assert exact == pytest.approx(quoted, abs=1e-4)
        got = gadget_probability(circ, {0: 0}, "", "0" * k)
        assert got == pytest.approx(exact, abs=1e-12)


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_probability_exact_on_engine(seed):
    from circuitsimulator_amd.probability import probability
    rng = np.random.default_rng(70 + seed)
    n = int(rng.integers(2, 5))
    while True:
        circ = random_clifford_t(rng, n, int(rng.integers(8, 20)), int(rng.integers(2, 7)))
        if valid_ts(circ):
            break
    for meas in ({0: 0}, {n - 1: 1}):
        want = sim.conditional_probability(circ, meas, {})
        got = probability(circ, dict(meas), config={"noapprox": True, "quiet": True, "seed": 5 + seed,
                                                    "rng": np.random.default_rng(seed)})
        assert got == pytest.approx(want, abs=1e-12)


@pytest.mark.gpu
def test_probability_ht_stack_on_engine():
    from circuitsimulator_amd.probability import probability
    H, T = sim.GATES["H"], sim.GATES["T"]
    for k in range(1, 9):
        circ = "H\n" + "T\nH\n" * k
        out = H
        for _ in range(k):
            out = H @ T @ out
        got = probability(circ, {0: 0}, config={"noapprox": True, "quiet": True, "seed": k,
                                                "rng": np.random.default_rng(k)})
        assert got == pytest.approx(abs(out[0, 0]) ** 2, abs=1e-12)


@pytest.mark.gpu
def test_probability_sampled_on_engine():
    # an |L> decomposition sampled with many draws lands near the exact value
    from circuitsimulator_amd.probability import probability
    rng = np.random.default_rng(91)
    while True:
        circ = random_clifford_t(rng, 3, 14, 6)
        if valid_ts(circ):
            break
    want = sim.conditional_probability(circ, {0: 0}, {})
    got = probability(circ, {0: 0}, config={"exact": False, "k": 3, "samples": 20000, "bins": 1,
                                            "forceSample": True, "quiet": True, "seed": 17,
                                            "rng": np.random.default_rng(3)})
    assert got == pytest.approx(want, abs=0.05)


@pytest.mark.gpu
def test_sample_qubits_on_engine():
    # a Bell pair: both qubits always agree
    from circuitsimulator_amd.sample import sampleQubits
    circ = "H_\nCX\n"
    for s in range(4):
        out = sampleQubits(circ, {}, [0, 1], config={"quiet": True, "noapprox": True, "exact": True, "seed": s,
                                                     "rng": np.random.default_rng(s)})
        assert out in ("00", "11")
