"""Brute-force state-vector model of compiled circuits, for tests.

Qubit q is bit q of the amplitude index.  A compiled line is one gate from
{H, X, S, T, CX} (C marks the control); a line with an M at qubit m runs only
when m's outcome is 1.  P(outcomes | circuit measurements x) =
||Pi_{outcomes, x} psi_x||^2 / ||Pi_x psi_x||^2, with psi_x the final state
when every M line runs according to x (its qubit is never touched after the
line, so measuring at the end is the same measurement).
"""
import numpy as np

SQ = 1 / np.sqrt(2)
GATES = {
    "H": np.array([[SQ, SQ], [SQ, -SQ]], dtype=complex),
    "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "S": np.array([[1, 0], [0, 1j]], dtype=complex),
    "T": np.array([[1, 0], [0, np.exp(1j * np.pi / 4)]], dtype=complex),
}


def groups(line):
    out = []
    for ch in line:
        if ch.islower():
            out[-1] += ch
        else:
            out.append(ch)
    return out


def apply_1q(psi, n, q, U):
    psi = psi.reshape([2] * n)
    ax = n - 1 - q  # bit q of the index = axis n-1-q of the reshaped array
    psi = np.moveaxis(np.tensordot(U, psi, axes=([1], [ax])), 0, ax)
    return psi.reshape(-1)


def apply_cx(psi, n, c, t):
    idx = np.arange(len(psi))
    sel = ((idx >> c) & 1) == 1
    out = psi.copy()
    out[idx[sel]] = psi[idx[sel] ^ (1 << t)]
    return out


def run(circuit, x=None):
    """Final state of a compiled circuit; x maps measured qubits to the
    outcome that decides their M lines."""
    lines = [ln for ln in circuit.splitlines() if ln]
    n = len(groups(lines[0]))
    psi = np.zeros(2 ** n, dtype=complex)
    psi[0] = 1
    for ln in lines:
        g = groups(ln)
        if "M" in g and (x or {}).get(g.index("M"), 0) == 0:
            continue
        ops = [(q, c) for q, c in enumerate(g) if c not in ("_", "M")]
        if len(ops) == 1:
            psi = apply_1q(psi, n, ops[0][0], GATES[ops[0][1]])
        else:
            c = [q for q, ch in ops if ch == "C"][0]
            t = [q for q, ch in ops if ch == "X"][0]
            psi = apply_cx(psi, n, c, t)
    return psi, n


def project(psi, outcomes):
    idx = np.arange(len(psi))
    keep = np.ones(len(psi), dtype=bool)
    for q, v in outcomes.items():
        keep &= ((idx >> q) & 1) == v
    return np.where(keep, psi, 0)


def conditional_probability(circuit, measure, x):
    psi, n = run(circuit, x)
    den = np.vdot(project(psi, x), project(psi, x)).real
    joint = dict(x)
    joint.update(measure)
    pj = project(psi, joint)
    return np.vdot(pj, pj).real / den


def magic_norm2(P, t):
    """||prod_g (1 + g)/2 |H>^t||^2, |H> = cos(pi/8)|0> + sin(pi/8)|1>,
    g = i^phase Z(z) X(x)."""
    h = np.array([np.cos(np.pi / 8), np.sin(np.pi / 8)])
    psi = np.array([1.0 + 0j])
    for _ in range(t):
        psi = np.kron(h, psi)  # qubit j = bit j
    idx = np.arange(2 ** t)
    for ph, xs, zs in zip(*P):
        xm = sum(int(b) << j for j, b in enumerate(xs))
        zm = sum(int(b) << j for j, b in enumerate(zs))
        sign = np.array([(-1) ** bin(i & zm).count("1") for i in idx])
        gpsi = (1j ** ph) * sign * psi[idx ^ xm]
        psi = (psi + gpsi) / 2
    return np.vdot(psi, psi).real
