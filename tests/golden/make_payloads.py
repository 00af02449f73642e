"""Writes tests/golden/ref_payloads.json: what the reference's own Python
front end produces for a set of seeded circuits, so that the front end and
the backend payload of this package can be compared with it byte for byte.

Run in the build container only, where the reference checkout is present
(nothing of the reference goes anywhere else; the fixture holds data only):

    python tests/golden/make_payloads.py [/root/reference]

For each case it records, from the reference's libcirc modules:
  - the compiled circuit (libcirc/compile/compilecirc.py compileCircuit, for
    the cases given in the circuit language);
  - G, H, n, t of libcirc/compile/projectors.py projectors() with fixed x, y;
  - G', u and H', v of truncate() (projectors.py:87-162);
  - the stdin payload libcirc/probability.py writes with config["file"]
    (probability.py:216-280; with "file" set it returns before running any
    backend), or the value it returns before writing one (0 when G' holds
    -I, 2^(v-u) when G' = H', probability.py:120-138).
The circuits and configurations are seeded and written for this fixture.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

# circuit-language cases: a small library of symbols written for this
# fixture (measured ancillas, a conditional T, a reused ancilla)
LANG_LIB = """\
CZ:
_H
CX
_H

Z:
H
X
H

Q:
CX
_H
ZM

V:
_H
TM

W:
CX
CX
_R
"""

LANG_CASES = [
    # (source of main, measure, x, y, config); one gate per line
    ("main:\nH__\nT__\nCZ_\n_V_\n_H_\n__T\n__Q\nT__\nH__\n", {0: 1}, "10", None, {"noapprox": True}),
    ("main:\nH_\n_H\nT_\n_T\nCZ\nV_\nZ_\n_T\nH_\nQ_\nT_\n", {1: 0}, "01", None,
     {"exact": False, "k": 2, "forceSample": True, "samples": 40, "bins": 3}),
    ("main:\nH__\nCX_\n_CX\nT__\n__T\n_W_\nH__\n__V\nT__\n_T_\n__T\nH__\n", {0: 0, 2: 1}, "1", None,
     {"exact": True}),
]


def to_lists(P):
    ph, xs, zs = P
    return [[int(v) for v in ph], [[int(b) for b in x] for x in xs], [[int(b) for b in z] for z in zs]]


def main():
    refroot = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    sys.dont_write_bytecode = True   # nothing is written under the reference
    sys.path.insert(0, refroot)
    sys.path.insert(1, ROOT)
    import tempfile

    import numpy as np
    from libcirc.compile import compilecirc as ref_compile
    from libcirc.compile import projectors as ref_proj
    from libcirc.probability import probability as ref_probability

    from circuitsimulator_amd.workloads import c1_circuit

    cases = []
    # random Clifford+T circuits: (seed, n, gates, T count, measure, config)
    rand = [
        (1001, 8, 40, 6, {0: 0, 1: 0}, {"noapprox": True, "exact": True}),
        (2001, 5, 30, 4, {0: 1}, {"exact": True}),
        (2002, 6, 36, 5, {2: 0, 4: 1}, {"exact": False, "k": 3, "samples": 100, "bins": 3, "forceSample": True}),
        (2003, 7, 50, 8, {1: 1}, {"exact": False, "k": 4, "samples": 20, "bins": 2}),
        (2004, 4, 24, 3, {0: 0, 3: 0}, {"noapprox": True, "exact": False, "k": 2, "forceL": True}),
        (2005, 9, 60, 7, {5: 1}, {"error": 0.1, "failprob": 0.01}),
        (2006, 6, 40, 6, {0: 1, 1: 1}, {"failprob": 0.001, "exact": False, "k": 3, "rank": True}),
        (2007, 10, 80, 9, {9: 0}, {"exact": False, "fidbound": 1e-3, "verbose": False, "quiet": True}),
        (2008, 8, 60, 8, {0: 0, 7: 1}, {"exact": True, "samples": 10, "bins": 5}),
        (2009, 6, 30, 5, {3: 1}, {"noapprox": True}),
        (2010, 9, 70, 10, {2: 0, 4: 0, 6: 1}, {"exact": False, "k": 5, "samples": 16, "bins": 1,
                                                "forceSample": True, "fidelity": True}),
        (2011, 5, 20, 4, {1: 0}, {"exact": False, "noapprox": True, "k": 2, "rank": True}),
        (2012, 7, 45, 7, {0: 1, 6: 0}, {"quiet": False}),
    ]
    for seed, n, gates, ts, measure, config in rand:
        circ, y = c1_circuit(seed=seed, n=n, gates=gates, ts=ts)
        cases.append({"name": "random-%d" % seed, "source": None, "circuit": circ, "measure": measure,
                      "x": "", "y": y, "config": config})
    rng = np.random.default_rng(3001)
    for i, (src, measure, x, _, config) in enumerate(LANG_CASES):
        circ = ref_compile.compileCircuit(raw=LANG_LIB + src)
        t, Ms, MTs = ref_proj.gadgetize.countY(circ)
        tt = t + sum(1 for M in MTs if x[Ms.index(M)] == "1")
        y = "".join(str(int(b)) for b in rng.integers(0, 2, size=tt))
        cases.append({"name": "language-%d" % i, "source": LANG_LIB + src, "circuit": circ, "measure": measure,
                      "x": x, "y": y, "config": config})

    out = []
    with tempfile.TemporaryDirectory() as tmp:
        for c in cases:
            G, H, n, t = ref_proj.projectors(c["circuit"], dict(c["measure"]), x=c["x"], y=c["y"])
            Gp, u = ref_proj.truncate(n, G)
            Hp, v = ref_proj.truncate(n, H)
            # a fresh file per case: probability() returns before writing
            # when G' holds -I (0) or G' = H' (2^(v-u)); that value is kept
            path = os.path.join(tmp, "payload-%s.txt" % c["name"])
            cfg = dict(c["config"], file=path, x=c["x"], y=c["y"])
            cfg.setdefault("quiet", True)
            result = ref_probability(c["circuit"], dict(c["measure"]), config=cfg)
            payload = None
            if os.path.exists(path):
                with open(path) as f:
                    payload = f.read()
            out.append({
                "name": c["name"], "source": c["source"], "circuit": c["circuit"],
                "measure": {str(q): int(b) for q, b in c["measure"].items()},
                "x": c["x"], "y": c["y"], "config": c["config"],
                "n": int(n), "t": int(t), "G": to_lists(G), "H": to_lists(H),
                "Gprime": to_lists(Gp), "u": int(u), "Hprime": to_lists(Hp), "v": int(v),
                "payload": payload, "result": None if result is None else float(result),
            })
            print("%s: n=%d t=%d |G'|=%d u=%d |H'|=%d v=%d, %s"
                  % (c["name"], n, t, len(Gp[0]), u, len(Hp[0]), v,
                     "payload %d bytes" % len(payload) if payload is not None else "returned %r" % result))
    with open(os.path.join(HERE, "ref_payloads.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_payloads.py", "cases": out}, f, indent=1)


if __name__ == "__main__":
    main()
