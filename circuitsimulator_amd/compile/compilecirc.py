"""The circuit language: symbol tables, imports and compilation to the gate
set {H, X, S, T, CX} with measurements (libcirc/compile/compilecirc.py).

A circuit file holds symbols.  A heading ``NAME:`` opens a symbol; every
following line is one step, one character group per qubit: an upper-case
letter plus any lower-case suffix (``Ti``) names a gate acting on that
position, ``_`` is an idle qubit, ``M`` marks the qubit measured on that line
(the step then runs only when the outcome is 1), ``R`` marks a qubit handed
back in |0> (it is reused by later calls of the same symbol).  A step whose
gate (its letters sorted, ``_`` and ``M`` dropped) is not in the gate set
calls the symbol of that name on the positions of its letters, in sorted
order; the symbol's extra positions are ancillas.  ``//`` and ``/* */``
comment; ``import file`` and ``from file import A B`` pull symbols from
other files (paths relative to the importing file).

compileCircuit returns the compiled circuit: one line per gate, one
character per qubit (main's qubits first, then the ancillas).  Unlike the
reference, which calls sys.exit on a syntax error (compilecirc.py:28-29),
errors raise SyntaxError.
"""
import os

# gates that need no symbol (compilecirc.py:67)
GATE_SET = ("H", "X", "S", "T", "CX")


def standardGate(line, listMode=False, lineMode=False):
    """Standard form of a step (compilecirc.py:242-259): the letter groups
    (an upper-case letter and its lower-case suffix), sorted and joined;
    listMode returns the sorted groups, lineMode the groups in position
    order with '_' and 'M' kept."""
    if line == "main":
        if listMode or lineMode:
            raise ValueError("Can't bring 'main' into list form.")
        return "main"
    if not lineMode:
        line = line.replace("_", "").replace("M", "")
    groups = []
    for ch in line:
        if ch.islower():
            groups[-1] += ch
        else:
            groups.append(ch)
    if lineMode:
        return groups
    groups = sorted(groups)
    return groups if listMode else "".join(groups)


def _strip_comments(line, in_block, where):
    """compilecirc.py:35-63: cut '//' comments, track '/* */' blocks (one
    delimiter per line), and drop spaces unless the line imports."""
    cut = line.find("//")
    if cut != -1:
        line = line[:cut]
    close = line.find("*/")
    if close != -1:
        if not in_block:
            raise SyntaxError(where + "Unexpected '*/'.")
        line = line[close + 2:]
        in_block = False
    opening = line.find("/*")
    if in_block:
        if opening != -1:
            raise SyntaxError(where + "Unexpected '/*' inside block comment.")
        return "", in_block
    if opening != -1:
        line = line[:opening]
        in_block = True
    if "import" not in line:
        line = line.replace(" ", "")
    return line, in_block


def _import(statement, path, fname, past, output, where):
    """An import line (compilecirc.py:89-136): the imported symbols and the
    direct dependencies of each."""
    words = statement.split()
    if words[0] == "import":
        wanted = []
    elif words[0] == "from" and len(words) > 2 and words[2] == "import":
        wanted = words[3:]
    else:
        raise SyntaxError(where + "Invalid import statement.")
    target = os.path.join(path, words[1])
    try:
        with open(target) as f:
            text = f.read()
    except FileNotFoundError:
        raise SyntaxError(where + "File %s not found." % target)
    table = circuitSymbols(text, os.path.dirname(target), os.path.basename(target),
                           past + [os.path.join(path, fname)])
    got, pulled_in = [], []
    for name, sym in table.items():
        if wanted and name not in wanted:
            continue
        if name in output and name not in pulled_in:
            raise SyntaxError(where + "Can't import %s because it is already defined!" % name)
        output[name] = sym
        got.append(name)
        for dep in sym["dependencies"]:
            if dep not in output:
                output[dep] = table[dep]
                pulled_in.append(dep)
                if dep in wanted:
                    got.append(dep)
    if wanted and len(wanted) != len(got):
        missing = ["'" + w + "'" for w in wanted if w not in got]
        raise SyntaxError(where + "Failed to import symbol(s): " + " ".join(missing))


def circuitSymbols(data, path, fname, past):
    """Symbol table of a circuit text (compilecirc.py:72-236): for each
    symbol its steps, the symbols it calls (once per call), the string
    positions it reuses, and its qubit count."""
    path = os.path.normpath(path)
    here = os.path.join(path, fname)
    if here in past:
        raise SyntaxError("Dependency loop detected: " + " -> ".join(past + [here]))
    output = {}
    in_block = False
    current = ""
    nq = 0
    for number, raw in enumerate(data.splitlines(), start=1):
        where = "Error at " + here + ":" + str(number) + " - "
        line, in_block = _strip_comments(raw, in_block, where)
        if not line:
            continue
        if "import" in line:
            _import(line, path, fname, past, output, where)
            continue
        if ":" in line:
            current = line[:-1]
            if current != standardGate(current):
                raise SyntaxError(where + "Symbol %s is not in standard form (expected %s)."
                                  % (current, standardGate(current)))
            if current in GATE_SET:
                raise SyntaxError(where + "Symbol %s is in gate set." % current)
            if current in output:
                raise SyntaxError(where + "Symbol %s already defined." % current)
            nq = 0
            output[current] = {"steps": [], "dependencies": [], "reused": [], "n": 0}
            continue
        if current == "":
            raise SyntaxError(where + "Expected function declaration.")
        width = sum(1 for ch in line if not ch.islower() and not ch.isdigit())
        if nq == 0:
            nq = width
            output[current]["n"] = nq
        elif width != nq:
            raise SyntaxError(where + "Bad number of qubits (expected %d)." % nq)
        gate = standardGate(line)
        if gate not in GATE_SET + ("R",):
            output[current]["dependencies"].append(gate)
        output[current]["steps"].append(line)
    if in_block:
        raise SyntaxError(where + "Expected '*/' before end of file.")

    # every called symbol is defined, and calls form no loop
    checked = set()

    def walk(chain):
        if chain[-1] not in output:
            raise SyntaxError("Undefined gate: %s" % chain[-1])
        for nxt in output[chain[-1]]["dependencies"]:
            if nxt in chain:
                raise SyntaxError("Dependency loop: %s -> %s" % (str(chain), nxt))
            walk(chain + [nxt])

    for sym in output.values():
        for dep in sym["dependencies"]:
            if dep not in checked:
                walk([dep])
                checked.add(dep)

    # measured and reused qubits are left alone afterwards; only ancillas
    # (string positions past the symbol's name) may be measured or reused.
    # Positions are string indices, as in the reference.
    for name, sym in output.items():
        measured = []
        for line in sym["steps"]:
            for pos in measured:
                if line[pos] not in ("_", "M"):
                    raise SyntaxError("In %s: cannot perform operation on measured qubit %d on line %s"
                                      % (name, pos, line))
            pos = line.find("M")
            if pos == -1:
                continue
            if pos < len(name):
                raise SyntaxError("In %s: cannot measure non-ancilla qubit on line %s" % (name, line))
            measured.append(pos)
    for name, sym in output.items():
        reused = []
        for line in sym["steps"]:
            for pos in reused:
                if line[pos] != "_":
                    raise SyntaxError("In %s: cannot perform operation on reused qubit %d on line %s"
                                      % (name, pos, line))
            pos = line.find("R")
            if pos == -1:
                continue
            if pos < len(name):
                raise SyntaxError("In %s: cannot reuse non-ancilla qubit on line %s" % (name, line))
            reused.append(pos)
        sym["reused"] = reused
    return output


def _all_calls(symbols, calls):
    """Every symbol call made, directly or through other symbols, counted
    with multiplicity (compilecirc.py:271-281)."""
    out = []
    frontier = list(calls)
    while frontier:
        nxt = [d for c in frontier for d in symbols[c]["dependencies"]]
        out += nxt
        frontier = nxt
    return list(calls) + out


def compile(symbols, name="main"):
    """Compiled circuit of a symbol (compilecirc.py:262-372)."""
    if name not in symbols:
        raise SyntaxError("Could not find symbol %s to compile" % name)
    n = symbols[name]["n"]
    counts = {}
    for dep in _all_calls(symbols, symbols[name]["dependencies"]):
        counts[dep] = counts.get(dep, 0) + 1

    # ancillas: fresh ones per call, reused ones once per symbol
    ancillas = 0
    for dep, cnt in counts.items():
        reused = len(symbols[dep]["reused"])
        fresh = symbols[dep]["n"] - len(standardGate(dep, listMode=True)) - reused
        ancillas += fresh * cnt + reused
    total = n + ancillas
    state = {"free": list(range(n, n + ancillas)), "owned": {dep: [] for dep in counts}}

    def build(line, qubits):
        groups = standardGate(line, lineMode=True)
        gate = standardGate(line)
        if gate == "R":
            return ""
        if gate in GATE_SET:
            row = ["_"] * total
            for pos in range(len(qubits)):
                row[qubits[pos]] = groups[pos]
            return "".join(row) + "\n"
        # the symbol's arguments: the qubits under its letters, sorted
        args = []
        for g in standardGate(line, listMode=True):
            at = groups.index(g)
            args.append(qubits[at])
            groups[at] = " "
        sym = symbols[gate]
        reused = sym["reused"]
        owned = state["owned"][gate]
        while len(reused) > len(owned):
            owned.append(state["free"].pop(0))
        for pos in range(len(gate), sym["n"]):
            if pos in reused:
                args.append(owned[reused.index(pos)])
                continue
            q = state["free"].pop(0)
            args.append(q)
            owned.append(q)
        text = ""
        for step in sym["steps"]:
            out = build(step, args)
            if "M" in groups:
                # the call is conditioned on a measurement: so is every gate
                mq = qubits[groups.index("M")]
                marked = ""
                for ln in out.splitlines():
                    ln = list(ln)
                    ln[mq] = "M"
                    marked += "".join(ln) + "\n"
                out = marked
            text += out
        return text

    return "".join(build(step, list(range(n))) for step in symbols[name]["steps"])


def compileCircuit(raw=None, fname=None, name="main"):
    """Compile a circuit given as text (raw) or as a file (fname)
    (compilecirc.py:12-29)."""
    if raw is None and fname is None:
        raise ValueError("Must specify either a file or raw circuit data.")
    if raw is None:
        with open(fname) as f:
            data = f.read()
        symbols = circuitSymbols(data, os.path.dirname(fname), os.path.basename(fname), [])
    else:
        symbols = circuitSymbols(raw, "", "<raw input>", [])
    return compile(symbols, name=name)
