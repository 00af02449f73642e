"""The circuit language and its compiler to the gate set {H, X, S, T, CX}
with measurements (the language of libcirc/compile/compilecirc.py).

Language.  A circuit text holds symbols.  A heading ``NAME:`` opens one;
each following line is a step with one token per qubit: an upper-case
letter with an optional lower-case suffix (``Ti``) is a gate acting on that
position, ``_`` leaves the qubit alone, ``M`` marks the position as
measured (the step then runs only when that outcome is 1), ``R`` hands the
position back in |0> for later calls of the same symbol.  A step whose
gate -- its tokens sorted and joined, ``_`` and ``M`` left out -- is not in
the gate set calls the symbol of that name, its tokens in sorted order
naming the callee's first positions; the callee's remaining positions are
ancillas.  ``//`` and ``/* ... */`` are comments, ``import FILE`` and
``from FILE import A B`` bring in symbols of other files (paths relative to
the importing file).

Design.  The text goes through three stages:
  1. a line reader that removes comments (``_decomment``);
  2. a symbol table of ``Symbol`` objects with their parsed steps
     (``_SymbolReader``, imports merged as whole dependency closures),
     checked for undefined calls and call cycles by one memoised walk
     (``_check_calls``) and for operations on measured or handed-back
     positions (``Symbol.check_positions``);
  3. an expansion of the compiled symbol by an explicit stack of call frames
     (``_expand``): an allocator hands out ancilla wires in call order (the
     fresh ancillas of every call, and the handed-back ones of a symbol once,
     at its first call), gates are collected as sparse rows, and the rows
     are rendered at the end, one character per wire, main's qubits first.

The wire numbering and the error messages are those of the reference, which
callers see; elsewhere this compiler differs where the reference fails:
errors raise SyntaxError instead of calling sys.exit
(compilecirc.py:28-29), an import brings in the symbols its imported
symbols need at every depth instead of one level, positions of measured and
handed-back qubits count tokens rather than characters, and a callee whose
name carries a lower-case suffix gets its ancillas.
"""
import os
import re

# gates that need no symbol (compilecirc.py:67)
GATE_SET = ("H", "X", "S", "T", "CX")

# a token: one character that is not a lower-case letter, with the
# lower-case letters that follow it
_TOKEN = re.compile(r"[^a-z][a-z]*")
_IDLE_OR_MEASURED = str.maketrans("", "", "_M")


def standardGate(line, listMode=False, lineMode=False):
    """A step in standard form (compilecirc.py:242-259): its tokens with '_'
    and 'M' left out, sorted and joined ("MTi_C" -> "CTi"); listMode gives
    the sorted tokens (["C", "Ti"]), lineMode every token in position order,
    '_' and 'M' included (["M", "Ti", "_", "C"])."""
    if line == "main":
        if listMode or lineMode:
            raise ValueError("Can't bring 'main' into list form.")
        return "main"
    if lineMode:
        return _TOKEN.findall(line)
    tokens = sorted(_TOKEN.findall(line.translate(_IDLE_OR_MEASURED)))
    return tokens if listMode else "".join(tokens)


def _decomment(line, in_block, where):
    """One line without its comments: a '//' comment, then a '*/' closing an
    open block, then a '/*' opening one (one delimiter of each kind per
    line); spaces are dropped except on import lines.  Returns the line and
    whether a block comment is open after it."""
    line = line.split("//", 1)[0]
    if "*/" in line:
        if not in_block:
            raise SyntaxError(where + "Unexpected '*/'.")
        line, in_block = line.split("*/", 1)[1], False
    elif in_block:
        if "/*" in line:
            raise SyntaxError(where + "Unexpected '/*' inside block comment.")
        return "", True
    if "/*" in line:
        line, in_block = line.split("/*", 1)[0], True
    return (line if "import" in line else line.replace(" ", "")), in_block


class Step:
    """One line of a symbol: its tokens in position order and its gate (the
    standard form; a name outside the gate set is a call)."""
    __slots__ = ("text", "tokens", "gate")

    def __init__(self, text):
        self.text = text
        self.tokens = standardGate(text, lineMode=True)
        self.gate = standardGate(text)

    @property
    def is_call(self):
        return self.gate not in GATE_SET and self.gate != "R"

    def arguments(self):
        """positions of the call's tokens in sorted token order (equal tokens
        in position order): the callee's positions 0, 1, ..."""
        used = [i for i, tok in enumerate(self.tokens) if tok not in ("_", "M")]
        return sorted(used, key=lambda i: self.tokens[i])


class Symbol:
    """A symbol: its name, qubit count, steps and the symbols it calls."""

    def __init__(self, name):
        self.name = name
        self.width = 0
        self.steps = []
        self.calls = []          # one entry per calling step
        self.handed_back = []    # positions marked 'R', in order of appearance

    @property
    def nargs(self):
        """positions taken by arguments: one per token of the name; main's
        first four positions count as such (the length of its name), as in
        the reference's measurement and hand-back checks"""
        if self.name == "main":
            return len(self.name)
        return len(standardGate(self.name, listMode=True))

    def add_step(self, text, where):
        # a step's width counts every character but lower-case letters and
        # digits
        w = sum(1 for ch in text if not (ch.islower() or ch.isdigit()))
        if self.width == 0:
            self.width = w
        elif w != self.width:
            raise SyntaxError(where + "Bad number of qubits (expected %d)." % self.width)
        step = Step(text)
        self.steps.append(step)
        if step.is_call:
            self.calls.append(step.gate)

    def check_positions(self):
        """A measured position takes no operation after its measurement line
        (it may be measured again), a handed-back one none at all; only
        ancilla positions (past the name's tokens) may be measured or handed
        back.  Sets handed_back."""
        measured, back = [], []
        for step in self.steps:
            tok = step.tokens
            for q in measured:
                if q < len(tok) and tok[q] not in ("_", "M"):
                    raise SyntaxError("In %s: cannot perform operation on measured qubit %d on line %s"
                                      % (self.name, q, step.text))
            if "M" in tok:
                q = tok.index("M")
                if q < self.nargs:
                    raise SyntaxError("In %s: cannot measure non-ancilla qubit on line %s" % (self.name, step.text))
                measured.append(q)
        for step in self.steps:
            tok = step.tokens
            for q in back:
                if q < len(tok) and tok[q] != "_":
                    raise SyntaxError("In %s: cannot perform operation on reused qubit %d on line %s"
                                      % (self.name, q, step.text))
            if "R" in tok:
                q = tok.index("R")
                if q < self.nargs:
                    raise SyntaxError("In %s: cannot reuse non-ancilla qubit on line %s" % (self.name, step.text))
                back.append(q)
        self.handed_back = back


class _SymbolReader:
    """Reads a circuit text into a symbol table (name -> Symbol, in order of
    definition), following its imports; chain is the list of files whose
    imports led here."""

    def __init__(self, path, fname, chain):
        self.path = os.path.normpath(path)
        self.origin = os.path.join(self.path, fname)
        if self.origin in chain:
            raise SyntaxError("Dependency loop detected: " + " -> ".join(chain + [self.origin]))
        self.chain = chain
        self.table = {}
        self.current = None

    def read(self, text):
        in_block = False
        where = "Error at " + self.origin + ":0 - "
        for number, raw in enumerate(text.splitlines(), start=1):
            where = "Error at %s:%d - " % (self.origin, number)
            line, in_block = _decomment(raw, in_block, where)
            if not line:
                continue
            if "import" in line:
                self.merge_import(line, where)
            elif ":" in line:
                self.open_symbol(line[:-1], where)
            elif self.current is None:
                raise SyntaxError(where + "Expected function declaration.")
            else:
                self.current.add_step(line, where)
        if in_block:
            raise SyntaxError(where + "Expected '*/' before end of file.")
        return self.table

    def open_symbol(self, name, where):
        std = standardGate(name)
        if name != std:
            raise SyntaxError(where + "Symbol %s is not in standard form (expected %s)." % (name, std))
        if name in GATE_SET:
            raise SyntaxError(where + "Symbol %s is in gate set." % name)
        if name in self.table:
            raise SyntaxError(where + "Symbol %s already defined." % name)
        self.current = self.table[name] = Symbol(name)

    def merge_import(self, line, where):
        words = line.split()
        if words[0] == "import":
            wanted = []
        elif words[0] == "from" and len(words) > 2 and words[2] == "import":
            wanted = words[3:]
        else:
            raise SyntaxError(where + "Invalid import statement.")
        target = os.path.join(self.path, words[1])
        try:
            with open(target) as f:
                text = f.read()
        except FileNotFoundError:
            raise SyntaxError(where + "File %s not found." % target)
        sub = _SymbolReader(os.path.dirname(target), os.path.basename(target), self.chain + [self.origin]).read(text)
        chosen = [name for name in sub if not wanted or name in wanted]
        # a chosen symbol may not clash with one defined here; the symbols it
        # calls (at every depth) come along unless already present
        for name in chosen:
            if name in self.table:
                raise SyntaxError(where + "Can't import %s because it is already defined!" % name)
        missing = [w for w in wanted if w not in sub]
        if missing:
            raise SyntaxError(where + "Failed to import symbol(s): " + " ".join("'" + w + "'" for w in missing))
        todo = list(chosen)
        while todo:
            name = todo.pop(0)
            if name in self.table or name not in sub:
                continue
            self.table[name] = sub[name]
            todo.extend(sub[name].calls)


def _check_calls(table):
    """Every called symbol exists and no symbol reaches itself.  The walk
    starts from each call of each symbol in turn, depth first; a symbol whose
    walk finished cleanly is not walked again."""
    clean = set()
    for sym in table.values():
        for start in sym.calls:
            if start in clean:
                continue
            path = [start]
            pending = []      # per path entry: the calls still to follow
            while path:
                name = path[-1]
                if len(pending) < len(path):
                    if name not in table:
                        raise SyntaxError("Undefined gate: %s" % name)
                    pending.append(list(table[name].calls))
                todo = pending[-1]
                while todo and todo[0] in clean:
                    todo.pop(0)
                if not todo:
                    clean.add(name)
                    path.pop()
                    pending.pop()
                    continue
                nxt = todo.pop(0)
                if nxt in path:
                    raise SyntaxError("Dependency loop: %s -> %s" % (str(path), nxt))
                path.append(nxt)


def read_symbols(text, path="", fname="<raw input>"):
    """The checked symbol table of a circuit text (imports resolved relative
    to path)."""
    table = _SymbolReader(path, fname, []).read(text)
    _check_calls(table)
    for sym in table.values():
        sym.check_positions()
    return table


class _Wires:
    """Ancilla wires, handed out in call order from the first wire after the
    compiled symbol's qubits: every call takes fresh wires for its ancilla
    positions; a symbol's handed-back positions take wires once, at its
    first call, before its fresh ones, and keep them for its later calls."""

    def __init__(self, first):
        self.next = first
        self.kept = {}

    def take(self):
        q = self.next
        self.next += 1
        return q

    def for_call(self, callee, args):
        if callee.name not in self.kept:
            self.kept[callee.name] = [self.take() for _ in callee.handed_back]
        kept = self.kept[callee.name]
        wires = list(args)
        for pos in range(callee.nargs, callee.width):
            if pos in callee.handed_back:
                wires.append(kept[callee.handed_back.index(pos)])
            else:
                wires.append(self.take())
        return wires


def _expand(table, name):
    """The gates of symbol `name` with every call inlined, as (rows, wire
    count): a row is a list of (wire, token) pairs.  A call's gates carry
    the measured wire of the calling step (and of every enclosing
    conditional call) as an 'M'."""
    root = table[name]
    wires = _Wires(root.width)
    rows = []
    # frame: (symbol, its wires, the measured wires it runs under, next step)
    stack = [(root, list(range(root.width)), (), 0)]
    while stack:
        sym, w, marks, at = stack.pop()
        if at == len(sym.steps):
            continue
        stack.append((sym, w, marks, at + 1))
        step = sym.steps[at]
        if step.gate == "R":
            continue
        if not step.is_call:
            row = [(w[pos], tok) for pos, tok in enumerate(step.tokens[:len(w)])]
            rows.append(row + [(q, "M") for q in marks])
            continue
        callee = table[step.gate]
        cw = wires.for_call(callee, [w[pos] for pos in step.arguments()])
        cmarks = marks + ((w[step.tokens.index("M")],) if "M" in step.tokens else ())
        stack.append((callee, cw, cmarks, 0))
    return rows, wires.next


def compile_symbol(table, name="main"):
    """The compiled circuit of one symbol: one line per gate, one character
    per wire (the symbol's qubits, then the ancillas), '_' on idle wires."""
    if name not in table:
        raise SyntaxError("Could not find symbol %s to compile" % name)
    rows, width = _expand(table, name)
    out = []
    for row in rows:
        line = ["_"] * width
        for q, tok in row:
            line[q] = tok
        out.append("".join(line) + "\n")
    return "".join(out)


def compileCircuit(raw=None, fname=None, name="main"):
    """Compile a circuit given as text (raw) or as a file (fname); the
    interface of compilecirc.py:12-29."""
    if raw is None and fname is None:
        raise ValueError("Must specify either a file or raw circuit data.")
    if raw is not None:
        return compile_symbol(read_symbols(raw), name)
    with open(fname) as f:
        text = f.read()
    return compile_symbol(read_symbols(text, os.path.dirname(fname), os.path.basename(fname)), name)
