"""Client side of the backend's stdin/stdout protocol.

The payload and the output handling follow the reference's C-backend caller
(libcirc/probability.py:216-312): 13 whitespace-separated parameters, then
the truncated projectors G' and H' (each: Nstabs, Nqubits, then per generator
a phase 0..3 followed by x z for every qubit); the backend prints
informational lines and then numerator and denominator as the last two
lines.  `run_backend` plays the role of the Popen block in probability.py and
raises RuntimeError when the last two lines are not numbers, as that caller
does.
"""
import os
import subprocess

DEFAULT_BACKEND = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cse_backend")


def _line(v):
    if v is None:
        v = 0
    if v is True:
        v = 1
    if v is False:
        v = 0
    return f"{v}\n"


def projector_payload(projector):
    """projector = (phases, xs, zs) with xs[g], zs[g] lists of 0/1 per qubit
    (the truncated-projector tuples of libcirc/compile/projectors.py)."""
    ph, xs, zs = projector
    out = _line(len(ph))
    out += _line(len(xs[0]) if len(ph) > 0 else 0)
    for g in range(len(ph)):
        out += _line(int(ph[g]))
        for j in range(len(xs[g])):
            out += _line(int(xs[g][j]))
            out += _line(int(zs[g][j]))
    return out


def make_payload(G, H, config=None, t=None):
    """stdin payload (probability.py:247-272).  config keys as in the
    reference's probability() config dict; unset values are sent as 0, as
    the reference's send() does (samples and bins included: probability()
    fills them in only when sampling is allowed)."""
    c = dict(config or {})
    if t is None:
        t = len(G[1][0]) if len(G[0]) else (len(H[1][0]) if len(H[0]) else 0)
    s = ""
    s += _line(1 if c.get("quiet") else 0)
    s += _line(1 if c.get("verbose") else 0)
    s += _line(1 if c.get("noapprox") else 0)
    s += _line(c.get("samples"))
    s += _line(c.get("bins"))
    s += _line(t)
    s += _line(c.get("k"))
    s += _line(1 if c.get("exact", True) else 0)
    s += _line(c.get("fidbound") if c.get("fidbound") else 1e-5)
    s += _line(1 if c.get("fidelity") else 0)
    s += _line(1 if c.get("rank") else 0)
    s += _line(1 if c.get("forceL") else 0)
    s += _line(1 if c.get("forceSample") else 0)
    s += projector_payload(G)
    s += projector_payload(H)
    return s


def run_backend(payload, cpath=DEFAULT_BACKEND, env=None, argv=("stdin",), timeout=600):
    """Runs the backend on the payload; returns (numerator, denominator,
    informational lines).  Raises RuntimeError like probability.py:302-312
    when the output does not end in two numbers."""
    p = subprocess.run([cpath, *argv], input=payload.encode(), stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, env=env, timeout=timeout)
    lines = p.stdout.decode().splitlines()
    info = []
    for ln in lines:
        try:
            float(ln)
        except ValueError:
            info.append(ln)
    try:
        num, den = float(lines[-2]), float(lines[-1])
    except (IndexError, ValueError):
        raise RuntimeError("C code encountered error. Aborting. (exit %d): %s"
                           % (p.returncode, " | ".join(lines[-3:])))
    return num, den, info
