"""MI355X (gfx950) stabilizer-rank inner-product engine.

A drop-in for the C backend of patrickrall/CircuitSimulator: the executable
`cse_backend` speaks the reference's stdin/stdout protocol
(libcirc/probability.c:44-221), and libcse.so exposes the engine through the
C ABI of include/cse.h.  `engine` is the Python binding of that ABI.

The circuit front end mirrors the reference's Python interface:
compile.compilecirc.compileCircuit, compile.projectors.projectors/truncate,
probability.probability and sample.sampleQubits, with the engine as backend.
"""
from . import engine  # noqa: F401
from .engine import Engine, EngineError, Projector  # noqa: F401

__all__ = ["engine", "Engine", "EngineError", "Projector"]
