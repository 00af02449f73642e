"""Circuit front end: the circuit language (compilecirc), T gadgets
(gadgetize) and the truncated projectors (projectors)."""
