"""The C ABI boundary: libcse.so loads on a machine without a GPU and
exports every entry point include/cse.h declares; the product code never
links or imports the oracle."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cse.h")
LIB = os.path.join(ROOT, "circuitsimulator_amd", "libcse.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cse_\w+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libcse.so not built")
def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 12
    lib = ctypes.CDLL(LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from circuitsimulator_amd.engine import ABI_SYMBOLS
    assert sorted(ABI_SYMBOLS) == names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libcse.so not built")
def test_no_gpu_create_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    from circuitsimulator_amd import Engine, EngineError
    with pytest.raises(EngineError):
        Engine(0)


def test_product_does_not_reference_the_oracle():
    # the oracle is test infrastructure; the product never names it
    pkg = os.path.join(ROOT, "circuitsimulator_amd")
    for dirpath, _, files in os.walk(pkg):
        if "_build" in dirpath or "__pycache__" in dirpath:
            continue
        for f in files:
            if f.endswith((".py", ".c", ".h", ".hip")):
                src = open(os.path.join(dirpath, f)).read()
                assert "liborc" not in src and "cse_oracle" not in src and "import orc" not in src, f
                assert "orc_" not in src, f


@pytest.mark.skipif(not os.path.exists(LIB), reason="libcse.so not built")
def test_library_links_no_oracle():
    out = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "liborc" not in out and "libref" not in out
    syms = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True).stdout
    assert " orc_" not in syms and " ref_" not in syms
