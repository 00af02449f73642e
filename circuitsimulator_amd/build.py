"""In-tree build of the engine for gfx950.

Products (next to this file):
  libcse.so     HIP kernels + C ABI (include/cse.h)
  cse_backend   drop-in backend executable speaking the reference's
                stdin/stdout protocol (libcirc/probability.c:44-221)

Every HIP source is compiled for gfx950 only, with packed-FP32 instructions
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I", INCLUDE]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu99", "-Wall", "-Wextra", "-I", INCLUDE]

LIB = os.path.join(HERE, "libcse.so")
BACKEND = os.path.join(HERE, "cse_backend")


def _run(cmd, quiet=False):
    if not quiet:
        print("+ " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    if out.strip() and not quiet:
        print(out)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    host_o = os.path.join(BUILD, "cse_host.o")
    eng_o = os.path.join(BUILD, "cse_engine.o")
    norms_o = os.path.join(BUILD, "cse_norms.o")
    _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, "cse_host.c"), "-o", host_o], not verbose)
    _run(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, "cse_norms.c"), "-o", norms_o], not verbose)
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "cse_engine.hip"), "-o", eng_o], not verbose)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, eng_o, host_o, norms_o, "-lm"], not verbose)
    os.replace(tmp, LIB)
    btmp = BACKEND + ".tmp"
    _run(["gcc", *C_FLAGS, "-o", btmp, os.path.join(CSRC, "backend_main.c"),
          "-L", HERE, "-lcse", "-Wl,-rpath,$ORIGIN", "-lm"], not verbose)
    os.replace(btmp, BACKEND)
    return LIB


if __name__ == "__main__":
    build()
