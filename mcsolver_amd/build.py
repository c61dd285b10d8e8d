"""In-tree build of libmcgpu.so for gfx950 (hipcc), and of the CPU oracle
(make -C oracle; test infrastructure, built beside the product)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmcgpu.so")
SRC = [os.path.join(HERE, "csrc", "mcgpu.hip")]
DEPS = SRC + [os.path.join(HERE, "csrc", "mc_math.h"), os.path.join(ROOT, "include", "mcgpu.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-std=c++17", "-O3", "-fPIC", "-shared",
    # bit-reproducible FP32: no contraction of a*b+c into fma behind our back
    "-ffp-contract=off",
    # no packed-FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build_lib(force: bool = True, verbose: bool = False) -> str:
    """Compile mcsolver_amd/libmcgpu.so (always recompiles when force)."""
    if not force and os.path.exists(LIB):
        lm = os.path.getmtime(LIB)
        if all(os.path.getmtime(d) <= lm for d in DEPS):
            return LIB
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + SRC
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmcgpu.so")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(with_reference: bool = True, force: bool = True) -> None:
    """make -C oracle (our restatement + checker); with_reference also builds
    oracle/_ref from /root/reference when that tree exists here. force
    recompiles even when the outputs are newer than their sources."""
    od = os.path.join(ROOT, "oracle")
    mk = ["make", "-s"] + (["-B"] if force else []) + ["-C", od]
    r = subprocess.run(mk, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    if with_reference and os.path.isdir("/root/reference/mcsolver"):
        r = subprocess.run(mk + ["ref"], capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("reference build (oracle/_ref) failed")


if __name__ == "__main__":
    build_lib(verbose=True)
    build_oracle()
    print("built", LIB)
