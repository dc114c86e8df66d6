"""Build the HIP transport library (gemini_amd/libgemhip.so) for gfx950 with hipcc.

Flags that matter for results:
  -ffp-contract=off   no fused multiply-add: every product and sum is rounded on its
                      own, as in the reference's x86-64 build (bit-identical parity),
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions anywhere in the device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgemhip.so")
SOURCES = ["gemhip_api.hip", "kernels_advec.hip", "kernels_diffusion.hip"]
ARCH = os.environ.get("GEMHIP_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Compile every source into the shared library (always recompiles).

    ``defines`` are extra ``NAME=VALUE`` preprocessor settings (tile sizes, pipeline
    depths) for tuning builds written to ``out``."""
    cmd = [hipcc(), "-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-std=c++17",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-result", "-o", out + ".tmp"]
    cmd += [f"-D{d}" for d in defines]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgemhip.so")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
