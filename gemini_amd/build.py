"""Build the HIP transport library (gemini_amd/libgemhip.so) for gfx950 with hipcc.

The HDF5 file layer (csrc/h5io.cpp) is compiled against hdf5.h and loads libhdf5 when a
file is first read or written, so the library has no load-time dependency on it.

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
SOURCES = ["gemhip_api.hip", "kernels_advec.hip", "kernels_diffusion.hip", "kernels_adjoint.hip",
           "kernels_block.hip", "kernels_sources.hip", "kernels_output.hip", "h5io.cpp"]
ARCH = os.environ.get("GEMHIP_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def hdf5_include():
    """Directory holding hdf5.h (the HDF5 C library's public header).  The library itself is
    loaded at run time, so only its declarations are needed to build."""
    for d in (os.environ.get("GEMHIP_HDF5_INCLUDE"), "/opt/conda/include", "/usr/include/hdf5/serial",
              "/usr/include"):
        if d and os.path.exists(os.path.join(d, "hdf5.h")):
            return d
    raise RuntimeError("hdf5.h not found; set GEMHIP_HDF5_INCLUDE to the HDF5 include directory")


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Compile every source into the shared library (always recompiles).

    ``defines`` are extra ``NAME=VALUE`` preprocessor settings (tile sizes, pipeline
    depths) for tuning builds written to ``out``."""
    cmd = [hipcc(), "-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-std=c++17",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-result", "-idirafter", hdf5_include(), "-o", out + ".tmp"]
    cmd += [f"-D{d}" for d in defines]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    cmd += ["-ldl"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgemhip.so")
    os.replace(out + ".tmp", out)
    return out


ROOT = os.path.dirname(HERE)
DRIVERS = os.path.join(ROOT, "tests", "drivers")
DRIVER_BIN = os.path.join(DRIVERS, "bin")


def flang():
    for c in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", "flang"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return None


def build_drivers(verbose: bool = False) -> list:
    """Compile the Fortran binding module and the two host drivers the test suite runs
    (tests/drivers/*): a Fortran program using gemini_amd/fortran/gemhip_bind.f90 and a
    C++ program using include/gemini3d_hip.h, both linked against libgemhip.so; and the
    HDF5 writer's test program (h5io_check)."""
    os.makedirs(DRIVER_BIN, exist_ok=True)
    rpath = "-Wl,-rpath,$ORIGIN/../../../gemini_amd"
    link = ["-L" + HERE, "-lgemhip", rpath]
    built = []
    cmds = [["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
             os.path.join(DRIVERS, "capi_driver.cpp"), "-o", os.path.join(DRIVER_BIN, "capi_driver")]
            + link]
    # the HDF5 writer on its own, for the CPU tests of its chunked, threaded compression
    cmds.append(["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                 "-idirafter", hdf5_include(), os.path.join(DRIVERS, "h5io_check.cpp"),
                 os.path.join(CSRC, "h5io.cpp"), "-o", os.path.join(DRIVER_BIN, "h5io_check"),
                 "-L/opt/rocm/lib", "-lamdhip64", "-ldl", "-pthread", "-Wl,-rpath,/opt/rocm/lib"])
    fc = flang()
    if fc:
        mod = os.path.join(DRIVER_BIN, "gemhip_bind.o")
        cmds.append([fc, "-O2", "-c", os.path.join(HERE, "fortran", "gemhip_bind.f90"), "-o", mod,
                     "-module-dir", DRIVER_BIN])
        cmds.append([fc, "-O2", "-I" + DRIVER_BIN, os.path.join(DRIVERS, "fortran_driver.f90"), mod,
                     "-o", os.path.join(DRIVER_BIN, "fortran_driver")] + link)
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=DRIVER_BIN)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"building {cmd[-len(link) - 1]} failed")
        built.append(cmd[cmd.index("-o") + 1])
    return built


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_drivers(verbose=True))
