"""Build the HIP library (libpfhip.so) and the Cython binding in-tree.

hipcc cross-compiles for gfx950 without a GPU, so this runs anywhere the
ROCm toolchain is installed.  Usage: ``python -m periodfind_amd.build``.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpfhip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: no implicit multiply-add fusion, so f32/f64 arithmetic
# follows the reference's operation order; explicit fma() calls stay fused.
# Packed-FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", f"-I{INCLUDE}",
]

SOURCES = ["pf_grid.hip", "pf_prep.hip", "pf_peaks.hip", "pf_stats.hip", "pf_lsdiff.hip",
           "pf_noalloc.hip", "pf_runtime.cpp"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build_hip(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_cython(verbose=False):
    """Compile periodfind_amd/_pfhip.pyx into an in-tree extension module."""
    from Cython.Build import cythonize
    import numpy as np
    import sysconfig

    pyx = os.path.join(HERE, "_pfhip.pyx")
    c_file = cythonize([pyx], language_level=3, quiet=not verbose, force=True)
    del c_file
    csrc = os.path.join(HERE, "_pfhip.c")
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_pfhip" + suffix)
    cmd = [
        "gcc", "-O2", "-fPIC", "-shared", csrc, "-o", out + ".tmp",
        f"-I{sysconfig.get_paths()['include']}", f"-I{np.get_include()}", f"-I{INCLUDE}",
        f"-L{HERE}", "-lpfhip", "-Wl,-rpath,$ORIGIN",
        "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION",
    ]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed for the Cython binding:\n{r.stderr}")
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {out}")
    return out


def build(verbose=False):
    build_hip(verbose)
    if os.path.exists(os.path.join(HERE, "_pfhip.pyx")):
        build_cython(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
