"""One HIP runtime per process.

PyTorch-ROCm wheels ship their own HIP and HSA runtimes (torch/lib/
libamdhip64.so, SONAME libamdhip64.so.7) and load them by the unversioned
name.  libpfhip.so links the system runtime by SONAME.  If libpfhip.so is
loaded first, a later `import torch` does not match the loaded runtime by
name and maps a second HIP + HSA runtime into the process, which then finds
no GPU.  If torch's runtime is loaded first, libpfhip.so's dependency is
satisfied by it (same SONAME) and both share one runtime.

`preload()` therefore maps torch's runtime (without importing torch) before
libpfhip.so is loaded, whenever torch is installed.  PF_HIP_RUNTIME=system
skips this and uses the system runtime.
"""

import ctypes
import importlib.util
import os
import sys

_done = False


def torch_runtime_path():
    """Path of the HIP runtime bundled with the installed torch, or None."""
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return None
    if spec is None or not spec.submodule_search_locations:
        return None
    for loc in spec.submodule_search_locations:
        p = os.path.join(loc, "lib", "libamdhip64.so")
        if os.path.exists(p):
            return p
    return None


def preload():
    """Map torch's HIP runtime first when torch is installed (idempotent)."""
    global _done
    if _done:
        return
    _done = True
    if os.environ.get("PF_HIP_RUNTIME", "") == "system":
        return
    if "torch" in sys.modules:  # torch already mapped its runtime
        return
    p = torch_runtime_path()
    if p is not None:
        ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
