"""The C ABI library (CPU-side checks, no device needed): it loads, exports
every function include/*.h declares, and fails cleanly without a GPU."""

import ctypes
import glob
import os
import re

import pytest

from periodfind_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(pf_\w+)\s*\(", src))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("pf_ctx_create", "pf_ctx_destroy", "pf_last_error", "pf_ce_grid", "pf_aov_grid",
              "pf_ls_grid", "pf_fpw_grid", "pf_bls_grid", "pf_peaks"):
        assert n in names


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_signatures_cover_header():
    """The ctypes table binds every declared function."""
    assert set(declared_functions()) <= set(capi.SIGNATURES)


def test_no_device_error_path():
    if capi.device_count() > 0:
        pytest.skip("a device is visible")
    h = ctypes.c_void_p()
    rc = capi.lib().pf_ctx_create(ctypes.byref(h), 0)
    assert rc == capi.PF_E_NODEVICE
    assert b"no HIP device" in capi.lib().pf_last_error()
    with pytest.raises(capi.PfError):
        capi.Context(0)


def test_null_context_is_an_error_not_a_crash():
    import numpy as np

    off = np.array([0, 3], np.int64)
    t = np.zeros(3, np.float32)
    out = np.zeros(3, np.float32)
    P = np.ones(3, np.float32)
    D = np.zeros(1, np.float32)
    rc = capi.lib().pf_ls_grid(None, 1, off, t, t, P, 3, D, 1, out)
    assert rc == capi.PF_E_ARG


def test_one_hip_runtime_with_torch():
    """Loading the library and then torch maps a single HIP runtime: two
    runtimes in one process leave the second without GPUs."""
    import subprocess
    import sys

    pytest.importorskip("torch")
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from periodfind_amd import capi; capi.lib()\n"
        "import periodfind_amd.gpu\n"
        "import torch\n"
        "maps = open('/proc/self/maps').read().splitlines()\n"
        "libs = {l.split()[-1] for l in maps if 'libamdhip64' in l}\n"
        "print(len(libs))\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines()[-1] == "1", r.stdout
