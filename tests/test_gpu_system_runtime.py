"""The product path on the system HIP runtime alone.

By default the package maps the HIP runtime bundled with torch (when torch is
installed) before libpfhip.so, so that torch and the library share one
runtime (periodfind_amd/_hipruntime.py).  PF_HIP_RUNTIME=system skips that:
libpfhip.so then runs on the system ROCm runtime it was linked against.  This
test runs grids in a fresh process with that setting, checks them against the
oracle bit for bit, and checks from /proc/self/maps that the only HIP runtime
mapped is not torch's and that torch was never imported.
"""

import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
from periodfind_amd import capi, synth
from oracle import oracle as O

ctx = capi.Context(0)
t, m, e, P, D = synth.config2(1024, n_curves=4)
P = P[::5].copy()
checked = []
for algo, p in (("aov", dict(n_phase=15)), ("fpw", dict(n_bins=10)),
                ("bls", dict(n_bins=50, qmin=0.01, qmax=0.5))):
    errs = e if algo in ("fpw", "bls") else None
    got = ctx.grid(algo, t, m, P, D, errs=errs, **p)
    if algo == "aov":
        ref = O.aov_grid(t, m, P, D, **p)
    elif algo == "fpw":
        ref = O.fpw_grid(t, m, e, P, D, **p)
    else:
        ref = O.bls_grid(t, m, e, P, D, **p)
    assert np.array_equal(got, ref), algo
    checked.append(algo)
ctx.close()
maps = open("/proc/self/maps").read().splitlines()
hip = sorted({ln.split()[-1] for ln in maps if "libamdhip64" in ln})
print(json.dumps({"hip": hip, "torch": "torch" in sys.modules, "checked": checked}))
"""


def test_grids_on_system_runtime():
    env = dict(os.environ, PF_HIP_RUNTIME="system")
    r = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["checked"] == ["aov", "fpw", "bls"]
    assert not info["torch"]
    assert info["hip"], "no HIP runtime mapped"
    assert all("torch" not in p for p in info["hip"]), info["hip"]
