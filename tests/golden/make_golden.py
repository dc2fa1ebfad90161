"""Writes tests/golden/golden.npz: small input/output vectors of the oracle.

Provenance: outputs of this repository's C++ restatement of periodfind's Rust
CPU backend (oracle/pf_oracle.cpp), not of the reference binary, which cannot
be built here.  They freeze the oracle so any change to it is visible, and the
GPU tests compare the HIP path against the same vectors.

    python -m tests.golden.make_golden
"""

import os

import numpy as np

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def inputs():
    rng = np.random.default_rng(424242)
    lens = [1, 5, 33, 120, 301]
    t, y, e = [], [], []
    for i, n in enumerate(lens):
        ti = np.sort(rng.uniform(0, 400, n)).astype(np.float32)
        yi = 15 + 0.3 * np.sin(2 * np.pi * ti / 1.7) if i % 2 == 0 else 15 + rng.normal(0, 0.1, n)
        t.append(ti)
        y.append(np.asarray(yi, np.float32) + rng.normal(0, 0.03, n).astype(np.float32))
        e.append(rng.uniform(0.02, 0.09, n).astype(np.float32))
    periods = np.geomspace(0.3, 12.0, 61).astype(np.float32)
    pdts = np.array([0.0, 2e-5], np.float32)
    return t, y, e, periods, pdts


def compute():
    t, y, e, P, D = inputs()
    out = {}
    off, tf = O.pack(t)
    out["in_offsets"] = off
    out["in_t"] = tf
    out["in_y"] = O.pack(y)[1]
    out["in_e"] = O.pack(e)[1]
    out["in_periods"] = P
    out["in_pdts"] = D
    out["ce"] = O.calc_grid("ce", t, y, P, D)
    out["ce_ext"] = O.calc_grid("ce", t, y, P, D, n_phase=7, n_mag=5, p_ext=2, m_ext=2)
    out["aov"] = O.calc_grid("aov", t, y, P, D, n_phase=15)
    out["ls"] = O.calc_grid("ls", t, y, P, D)
    out["fpw"] = O.calc_grid("fpw", t, y, P, D, errs=e)
    out["bls"] = O.calc_grid("bls", t, y, P, D, errs=e)
    for algo in ("ce", "bls"):
        idx, val = O.find_top_peaks_batched(out[algo], 8, 2, use_max=(algo != "ce"))
        out[algo + "_peak_idx"] = idx
        out[algo + "_peak_val"] = val
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **compute())
    print("wrote", os.path.join(HERE, "golden.npz"))
