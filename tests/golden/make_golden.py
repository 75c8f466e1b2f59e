"""Writes the golden vectors of the 32x32 -> 64x64 FP64 case (sign +1,
lambda 632.8 nm) in the reference's .dat layout (raw little-endian float64,
cuda/util.h:248-337): inputs x0 amp/phase, positions u/v, and the oracle's
complex field y (interleaved re, im).  Run from the repository root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402
import oracle  # noqa: E402

wl = __graft_entry__.load_package()
from holo_amd import workloads  # noqa: E402


def main():
    st = workloads.plain_small(0).stages[0]
    y = oracle.transform_full(st.sign, st.lam, st.amp, st.phase, st.u, st.v)
    out = {
        "plain_small_x_amp.dat": st.amp,
        "plain_small_x_phase.dat": st.phase,
        "plain_small_u.dat": st.u.reshape(-1),
        "plain_small_v.dat": st.v.reshape(-1),
        "plain_small_y_oracle.dat": y.view(np.float64),
    }
    for name, arr in out.items():
        np.ascontiguousarray(arr, dtype="<f8").tofile(os.path.join(HERE, name))
    print("wrote", ", ".join(out))


if __name__ == "__main__":
    main()
