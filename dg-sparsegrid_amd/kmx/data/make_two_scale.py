"""Extract the multiwavelet two-scale filters H0, G0 (deg 1..8, 10) from the
reference's MATLAB v5 data files two_scale/two_scale_rel_<deg>.mat into
two_scale.npz beside this script.  scipy.io.loadmat reads plain numeric
arrays; nothing in the files is executed.

    python dg-sparsegrid_amd/kmx/data/make_two_scale.py [REF_ROOT]
"""
import os
import sys

import numpy as np
import scipy.io

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
out = {}
for deg in list(range(1, 9)) + [10]:
    m = scipy.io.loadmat(os.path.join(REF, "two_scale", f"two_scale_rel_{deg}.mat"))
    out[f"H0_{deg}"] = np.asarray(m["H0"], dtype=np.float64)
    out[f"G0_{deg}"] = np.asarray(m["G0"], dtype=np.float64)
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "two_scale.npz")
np.savez_compressed(dst, **out)
print("wrote", dst, os.path.getsize(dst), "bytes")
