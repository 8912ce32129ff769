"""Pack the reference's own gold fixtures that pin the operator-apply path into
one compressed numpy archive, ``gold_fixtures.npz``, next to this script.

The reference (project-asgard/DG-SparseGrid) keeps MATLAB-generated fixtures
under ``gold/generated-inputs/`` as octave-style text (a ``# name/type/rows/
columns`` header followed by ``%.16g`` rows; gold/.. write_octave_like_output.m
:56-85).  This script only reads those data files; nothing of the reference is
executed.  Keys in the archive are the paths relative to
``gold/generated-inputs`` without the ``.dat`` suffix.

Run from the repository root:  python tests/golden/make_golden.py [REF_ROOT]
"""
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
GOLD = os.path.join(REF, "gold", "generated-inputs")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gold_fixtures.npz")

# which fixtures pin which part of the path (see the test files for the use)
PATTERNS = [
    r"element_table/(table|ids|child_ids)_[123]d_(SG|FG)$",
    r"element_table/1d_index_\d+_\d+$",
    r"connectivity/connect_1_\d+$",
    r"connectivity/connect_n_2_3_FG_\d+$",
    r"connectivity/connect_n_3_4_SG_\d+$",
    r"permutations/.*$",
    r"coefficients/continuity1_coefficients_(norotate_)?l2_d2_\d_\d$",
    r"coefficients/continuity2_coefficients_(norotate_)?l4_[45]_d3_\d_\d$",
    r"coefficients/continuity3_coefficients_(norotate_)?l2_3_2_d4_\d_\d$",
    r"coefficients/diffusion2_coefficients_(norotate_)?l(2_3|3_3)_d5_\d_\d$",
    r"transformations/operator_two_scale_\d_\d$",
    r"transformations/multiwavelet_\d_(h0|g0|h1|g1)$",
    r"time_advance/continuity1_sg_l2_d2_t\d$",
    r"time_advance/continuity2_sg_l(2_d2|4_d3)_t\d$",
    r"time_advance/continuity2_fg_l2_d2_t\d$",
    r"time_advance/continuity3_sg_l2_d2_t\d$",
    r"time_advance/continuity2_implicit_l2_d2_t\d$",
    r"time_advance/diffusion2_sg_l2_d2_t\d$",
    r"time_advance/diffusion2_implicit_sg_l3_d3_t\d$",
    r"time_advance/diffusion1_sg_l(2_d2|3_d3)_t\d$",
    r"time_advance/diffusion2_sg_l3_d3_t\d$",
    r"boundary_conditions/(vector|compute)_diffusion1_.*$",
    r"quadrature/lgwt_(roots|weights)_\d+_\w+$",
]

# Fixtures too large to keep whole, kept as regular samples (the stride is stored
# beside them under "<key>/stride"): the evaluation matrices of matrix_plot_D.m
# (gold_matrix_plot_D.m) as every 8th row, and the wavelet_to_realspace vectors
# (gold_wavelet_to_realspace.m) as every value, or every 8th of the 3D one.
SAMPLED = {
    **{f"transformations/matrix_plot_D/continuity_{k}/matrix_plot_D_{d}": 8 for k in (1, 2, 3) for d in range(k)},
    "transformations/wavelet_to_realspace_continuity_1": 1,
    "transformations/wavelet_to_realspace_continuity_2": 1,
    "transformations/wavelet_to_realspace_continuity_3": 8,
}


def read_octave_dat(path):
    """Parse one octave-like text fixture into a 2-D float64 array."""
    rows = cols = None
    kind = None
    data = []
    with open(path, "r") as fh:
        for line in fh:
            s = line.strip()
            if not s:
                continue
            if s.startswith("#"):
                m = re.match(r"#\s*(\w+):\s*(\S+)", s)
                if m:
                    key, val = m.group(1), m.group(2)
                    if key == "type":
                        kind = val
                    elif key == "rows":
                        rows = int(val)
                    elif key == "columns":
                        cols = int(val)
                continue
            data.extend(float(v) for v in re.split(r"[\s,]+", s) if v)
    arr = np.asarray(data, dtype=np.float64)
    if kind == "scalar":
        return arr.reshape(1, 1)
    if rows is None or cols is None:
        raise ValueError(f"{path}: no rows/columns header")
    if arr.size != rows * cols:
        raise ValueError(f"{path}: {arr.size} values for {rows}x{cols}")
    return arr.reshape(rows, cols)


def main():
    pats = [re.compile(p) for p in PATTERNS]
    out = {}
    for sub in sorted(os.listdir(GOLD)):
        d = os.path.join(GOLD, sub)
        if not os.path.isdir(d):
            continue
        for fn in sorted(os.listdir(d)):
            if not fn.endswith(".dat"):
                continue
            key = f"{sub}/{fn[:-4]}"
            if any(p.match(key) for p in pats):
                out[key] = read_octave_dat(os.path.join(d, fn))
    for key, stride in SAMPLED.items():
        a = read_octave_dat(os.path.join(GOLD, key + ".dat"))
        out[key] = a[::stride].copy()
        out[key + "/stride"] = np.asarray([stride], np.int64)
    np.savez_compressed(OUT, **out)
    print(f"wrote {len(out)} fixtures to {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
