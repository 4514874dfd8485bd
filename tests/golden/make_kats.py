"""Writes reference_kats.json: the reference's own LU known-answer tests as data.

Transcribed from the reference's test file (values and expected status only):
  Testing/Cxx/TestLUDecomposition.cxx:283-299  zero-pivot KATs A0, A1, Z and their info
  Testing/Cxx/TestLUDecomposition.cxx:310-315  1e-8 * I (info 0)
  Testing/Cxx/TestLUDecomposition.cxx:329-341  subnormal pivot [[t, 1], [t/2, 1]],
                                               t = 2^(min_exponent - 4) (info 0)
"""
import json
import os

import numpy as np

KATS = {
    "source": "Testing/Cxx/TestLUDecomposition.cxx:276-342",
    "zero_pivot": [
        {"name": "A0", "a": [[0, 1, 2], [0, 3, 4], [0, 5, 7]], "info": 1},
        {"name": "A1", "a": [[1, 2, 3], [2, 4, 5], [4, 8, 1]], "info": 2},
        {"name": "Z", "a": [[0, 0], [0, 0]], "info": 1},
    ],
    "small_scale": {"a": [[1e-8, 0.0], [0.0, 1e-8]], "info": 0},
    # numpy's minexp is C's DBL_MIN_EXP - 1, so min_exponent - 4 == minexp - 3
    "subnormal": {"exp2_t": int(np.finfo(np.float64).minexp) - 3, "info": 0},
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
    with open(out, "w") as f:
        json.dump(KATS, f, indent=1)
    print(out)
