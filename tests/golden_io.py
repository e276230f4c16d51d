"""Load the golden fixtures written by tests/golden/make_golden.py."""
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["v4_2d_geo", "v4_3d_dam", "v4_2d_jit"]


def load(name):
    z = np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["params"] = json.loads(str(d["params"]))
    return d
