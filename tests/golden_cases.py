"""Loader of the golden vectors under tests/golden/ (configuration + arrays)."""
import json
import os

import numpy as np

from fiesta_amd.config import make_config

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))
CASE_NAMES = sorted(_MANIFEST["cases"])


def load_case(name):
    """(config, arrays, steps) of one fixture."""
    c = _MANIFEST["cases"][name]
    with np.load(os.path.join(GOLDEN, c["file"])) as z:  # allow_pickle stays False
        arrays = {k: z[k] for k in z.files}
    kw = json.loads(str(arrays.pop("config")))
    for k in ("gamma", "M", "mu", "bc"):
        if k in kw:
            kw[k] = tuple(kw[k])
    return make_config(**kw), arrays, int(c["steps"])
