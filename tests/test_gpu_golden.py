"""GPU parity against the committed golden vectors (tests/golden/), bit for bit.

Each fixture holds a seeded input and the oracle's outputs.  The HIP path runs
the same input through the C-ABI and must reproduce every value (ghosts
included): the metrics and one compute()'s dvar, and the state after several
rkAdvance steps through the reference kernel sequence and, for the cartesian
3D module, through the fused stage kernels.
"""
import sys

import pytest

from fiesta_amd.config import FST_DVAR, FST_GRID, FST_METRICS, FST_VAR
from fiesta_amd.runtime import Module
from golden_cases import CASE_NAMES, load_case
from helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def _module(cfg, a):
    m = Module(cfg)
    if cfg.grid == 1:
        m.upload(FST_GRID, a["grid"])
    m.upload(FST_VAR, a["var_in"])
    m.pre_sim()
    return m


@pytest.mark.parametrize("name", CASE_NAMES)
def test_hip_path_reproduces_fixture(name):
    cfg, a, nsteps = load_case(name)
    print("case golden", name, flush=True)
    sys.stdout.flush()
    m = _module(cfg, a)
    if cfg.grid == 1:
        assert_bitwise(m.download(FST_METRICS), a["metrics"], name + " metrics")
    m.compute()
    assert_bitwise(m.download(FST_DVAR), a["dvar"], name + " dvar")
    m.close()
    paths = [False, True] if (cfg.ndim == 3 and cfg.grid == 0) else [None]
    for fused in paths:
        m = _module(cfg, a)
        if fused is not None:
            m.set_fused(fused)
        m.step(nsteps)
        label = {None: "", False: " (reference sequence)", True: " (fused)"}[fused]
        assert_bitwise(m.download(FST_VAR), a["var_steps"], name + " var after %d steps%s" % (nsteps, label))
        m.close()
