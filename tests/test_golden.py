"""Golden vectors (tests/golden/): the committed fixtures are intact, and the
oracle still reproduces every one of them bit for bit.

The fixtures are the oracle's outputs on seeded inputs (tests/golden/make_golden.py):
a change to the oracle that moves any value, or a changed fixture file, fails
here.  The GPU side of the same fixtures is tests/test_gpu_golden.py.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from fiesta_amd.config import FST_DVAR, FST_GRID, FST_METRICS, FST_VAR, Dims
from golden_cases import CASE_NAMES, GOLDEN, load_case
from helpers import assert_bitwise


def test_manifest_lists_every_fixture_with_its_hash():
    man = json.load(open(os.path.join(GOLDEN, "manifest.json")))
    files = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))
    assert sorted(c["file"] for c in man["cases"].values()) == files
    for name, c in man["cases"].items():
        h = hashlib.sha256(open(os.path.join(GOLDEN, c["file"]), "rb").read()).hexdigest()
        assert h == c["sha256"], name


@pytest.mark.parametrize("name", CASE_NAMES)
def test_fixture_shapes_follow_the_layout(name):
    cfg, arrays, nsteps = load_case(name)
    d = Dims(cfg)
    assert arrays["var_in"].shape == d.shape(d.nvt)
    assert arrays["dvar"].shape == d.shape(d.nvt)
    assert arrays["var_steps"].shape == d.shape(d.nvt)
    if cfg.grid == 1:
        assert arrays["grid"].shape == d.grid_shape()
        assert arrays["metrics"].shape == d.metrics_shape()
    assert nsteps >= 3


@pytest.mark.parametrize("name", CASE_NAMES)
def test_oracle_reproduces_fixture(name):
    cfg, a, nsteps = load_case(name)
    o = oracle.Oracle(cfg)
    if cfg.grid == 1:
        o.set(FST_GRID, a["grid"])
    o.set(FST_VAR, a["var_in"])
    o.pre_sim()
    if cfg.grid == 1:
        assert_bitwise(o.get(FST_METRICS), a["metrics"], name + " metrics")
    o.compute()
    assert_bitwise(o.get(FST_DVAR), a["dvar"], name + " dvar")
    o = oracle.Oracle(cfg)
    if cfg.grid == 1:
        o.set(FST_GRID, a["grid"])
    o.set(FST_VAR, a["var_in"])
    o.pre_sim()
    o.step(nsteps)
    assert_bitwise(o.get(FST_VAR), a["var_steps"], name + " var after %d steps" % nsteps)


def test_fixture_dvar_is_not_trivial():
    """Every fixture's RHS moves every flow variable somewhere (a fixture of
    zeros would pin nothing)."""
    for name in CASE_NAMES:
        cfg, a, _ = load_case(name)
        d = Dims(cfg)
        for v in range(d.nv):
            assert np.abs(a["dvar"][v]).max() > 0, (name, v)
        assert not np.array_equal(a["var_steps"], a["var_in"]), name
