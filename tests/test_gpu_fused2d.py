"""The fused cart2d stage (fst_fused2d.hip) against the oracle and against the
library's own reference-sequence kernels, bit for bit, on every cell:

- several tiles with partial edge tiles in both directions, RK2 and SSP-RK3,
  one to three species, with and without the viscous terms, buoyancy and the
  C-equation, periodic and physical boundaries;
- the configuration-1 Sod tube (1000 x 4, one species, SSP-RK3) against the
  oracle;
- switching between the fused path and the kernel sequence between steps
  (the maxima the fused stage carries from one stage to the next are rebuilt
  after the state changed);
- the learned 2D form, fused against the kernel sequence with the same
  network.
"""
import numpy as np
import pytest

import oracle
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_VAR, Dims
from fiesta_amd.runtime import Module, Trainer
from helpers import assert_bitwise, cfg2d

pytestmark = pytest.mark.gpu

CASES = {
    "tiles_rk2": dict(nci=150, ncj=21),
    "tiles_rk3": dict(nci=150, ncj=21, rk=3),
    "ns3_walls_rk3": dict(nci=70, ncj=17, ns=3, rk=3, per=(False, False),
                          bc=("reflective", "noslip", "hydrostatic", "outflow", "outflow", "outflow")),
    "ns1_no_visc": dict(nci=65, ncj=9, ns=1, visc=False, buoyancy=False, per=(True, False)),
    "no_ceq": dict(nci=100, ncj=12, ceq=False, rk=3, per=(True, True)),
    "one_tile": dict(nci=16, ncj=8),
}


def _say(*a):
    print("case:", *a, flush=True)


def _cfg(d):
    d = dict(d)
    rk = d.pop("rk", 2)
    cfg = cfg2d(**d)
    cfg.rk_order = rk
    return cfg


@pytest.mark.parametrize("name", sorted(CASES))
def test_fused2d_equals_kernel_sequence(name):
    cfg = _cfg(CASES[name])
    _say("fused 2D vs kernel sequence", name)
    var = W.random_state(cfg, 21)
    a = Module(cfg)
    b = Module(cfg)
    for m in (a, b):
        m.upload(FST_VAR, var)
        m.pre_sim()
    b.set_fused(False)
    a.step(5)
    b.step(5)
    assert_bitwise(a.download(FST_VAR), b.download(FST_VAR), "var after 5 steps (all cells)")
    if cfg.ceq:
        assert a.maxima() == b.maxima()
    a.close()
    b.close()


@pytest.mark.parametrize("name", ["tiles_rk3", "ns3_walls_rk3", "ns1_no_visc"])
def test_fused2d_equals_oracle(name):
    cfg = _cfg(CASES[name])
    _say("fused 2D vs oracle", name)
    var = W.random_state(cfg, 8)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.pre_sim()
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.pre_sim()
    o.step(3)
    m.step(3)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after 3 steps")
    m.close()


def test_fused2d_sod_tube_against_oracle():
    cfg = W.sod_config()
    cfg.rk_order = 3
    _say("fused 2D Sod 1000x4 SSP-RK3, 150 steps")
    var = W.sod_state(cfg)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.pre_sim()
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.pre_sim()
    o.step(150)
    m.step(150)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "sod var after 150 steps")
    m.close()


def test_fused2d_interleaved_with_kernel_sequence():
    cfg = _cfg(CASES["tiles_rk3"])
    _say("fused 2D interleaved with the kernel sequence")
    var = W.random_state(cfg, 4)
    a = Module(cfg)
    b = Module(cfg)
    for m in (a, b):
        m.upload(FST_VAR, var)
        m.pre_sim()
    b.set_fused(False)
    a.step(2)
    a.set_fused(False)
    a.step(1)
    a.set_fused(True)
    a.step(2)
    b.step(5)
    assert_bitwise(a.download(FST_VAR), b.download(FST_VAR), "var after 2 fused + 1 + 2 fused steps")
    # a new state uploaded between fused steps
    var2 = W.random_state(cfg, 5)
    for m in (a, b):
        m.upload(FST_VAR, var2)
        m.step(2)
    assert_bitwise(a.download(FST_VAR), b.download(FST_VAR), "var after an upload and 2 steps")
    a.close()
    b.close()


@pytest.mark.parametrize("rk", [2, 3])
def test_fused2d_learned_equals_kernel_sequence(rk):
    _say("fused 2D learned vs kernel sequence, rk", rk)
    n, nb = 96, 3
    cfg = W.multiblast2d_config(n, learned_c=True, nblobs=nb, cfl=0.3, rk=rk)
    var = W.multiblast2d_state(cfg, 4, nb)
    w = Trainer(seed=12, ndim=2).weights()
    a = Module(cfg)
    b = Module(cfg)
    for m in (a, b):
        m.upload(FST_VAR, var)
        m.set_mlp_weights(w)
        m.pre_sim()
    b.set_fused(False)
    a.step(4)
    b.step(4)
    assert_bitwise(a.download(FST_VAR), b.download(FST_VAR), "learned var after 4 steps")
    assert a.learned_stats() == b.learned_stats()
    a.close()
    b.close()
