"""GPU parity: the HIP path against the oracle, bit for bit (FP64).

Every case runs the same seeded state through the oracle (CPU restatement of
the reference kernels) and through the C-ABI, and requires equal values in
every cell (ghosts included) -- the FP64 bar is bitwise, far inside the
1e-12 relative-L2 requirement.
"""
import sys

import numpy as np
import pytest

import oracle
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_DVAR, FST_GRADRHO, FST_P, FST_RHO, FST_T, FST_VAR, FST_VARX, FST_VEL
from fiesta_amd.runtime import Module
from helpers import assert_bitwise, cfg2d, cfg3d

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES_3D = {
    "rt_like": dict(),
    "all_walls": dict(per=(False, False, False),
                      bc=("outflow", "reflective", "hydrostatic", "noslip", "reflective", "outflow")),
    "periodic_all_noceq": dict(per=(True, True, True), ceq=False),
    "inviscid_ns1": dict(ns=1, visc=False, buoyancy=False),
    "ns3_ragged": dict(ns=3, nci=7, ncj=5, nck=9),
}

CASES_2D = {
    "weno_full": dict(),
    "centered4": dict(scheme="centered4"),
    "quick": dict(scheme="quick", ns=1),
    "walls": dict(per=(False, False), bc=("reflective", "noslip", "hydrostatic", "outflow", "outflow", "outflow")),
    "ns3_periodic": dict(ns=3, per=(True, True)),
}


def _pair(cfg, seed):
    var = W.random_state(cfg, seed)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    m = Module(cfg)
    m.upload(FST_VAR, var)
    return o, m


@pytest.mark.parametrize("name", sorted(CASES_3D))
def test_compute_3d(name):
    cfg = cfg3d(**CASES_3D[name])
    _say("case 3d compute", name)
    o, m = _pair(cfg, 11)
    o.pre_sim(); m.pre_sim()
    assert_bitwise(m.download(FST_VARX), o.get(FST_VARX), "varx after pre_sim", nan_ok=True)
    o.compute(); m.compute()
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after ghost fill")
    for a, n in ((FST_P, "p"), (FST_RHO, "rho"), (FST_T, "T"), (FST_VEL, "vel")):
        assert_bitwise(m.download(a), o.get(a), n)
    if cfg.ceq:
        assert_bitwise(m.download(FST_GRADRHO), o.get(FST_GRADRHO), "gradRho")
        ms, mch = m.maxima()
        oms, omch, _ = o.maxima()
        assert (ms, mch) == (oms, omch)
    assert_bitwise(m.download(FST_DVAR), o.get(FST_DVAR), "dvar")


@pytest.mark.parametrize("name", sorted(CASES_2D))
def test_compute_2d(name):
    cfg = cfg2d(**CASES_2D[name])
    _say("case 2d compute", name)
    o, m = _pair(cfg, 12)
    o.pre_sim(); m.pre_sim()
    assert_bitwise(m.download(FST_VARX), o.get(FST_VARX), "varx after pre_sim")
    o.compute(); m.compute()
    assert_bitwise(m.download(FST_DVAR), o.get(FST_DVAR), "dvar")


@pytest.mark.parametrize("name", sorted(CASES_3D))
def test_steps_3d_reference_sequence(name):
    cfg = cfg3d(**CASES_3D[name])
    _say("case 3d steps (reference sequence)", name)
    o, m = _pair(cfg, 21)
    o.pre_sim(); m.pre_sim()
    m.set_fused(False)
    o.step(3); m.step(3)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after 3 steps")


@pytest.mark.parametrize("name", sorted(CASES_2D))
def test_steps_2d(name):
    cfg = cfg2d(**CASES_2D[name])
    _say("case 2d steps", name)
    o, m = _pair(cfg, 22)
    o.pre_sim(); m.pre_sim()
    o.step(4); m.step(4)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after 4 steps")


def test_sod_shock_tube_200_steps():
    cfg = W.sod_config()
    _say("case sod 1000x4, 200 steps")
    var = W.sod_state(cfg)
    o = oracle.Oracle(cfg); o.set(FST_VAR, var); o.pre_sim()
    m = Module(cfg); m.upload(FST_VAR, var); m.pre_sim()
    o.step(200); m.step(200)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "sod var after 200 steps")


def test_rk_stage_split_matches_step():
    """compute()/rk_stage() driven from the host equals fst_step (rkAdvance)."""
    cfg = cfg3d()
    _say("case rkAdvance split vs fst_step")
    var = W.random_state(cfg, 5)
    a = Module(cfg); a.upload(FST_VAR, var); a.pre_sim()
    b = Module(cfg); b.upload(FST_VAR, var); b.pre_sim()
    a.compute(); a.rk_stage(1); a.compute(); a.rk_stage(2)
    b.set_fused(False); b.step(1)
    assert_bitwise(a.download(FST_VAR), b.download(FST_VAR), "split vs step")


FUSED_CASES = dict(CASES_3D)
FUSED_CASES["multi_tile"] = dict(nci=70, ncj=19, nck=37)
FUSED_CASES["multi_tile_walls"] = dict(nci=130, ncj=9, nck=20, per=(False, False, False),
                                       bc=("reflective", "outflow", "hydrostatic", "hydrostatic", "noslip", "outflow"))


@pytest.mark.parametrize("name", sorted(FUSED_CASES))
def test_steps_3d_fused(name):
    """The fused stage kernels equal the oracle bit for bit over several steps."""
    cfg = cfg3d(**FUSED_CASES[name])
    _say("case 3d steps (fused)", name)
    o, m = _pair(cfg, 31)
    o.pre_sim(); m.pre_sim()
    m.set_fused(True)
    o.step(2); m.step(2)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after 2 steps")
    o.step(1); m.step(1)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after 3 steps")
    # maxima of the new state equal a fresh reduction of the oracle's state
    m.compute(); o.compute()
    if cfg.ceq:
        assert m.maxima() == o.maxima()[:2]


def test_fused_interleaved_with_reference_sequence():
    """Switching between the fused path and compute()/rk_stage() keeps parity."""
    cfg = cfg3d(nci=20, ncj=12, nck=10)
    _say("case fused/reference interleave")
    o, m = _pair(cfg, 41)
    o.pre_sim(); m.pre_sim()
    m.step(1)
    m.compute(); m.rk_stage(1); m.compute(); m.rk_stage(2)
    m.step(1)
    o.step(3)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after mixed steps")


RK3_CASES = {
    "rt_like": dict(),
    "all_walls": dict(per=(False, False, False),
                      bc=("outflow", "reflective", "hydrostatic", "noslip", "reflective", "outflow")),
    "multi_tile": dict(nci=70, ncj=19, nck=37),
}


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("name", sorted(RK3_CASES))
def test_steps_3d_ssp_rk3(name, fused):
    """SSP-RK3 (rk_order = 3) on both kernel paths equals the oracle's SSP-RK3
    bit for bit, ghosts included; an odd number of steps per call checks the
    fused path's alternating primitive/maxima sets."""
    cfg = cfg3d(**RK3_CASES[name])
    cfg.rk_order = 3
    _say("case 3d ssp-rk3", name, "fused" if fused else "reference sequence")
    o, m = _pair(cfg, 51)
    o.pre_sim(); m.pre_sim()
    m.set_fused(fused)
    o.step(1); m.step(1)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after 1 step")
    o.step(2); m.step(2)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after 3 steps")


def test_steps_2d_ssp_rk3():
    cfg = cfg2d()
    cfg.rk_order = 3
    _say("case 2d ssp-rk3")
    o, m = _pair(cfg, 52)
    o.pre_sim(); m.pre_sim()
    o.step(4); m.step(4)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after 4 steps")


def test_ssp_rk3_differs_from_rk2():
    """The option changes the integrator (a guard against rk_order being ignored)."""
    cfg = cfg3d()
    _say("case rk2 vs ssp-rk3")
    var = W.random_state(cfg, 53)
    outs = []
    for rk in (2, 3):
        cfg.rk_order = rk
        m = Module(cfg)
        m.upload(FST_VAR, var)
        m.pre_sim()
        m.step(1)
        outs.append(m.download(FST_VAR))
        m.close()
    assert not np.array_equal(outs[0], outs[1])
