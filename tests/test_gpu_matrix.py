"""Configuration matrix of the fused stages.

Cartesian 3D: every combination of species count (1-3), viscous terms,
C-equation and gravity, 2 RK2 steps on a ragged 21 x 11 x 9 grid (x and z
periodic, hydrostatic walls in y), in three forms against one oracle run:
FP64 bit for bit, the fast FP64 form within 1e-12 relative L2 on every
conserved variable, the FP32 form within 1e-4 relative L2 on density,
pressure and mass fractions (the requirement).  gen3d on a terrain grid:
species 1-3, plain / gen_ext / gen_ext with viscous terms, FP64 bit for bit
and FP32 within 1e-4.  The fused cart2d stage (FP64, bit for bit) on the
cartesian combinations.  The learned C-indicator form on species 1-3 with
and without viscous terms and gravity (FP64 bit for bit against the
oracle's learned steps, FP32 close to FP64), in 3D and on the fused cart2d
stage.  Combinations that no other
test runs are where a kernel can read a plane the configuration does not
have (the FP32 form without the C-equation once did).
"""
import itertools
import sys

import numpy as np
import pytest

import oracle
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_FP32_BULK, FST_FP64_FAST, FST_GRID, FST_VAR, Dims
from fiesta_amd.runtime import Module
from helpers import assert_bitwise, cfg3d

pytestmark = pytest.mark.gpu

RU = 8.314462


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(cfg, var):
    d = Dims(cfg)
    v = var[d.interior()]
    rs = v[4:4 + cfg.ns]
    rho = rs.sum(axis=0)
    Y = rs / rho
    cp = sum(Y[s] * cfg.gamma[s] * (RU / cfg.M[s]) / (cfg.gamma[s] - 1) for s in range(cfg.ns))
    cv = sum(Y[s] * (RU / cfg.M[s]) / (cfg.gamma[s] - 1) for s in range(cfg.ns))
    p = (cp / cv - 1) * (v[3] - 0.5 / rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    return rho, p, Y


def _rel_l2(a, b):
    den = np.sqrt((b ** 2).sum())
    return np.sqrt(((a - b) ** 2).sum()) / den if den > 0 else np.sqrt(((a - b) ** 2).sum())


def _run(cfg, var, nodes=None, steps=2):
    m = Module(cfg)
    if nodes is not None:
        m.upload(FST_GRID, nodes)
    m.upload(FST_VAR, var)
    m.pre_sim()
    m.step(steps)
    out = m.download(FST_VAR)
    m.close()
    return out


CART = list(itertools.product([1, 2, 3], [False, True], [False, True], [False, True]))


@pytest.mark.parametrize("ns,visc,ceq,buoy", CART)
def test_cart3d_forms(ns, visc, ceq, buoy):
    cfg = cfg3d(nci=21, ncj=11, nck=9, ns=ns, visc=visc, ceq=ceq, buoyancy=buoy)
    _say("case cart3d ns %d visc %d ceq %d buoyancy %d" % (ns, visc, ceq, buoy))
    var = W.random_state(cfg, 5 + ns)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.pre_sim()
    o.step(2)
    want = o.get(FST_VAR)
    _say("  fp64")
    assert_bitwise(_run(cfg, var), want, "fp64 after 2 steps")
    _say("  fp64_fast")
    cfg.precision = FST_FP64_FAST
    got = _run(cfg, var)
    d = Dims(cfg)
    gi, wi = got[d.interior()], want[d.interior()]
    for v in range(d.nvt):
        e = _rel_l2(gi[v], wi[v])
        assert e <= 1e-12, ("fast64 variable", v, e)
    _say("  fp32")
    cfg.precision = FST_FP32_BULK
    got = _run(cfg, var)
    assert np.isfinite(got).all()
    for a, b, label in zip(_fields(cfg, got), _fields(cfg, want), ("rho", "p", "Y")):
        assert _rel_l2(a, b) < 1e-4, (label, _rel_l2(a, b))


GEN = [(ns, mode) for ns in (1, 2, 3) for mode in ("plain", "ext", "ext_visc")]


@pytest.mark.parametrize("ns,mode", GEN)
def test_gen3d_forms(ns, mode):
    _say("case gen3d ns %d %s" % (ns, mode))
    from fiesta_amd.config import make_config
    kw = dict(ndim=3, ns=ns, nci=21, ncj=11, nck=9, grid=1, R=RU, gamma=(1.4, 1.67, 1.092)[:ns],
              M=(0.02896, 0.004, 0.14606)[:ns], mu=(1.8e-5, 2.1e-5, 1.5e-5)[:ns], dt=2e-7)
    if mode != "plain":
        kw.update(gen_ext=True, buoyancy=True, g=9.81, rhoRef=0.6, visc=mode == "ext_visc",
                  bc=("reflective", "outflow", "noslip", "outflow", "reflective", "outflow"))
    cfg = make_config(**kw)
    nodes = W.terrain_grid(cfg)
    var = W.random_state_all(cfg, seed=3 + ns, umag=20.0)
    o = oracle.Oracle(cfg)
    o.set(FST_GRID, nodes)
    o.set(FST_VAR, var)
    o.pre_sim()
    o.step(2)
    want = o.get(FST_VAR)
    _say("  fp64")
    assert_bitwise(_run(cfg, var, nodes), want, "gen3d fp64 after 2 steps")
    _say("  fp32")
    cfg.precision = FST_FP32_BULK
    got = _run(cfg, var, nodes)
    assert np.isfinite(got).all()
    for a, b, label in zip(_fields(cfg, got), _fields(cfg, want), ("rho", "p", "Y")):
        assert _rel_l2(a, b) < 1e-4, (label, _rel_l2(a, b))


@pytest.mark.parametrize("ns,visc,ceq,buoy", CART)
def test_cart2d_fused(ns, visc, ceq, buoy):
    """The fused cart2d stage (FP64) on the same combinations, 37 x 13 cells,
    x periodic and walls in y, 2 RK2 steps, bit for bit against the oracle."""
    from helpers import cfg2d
    cfg = cfg2d(nci=37, ncj=13, ns=ns, visc=visc, ceq=ceq, buoyancy=buoy, per=(True, False),
                bc=("outflow", "outflow", "reflective", "noslip", "outflow", "outflow"))
    _say("case cart2d ns %d visc %d ceq %d buoyancy %d" % (ns, visc, ceq, buoy))
    var = W.random_state(cfg, 7 + ns)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.pre_sim()
    o.step(2)
    want = o.get(FST_VAR)
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.pre_sim()
    m.step(2)
    got = m.download(FST_VAR)
    m.close()
    assert_bitwise(got, want, "cart2d fused after 2 steps")


LEARNED = list(itertools.product([1, 2, 3], [False, True], [False, True]))


@pytest.mark.parametrize("ns,visc,buoy", LEARNED)
def test_learned_forms(ns, visc, buoy):
    """The learned C-indicator form (untrained network) on species 1-3, with
    and without viscous terms and gravity: FP64 learned steps equal the
    oracle's learned steps given the GPU's C_hat and direction fields (2 RK2
    steps, interior), and the FP32 learned form stays within 1e-5 relative L2
    of the FP64 learned form in density."""
    from fiesta_amd.runtime import Trainer
    cfg = cfg3d(nci=21, ncj=11, nck=9, ns=ns, visc=visc, ceq=True, buoyancy=buoy)
    cfg.learned_c = 1
    _say("case learned ns %d visc %d buoyancy %d" % (ns, visc, buoy))
    var = W.random_state(cfg, 13 + ns)
    d = Dims(cfg)
    sl = d.interior()
    tr = Trainer(seed=11)
    w = tr.weights()
    tr.close()
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.set_mlp_weights(w)
    m.pre_sim()
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.pre_sim()
    for s in range(2):
        m.step(1)
        chat, dr, _, _ = m.learned_fields()
        _, _, mch = m.learned_stats()
        o.set_learned(chat, dr, mch)
        o.step(1)
        got, want = m.download(FST_VAR), o.get(FST_VAR)
        assert np.array_equal(got[sl], want[sl]), ("step", s, int((got[sl] != want[sl]).sum()))
    m.close()
    cfg.precision = FST_FP32_BULK
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.set_mlp_weights(w)
    m.pre_sim()
    m.step(2)
    g32 = m.download(FST_VAR)
    m.close()
    assert np.isfinite(g32).all()
    r32, r64 = g32[sl][4:4 + ns].sum(axis=0), got[sl][4:4 + ns].sum(axis=0)
    assert _rel_l2(r32, r64) < 1e-5, _rel_l2(r32, r64)


@pytest.mark.parametrize("ns,visc,buoy", LEARNED)
def test_learned2d_forms(ns, visc, buoy):
    """The 2D learned form (fused cart2d stage, untrained network) on species
    1-3 with and without viscous terms and gravity: 2 RK2 steps equal the
    oracle's learned steps given the GPU's fields (interior)."""
    from fiesta_amd.runtime import Trainer
    from helpers import cfg2d
    cfg = cfg2d(nci=37, ncj=13, ns=ns, visc=visc, ceq=True, buoyancy=buoy, per=(True, False),
                bc=("outflow", "outflow", "reflective", "noslip", "outflow", "outflow"))
    cfg.learned_c = 1
    _say("case learned 2D ns %d visc %d buoyancy %d" % (ns, visc, buoy))
    var = W.random_state(cfg, 17 + ns)
    d = Dims(cfg)
    sl = d.interior()
    tr = Trainer(seed=11, ndim=2)
    w = tr.weights()
    tr.close()
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.set_mlp_weights(w)
    m.pre_sim()
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.pre_sim()
    for s in range(2):
        m.step(1)
        chat, dr, _, _ = m.learned_fields()
        _, _, mch = m.learned_stats()
        o.set_learned(chat, dr, mch)
        o.step(1)
        got, want = m.download(FST_VAR), o.get(FST_VAR)
        assert np.array_equal(got[sl], want[sl]), ("step", s, int((got[sl] != want[sl]).sum()))
    m.close()
