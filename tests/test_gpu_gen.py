"""GPU parity of the generalized-grid modules (gen2d/gen3d) against the oracle.

Metrics (preSim), one compute() (dvar), multi-step runs through the
reference sequence and postStep's secondary variables are compared bit for
bit on curvilinear grids: terrain-following nodes with a tanh-stretched
vertical axis, and a sheared/wavy 2D grid.  Inputs carry ghost values, since
the generalized modules never fill ghosts (src/gen3d.cpp:341-369,
src/gen2d.cpp:100-131).
"""
import sys

import numpy as np
import pytest

import oracle
from fiesta_amd import workloads as W
from fiesta_amd.config import (FST_DVAR, FST_FP32_BULK, FST_GRID, FST_METRICS, FST_P, FST_VAR, FST_VARX, FST_VEL,
                               make_config)
from fiesta_amd.runtime import FiestaError, Module
from helpers import assert_bitwise

pytestmark = pytest.mark.gpu

RU = 8.314462


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cfg(ndim, n, ns, scheme="weno5", dt=2e-7):
    return make_config(ndim=ndim, ns=ns, nci=n[0], ncj=n[1], nck=n[2], grid=1, R=RU,
                       gamma=(1.4, 1.67, 1.092, 1.3)[:ns], M=(0.02896, 0.004, 0.14606, 0.044)[:ns],
                       mu=(1.8e-5, 2.1e-5, 1.5e-5, 1.4e-5)[:ns], scheme=scheme, dt=dt)


def _wavy2d(cfg):
    d = W.Dims(cfg)
    return W.grid_nodes(cfg, lambda i, j, k: (i / d.nci + 0.03 * np.sin(2 * np.pi * j / d.ncj),
                                              j / d.ncj + 0.04 * np.sin(2 * np.pi * i / d.nci) + 0.1 * i / d.nci,
                                              None))


CASES = {
    "gen3d_terrain_ns2": dict(ndim=3, n=(20, 14, 11), ns=2, nodes="terrain"),
    "gen3d_terrain_ns3_ragged": dict(ndim=3, n=(67, 9, 13), ns=3, nodes="terrain"),
    # 4 species: outside the fused kernel's range, the per-variable flux kernels
    "gen3d_terrain_ns4_per_variable": dict(ndim=3, n=(18, 11, 9), ns=4, nodes="terrain"),
    "gen2d_terrain_weno": dict(ndim=2, n=(24, 18, 1), ns=2, nodes="terrain"),
    "gen2d_wavy_quick": dict(ndim=2, n=(19, 23, 1), ns=1, nodes="wavy", scheme="quick"),
    "gen2d_wavy_centered4": dict(ndim=2, n=(16, 16, 1), ns=3, nodes="wavy", scheme="centered4"),
}


def _setup(name, seed=5):
    c = CASES[name]
    cfg = _cfg(c["ndim"], c["n"], c["ns"], c.get("scheme", "weno5"))
    nodes = W.terrain_grid(cfg) if c["nodes"] == "terrain" else _wavy2d(cfg)
    var = W.random_state_all(cfg, seed=seed, umag=20.0)
    o = oracle.Oracle(cfg)
    o.set(FST_GRID, nodes)
    o.set(FST_VAR, var)
    m = Module(cfg)
    m.upload(FST_GRID, nodes)
    m.upload(FST_VAR, var)
    o.pre_sim()
    m.pre_sim()
    return cfg, o, m


@pytest.mark.parametrize("name", sorted(CASES))
def test_gen_metrics_and_compute(name):
    _say("case gen compute", name)
    cfg, o, m = _setup(name)
    assert_bitwise(m.download(FST_METRICS), o.get(FST_METRICS), name + " metrics")
    o.compute()
    m.compute()
    assert_bitwise(m.download(FST_VEL), o.get(FST_VEL), name + " contravariant velocity")
    assert_bitwise(m.download(FST_P), o.get(FST_P), name + " p")
    assert_bitwise(m.download(FST_DVAR), o.get(FST_DVAR), name + " dvar")
    m.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_gen_steps_and_post_step(name):
    _say("case gen steps", name)
    cfg, o, m = _setup(name, seed=8)
    o.step(3)
    m.step(3)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), name + " var after 3 steps")
    o.post_step()
    m.post_step()
    assert_bitwise(m.download(FST_VARX), o.get(FST_VARX), name + " varx")
    m.close()


def test_gen_rejects_fused_and_learned():
    """The fused path covers gen3d with ns <= 3 only: gen2d and 4 species
    are refused, as is the learned C-indicator on generalized grids."""
    for cfg in (_cfg(2, (8, 8, 1), 2), _cfg(3, (8, 8, 8), 4)):
        m = Module(cfg)
        with pytest.raises(FiestaError):
            m.set_fused(True)
        m.close()
    m = Module(_cfg(3, (8, 8, 8), 2))
    m.set_fused(True)
    m.set_fused(False)
    m.close()
    bad = _cfg(3, (8, 8, 8), 2)
    bad.learned_c = 1
    bad.ceq = 1
    with pytest.raises(FiestaError):
        Module(bad)


GEN_EXT = {
    "gen3d_terrain_ns2_rk2": dict(ndim=3, n=(20, 14, 11), ns=2, rk=2),
    "gen3d_terrain_ns3_rk3": dict(ndim=3, n=(67, 9, 13), ns=3, rk=3),
    "gen2d_terrain_ns2_rk3": dict(ndim=2, n=(24, 18, 1), ns=2, rk=3),
}
EXT_BC = ("reflective", "outflow", "noslip", "outflow", "reflective", "outflow")


def _ext_cfg(c, **kw):
    cfg = _cfg(c["ndim"], c["n"], c["ns"])
    cfg.rk_order = c["rk"]
    for name, v in zip(("bcL", "bcR", "bcB", "bcT", "bcH", "bcF"), EXT_BC):
        setattr(cfg, name, {"outflow": 0, "reflective": 1, "noslip": 2}[v])
    cfg.gen_ext = 1
    cfg.buoyancy = 1
    cfg.rhoRef = 0.6
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


@pytest.mark.parametrize("name", sorted(GEN_EXT))
def test_gen_ext_terrain_against_oracle(name):
    """gen_ext: the cartesian ghost fill and buoyancy on a terrain grid, over
    several steps of the reference sequence, bit for bit on every cell."""
    _say("case gen_ext steps", name)
    cfg = _ext_cfg(GEN_EXT[name])
    nodes = W.terrain_grid(cfg)
    var = W.random_state_all(cfg, seed=31, umag=20.0)
    o = oracle.Oracle(cfg)
    m = Module(cfg)
    for x in (o, m):
        (x.set if x is o else x.upload)(FST_GRID, nodes)
        (x.set if x is o else x.upload)(FST_VAR, var)
        x.pre_sim()
    o.compute()
    m.compute()
    assert_bitwise(m.download(FST_DVAR), o.get(FST_DVAR), name + " dvar")
    o.step(4)
    m.step(4)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), name + " var after 4 steps")
    m.close()


@pytest.mark.parametrize("ndim", [2, 3])
def test_gen_ext_uniform_grid_equals_cartesian_module(ndim):
    """On a uniform grid of spacing 2^-5 the generalized module with gen_ext
    and the cartesian module give the same state bit for bit."""
    _say("case gen_ext uniform vs cart", ndim)
    h = 2.0 ** -5
    c = dict(ndim=ndim, n=(26, 15, 9 if ndim == 3 else 1), ns=2, rk=3)
    gen = _ext_cfg(c)
    cart = _ext_cfg(c, grid=0, gen_ext=0, dx=h, dy=h, dz=h)
    nodes = W.grid_nodes(gen, lambda i, j, k: (h * i, h * j, h * k if ndim == 3 else None))
    var = W.random_state_all(gen, seed=6, umag=20.0)
    a = Module(gen)
    a.upload(FST_GRID, nodes)
    b = Module(cart)
    for m in (a, b):
        m.upload(FST_VAR, var)
        m.pre_sim()
    b.set_fused(False)
    a.step(5)
    b.step(5)
    assert_bitwise(a.download(FST_VAR), b.download(FST_VAR), "gen_ext vs cart after 5 steps")
    a.close()
    b.close()


def test_terrain_expansion_against_oracle():
    """The terrain expansion workload (terrain_expansion_config: gen3d +
    gen_ext, 3 species, gravity, viscous, no-slip ground and lid, SSP-RK3) at
    40 x 20 x 40 cells: 12 steps of the blast on the GPU equal the oracle's
    bit for bit on every cell."""
    _say("case terrain expansion 40x20x40 vs oracle, 12 steps")
    cfg = W.terrain_expansion_config((40, 20, 40))
    nodes = W.terrain_grid(cfg, vertical=1)
    var, _ = W.terrain_expansion_state(cfg, nodes)
    cfg.dt = W.gen_cfl_dt(cfg, var, nodes, 0.4)
    o = oracle.Oracle(cfg)
    o.set(FST_GRID, nodes)
    o.set(FST_VAR, var)
    o.pre_sim()
    m = Module(cfg)
    m.upload(FST_GRID, nodes)
    m.upload(FST_VAR, var)
    m.pre_sim()
    o.step(12)
    m.step(12)
    assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "terrain expansion var after 12 steps")
    m.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_terrain_expansion_keeps_species_mass(precision):
    """The same workload at 128 x 64 x 128 cells, 60 steps on the GPU: every
    species' interior total stays what it was to 1e-13 relative in FP64, and
    to 1e-11 in the FP32 form (its flux differences are rounded to binary32,
    so they telescope only to that rounding)."""
    _say("case terrain expansion 128x64x128 mass, 60 steps, precision", precision)
    cfg = W.terrain_expansion_config((128, 64, 128))
    nodes = W.terrain_grid(cfg, vertical=1)
    var, _ = W.terrain_expansion_state(cfg, nodes)
    cfg.dt = W.gen_cfl_dt(cfg, var, nodes, 0.4)
    cfg.precision = precision
    d = W.Dims(cfg)
    sl = d.interior()
    m = Module(cfg)
    m.upload(FST_GRID, nodes)
    m.upload(FST_VAR, var)
    m.pre_sim()
    m0 = m.download(FST_VAR)[sl][4:].sum(axis=(1, 2, 3))
    m.step(60)
    v = m.download(FST_VAR)
    m.close()
    assert np.isfinite(v).all()
    np.testing.assert_allclose(v[sl][4:].sum(axis=(1, 2, 3)), m0, rtol=1e-13 if precision == 0 else 1e-11, atol=0)


FUSED_GEN = {
    "terrain_ns2_rk2": (dict(ndim=3, n=(20, 14, 11), ns=2, rk=2), False),
    "terrain_ns3_ragged_rk3": (dict(ndim=3, n=(67, 9, 13), ns=3, rk=3), False),
    "terrain_ns1_ragged_rk2_ext": (dict(ndim=3, n=(70, 15, 10), ns=1, rk=2), True),
    "terrain_ns3_rk3_ext": (dict(ndim=3, n=(20, 14, 11), ns=3, rk=3), True),
    # gen_ext with the viscous terms in metric form
    "terrain_ns2_rk2_ext_visc": (dict(ndim=3, n=(20, 14, 11), ns=2, rk=2, visc=True), True),
    "terrain_ns3_ragged_rk3_ext_visc": (dict(ndim=3, n=(67, 9, 13), ns=3, rk=3, visc=True), True),
}


@pytest.mark.parametrize("name", sorted(FUSED_GEN))
def test_gen3d_fused_steps_equal_sequence_and_oracle(name):
    """gen3d on the fused kernels (one EOS / contravariant-velocity pass and
    one advection + stage-update launch per stage, fst_step's default) and on
    the reference kernel sequence (fst_set_fused(0)): both equal the oracle
    bit for bit on every cell, ghosts included, after 4 steps; with and
    without gen_ext (ghost fill + buoyancy, and the viscous terms in metric
    form), RK2 and SSP-RK3.  The fused run calls postStep between its steps:
    postStep's EOS pass writes the cartesian velocity over the buffer that
    holds the fused stage's contravariant velocity, which the next step must
    not reuse."""
    _say("case gen3d fused vs sequence vs oracle", name)
    c, ext = FUSED_GEN[name]
    if ext:
        cfg = _ext_cfg(c, visc=int(c.get("visc", False)))
    else:
        cfg = _cfg(c["ndim"], c["n"], c["ns"])
        cfg.rk_order = c["rk"]
    nodes = W.terrain_grid(cfg)
    var = W.random_state_all(cfg, seed=17, umag=20.0)
    o = oracle.Oracle(cfg)
    o.set(FST_GRID, nodes)
    o.set(FST_VAR, var)
    o.pre_sim()
    o.step(4)
    want = o.get(FST_VAR)
    for fused in (True, False):
        m = Module(cfg)
        m.upload(FST_GRID, nodes)
        m.upload(FST_VAR, var)
        m.pre_sim()
        m.set_fused(fused)
        if fused:
            m.step(2)
            m.post_step()
            m.step(2)
        else:
            m.step(4)
        assert_bitwise(m.download(FST_VAR), want, "%s %s var after 4 steps" % (name, "fused" if fused else "sequence"))
        m.close()


def _fields(cfg, var):
    """Density, pressure and mass fractions of the interior (FP64, numpy)."""
    d = W.Dims(cfg)
    v = var[d.interior()]
    rs = v[4:4 + cfg.ns]
    rho = rs.sum(axis=0)
    Y = rs / rho
    cp = sum(Y[s] * cfg.gamma[s] * (RU / cfg.M[s]) / (cfg.gamma[s] - 1) for s in range(cfg.ns))
    cv = sum(Y[s] * (RU / cfg.M[s]) / (cfg.gamma[s] - 1) for s in range(cfg.ns))
    p = (cp / cv - 1) * (v[3] - 0.5 / rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    return rho, p, Y


@pytest.mark.parametrize("name", sorted(FUSED_GEN))
def test_gen3d_fp32_form_against_fp64_oracle(name):
    """The FP32 form of the fused gen3d stage (binary32 fluxes, float
    primitive planes; FP64 metric pressure gradient, gen_ext viscous terms,
    EOS and RK accumulation) after 4 steps on a terrain grid: relative L2
    error of density, pressure and mass fractions against the FP64 oracle
    below 1e-5 (the requirement is 1e-4), and not bit-equal to it."""
    _say("case gen3d fp32 form vs FP64 oracle", name)
    c, ext = FUSED_GEN[name]
    if ext:
        cfg = _ext_cfg(c, visc=int(c.get("visc", False)))
    else:
        cfg = _cfg(c["ndim"], c["n"], c["ns"])
        cfg.rk_order = c["rk"]
    nodes = W.terrain_grid(cfg)
    var = W.random_state_all(cfg, seed=17, umag=20.0)
    o = oracle.Oracle(cfg)
    o.set(FST_GRID, nodes)
    o.set(FST_VAR, var)
    o.pre_sim()
    o.step(4)
    want = o.get(FST_VAR)
    cfg.precision = FST_FP32_BULK
    m = Module(cfg)
    m.upload(FST_GRID, nodes)
    m.upload(FST_VAR, var)
    m.pre_sim()
    m.step(2)
    m.post_step()
    m.step(2)
    got = m.download(FST_VAR)
    m.close()
    assert np.isfinite(got).all()
    assert not np.array_equal(got, want)
    for a, b, label in zip(_fields(cfg, got), _fields(cfg, want), ("rho", "p", "Y")):
        l2 = np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())
        linf = np.abs(a - b).max() / np.abs(b).max()
        print("  %s: rel L2 %.3e  rel Linf %.3e" % (label, l2, linf), flush=True)
        assert l2 < 1e-5, (label, l2)


def test_gen3d_fp32_form_runs_through_fst_step_only():
    """An FP32 gen3d context refuses the single-RHS call and the reference
    sequence, which compute in FP64 whatever the precision field says."""
    cfg = _cfg(3, (12, 10, 9), 2)
    cfg.precision = FST_FP32_BULK
    m = Module(cfg)
    with pytest.raises(FiestaError, match="fst_step only"):
        m.compute()
    with pytest.raises(FiestaError, match="fst_step only"):
        m.set_fused(False)
    m.set_fused(True)
    m.close()


@pytest.mark.parametrize("rk", [2, 3])
def test_gen3d_fp32_equals_cartesian_fp32_on_uniform_power_of_two_grid(rk):
    """The FP32 form of the fused gen3d stage with gen_ext on a uniform grid of
    spacing 2^-5 follows the cartesian FP32 form with the same walls and
    gravity bit for bit over 3 steps, ghosts included: the metrics are exact
    powers of two (in FP64 and as floats), so every contravariant velocity,
    binary32 face velocity and flux is the cartesian one scaled by a power of
    two, and the metric pressure gradient equals the cartesian one as in FP64.
    This ties the generalized FP32 kernels (k_adv3d_sgen, k_rk_eos_gen<true>)
    to the cartesian FP32 kernels, which are held to the FP64 oracle."""
    _say("case gen3d fp32 vs cart3d fp32 on a uniform 2^-5 grid, rk", rk)
    h = 2.0 ** -5
    n = (70, 15, 10)
    bc = ("reflective", "outflow", "noslip", "outflow", "reflective", "outflow")
    kw = dict(ndim=3, ns=2, nci=n[0], ncj=n[1], nck=n[2], R=RU, gamma=(1.4, 1.67), M=(0.02896, 0.004),
              mu=(0.0, 0.0), dt=2e-7, buoyancy=True, g=9.81, rhoRef=0.5, bc=bc, zper=True, rk=rk,
              precision=FST_FP32_BULK)
    cart = make_config(dx=h, dy=h, dz=h, **kw)
    gen = make_config(grid=1, gen_ext=True, **kw)
    var = W.random_state(cart, seed=23)
    nodes = W.grid_nodes(gen, lambda i, j, k: (h * i, h * j, h * k))
    out = []
    for cfg, grid in ((cart, None), (gen, nodes)):
        m = Module(cfg)
        if grid is not None:
            m.upload(FST_GRID, grid)
        m.upload(FST_VAR, var)
        m.pre_sim()
        m.step(3)
        out.append(m.download(FST_VAR))
        m.close()
    assert np.isfinite(out[0]).all()
    assert_bitwise(out[1], out[0], "gen3d fp32 vs cart3d fp32 after 3 steps")
    # and it is the FP32 arithmetic, not the FP64 one
    gen.precision = 0
    m = Module(gen)
    m.upload(FST_GRID, nodes)
    m.upload(FST_VAR, var)
    m.pre_sim()
    m.step(3)
    assert not np.array_equal(m.download(FST_VAR), out[0])
    m.close()
