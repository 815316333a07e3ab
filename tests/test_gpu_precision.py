"""Precision form (FP32 stencil/flux/viscous/C-equation arithmetic, FP64 EOS
and RK accumulation) against the FP64 oracle: relative L2 and Linf errors of
density, pressure and mass fractions after several steps.  Tolerance: 1e-4
relative L2 on density (the requirement), and the same bar for pressure and
mass fractions on these rough random states."""
import numpy as np
import pytest

import oracle
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_FP32_BULK, FST_VAR, Dims
from fiesta_amd.runtime import Module
from helpers import cfg3d

pytestmark = pytest.mark.gpu

RU = 8.314462


def fields(cfg, var):
    """Density, pressure and mass fractions of the interior (FP64, numpy)."""
    d = Dims(cfg)
    v = var[d.interior()]
    rs = v[4:4 + cfg.ns]
    rho = rs.sum(axis=0)
    Y = rs / rho
    cp = sum(Y[s] * cfg.gamma[s] * (RU / cfg.M[s]) / (cfg.gamma[s] - 1) for s in range(cfg.ns))
    cv = sum(Y[s] * (RU / cfg.M[s]) / (cfg.gamma[s] - 1) for s in range(cfg.ns))
    p = (cp / cv - 1) * (v[3] - 0.5 / rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    return rho, p, Y


def rel_errors(a, b):
    l2 = np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())
    linf = np.abs(a - b).max() / np.abs(b).max()
    return l2, linf


@pytest.mark.parametrize("name,kw", [("rt_like", dict()), ("walls_ns3", dict(ns=3, per=(False, False, False),
                                     bc=("outflow", "reflective", "hydrostatic", "noslip", "reflective", "outflow"))),
                                     ("rt_like_ssp_rk3", dict(rk=3)),
                                     # no C-equation: var holds no C planes, which the ring must not read
                                     ("rt_like_no_ceq", dict(ceq=False)),
                                     ("rt_like_no_ceq_no_visc", dict(ceq=False, visc=False))])
def test_fp32_form_error_against_fp64_oracle(name, kw):
    kw = dict(kw)
    rk = kw.pop("rk", 2)
    cfg = cfg3d(nci=40, ncj=21, nck=18, **kw)
    cfg.rk_order = rk
    print("case fp32 form", name, flush=True)
    var = W.random_state(cfg, 51)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.pre_sim()
    o.step(4)
    want = o.get(FST_VAR)
    cfg.precision = FST_FP32_BULK
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.pre_sim()
    m.set_fused(True)
    m.step(4)
    got = m.download(FST_VAR)
    assert np.isfinite(got).all()
    assert not np.array_equal(got, want)  # it really is a different arithmetic form
    for (a, b), label in zip(zip(fields(cfg, got), fields(cfg, want)), ("rho", "p", "Y")):
        l2, linf = rel_errors(a, b)
        print("  %s: rel L2 %.3e  rel Linf %.3e" % (label, l2, linf), flush=True)
        assert l2 < 1e-4, (label, l2)


@pytest.mark.parametrize("name,kw,learned,prec", [
    ("rt_like_plain_fp32", dict(), False, FST_FP32_BULK),
    ("rt_like_plain_fp64", dict(), False, 0),
    ("walls_ns3_plain_fp32", dict(ns=3, per=(False, False, False),
                                  bc=("outflow", "reflective", "hydrostatic", "noslip", "reflective", "outflow")),
     False, FST_FP32_BULK),
    ("walls_ns3_plain_fp64", dict(ns=3, per=(False, False, False),
                                  bc=("outflow", "reflective", "hydrostatic", "noslip", "reflective", "outflow")),
     False, 0),
    ("ragged_learned_fp32", dict(nci=70, ncj=19, nck=13), True, FST_FP32_BULK),
    ("ragged_learned_fp64", dict(nci=70, ncj=19, nck=13), True, 0)])
def test_ring_kernel_equals_global_kernel(name, kw, learned, prec, monkeypatch):
    """The stage kernels that stage their stencils in LDS (the advection
    plane, the viscous/C-equation ring) give the same bits as the kernels
    that read every stencil value from memory."""
    from fiesta_amd.runtime import Trainer
    base = dict(nci=40, ncj=21, nck=18)
    base.update(kw)
    cfg = cfg3d(**base)
    cfg.precision = prec
    if learned:
        cfg.learned_c = 1
    print("case ring vs global", name, flush=True)
    d = Dims(cfg)
    var = W.random_state(cfg, 77)
    w = Trainer(seed=2).weights() if learned else None
    outs = []
    for glob in (False, True):
        for var_name in ("FST_VCR_GLOBAL", "FST_ADV_GLOBAL"):
            if glob:
                monkeypatch.setenv(var_name, "1")
            else:
                monkeypatch.delenv(var_name, raising=False)
        m = Module(cfg)
        m.upload(FST_VAR, var[:d.nvt])
        if learned:
            m.set_mlp_weights(w)
        m.pre_sim()
        m.step(3)
        outs.append(m.download(FST_VAR))
        m.close()
    assert np.array_equal(outs[0], outs[1])
