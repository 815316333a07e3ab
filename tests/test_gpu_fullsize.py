"""Properties at the full 256^3 Rayleigh-Taylor workload (2 species, viscous,
C-equation, gravity; the benchmark's configuration), where the oracle would
take minutes:

- the fused stage kernels and the reference kernel sequence, two independent
  implementations, give the same bits after a step;
- a shift of the state by whole cells along the periodic x axis shifts the
  result by the same cells, bit for bit (no dependence on the tiling, the
  block order or the XCD mapping), in the FP64 and the learned FP32 forms;
- species masses change only by round-off over several steps (x/z periodic,
  reflecting y walls);
- the FP32 precision form stays within 1e-4 relative L2 of FP64 on density,
  pressure and mass fractions.
"""
import sys

import numpy as np
import pytest

from fiesta_amd import workloads as W
from fiesta_amd.config import FST_FP32_BULK, FST_VAR, Dims
from fiesta_amd.runtime import Module, Trainer
from helpers import assert_bitwise

pytestmark = pytest.mark.gpu

N = 256


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rt():
    cfg = W.rt3d_config(N)
    _say("case full size: generating the %d^3 Rayleigh-Taylor state" % N)
    return cfg, W.rt3d_state(cfg)


def _run(cfg, var, steps, fused=True, weights=None):
    d = Dims(cfg)
    m = Module(cfg)
    m.upload(FST_VAR, var[:d.nvt])
    if weights is not None:
        m.set_mlp_weights(weights)
    m.pre_sim()
    m.set_fused(fused)
    m.step(steps)
    out = m.download(FST_VAR)
    m.close()
    return out


def test_fullsize_fused_equals_reference_sequence(rt):
    cfg, var = rt
    _say("case full size: fused vs reference sequence, 1 step")
    a = _run(cfg, var, 1, fused=True)
    b = _run(cfg, var, 1, fused=False)
    assert np.array_equal(a, b)


def _roll_x(a, d, s):
    """Shift the interior columns by s cells along x (ghost columns are
    refilled by the periodic ghost fill of every stage)."""
    out = a.copy()
    g = d.ng
    out[..., g:g + d.nci] = np.roll(a[..., g:g + d.nci], s, axis=-1)
    return out


@pytest.mark.parametrize("form", ["fp64", "fp32", "fp64_learned", "fp32_learned"])
def test_fullsize_periodic_shift_invariance(rt, form):
    cfg, var = rt
    cfg = W.rt3d_config(N)
    weights = None
    if form.startswith("fp32"):
        cfg.precision = FST_FP32_BULK
    if form.endswith("learned"):
        cfg.learned_c = 1
        tr = Trainer(seed=5)
        weights = tr.weights()
        tr.close()
    d = Dims(cfg)
    _say("case full size: x shift invariance,", form)
    s = 37  # not a multiple of the 64-cell tile width
    base = _run(cfg, var, 2, weights=weights)
    shifted = _run(cfg, _roll_x(var[:d.nvt], d, s), 2, weights=weights)
    sl = d.interior()
    assert_bitwise(shifted[sl], _roll_x(base, d, s)[sl], form + " shifted run vs shifted result")


def _roll_z(a, d, s):
    out = a.copy()
    g = d.ng
    out[:, g:g + d.nck] = np.roll(a[:, g:g + d.nck], s, axis=1)
    return out


@pytest.mark.parametrize("form", ["fp64_rk3", "fp32_learned_rk3"])
def test_fullsize_z_shift_invariance_ssp_rk3(rt, form):
    """The same along periodic z (across the k-chunks of the 2.5D march), with
    the SSP-RK3 option (blend kernels, ghost-cell update)."""
    cfg, var = rt
    cfg = W.rt3d_config(N)
    cfg.rk_order = 3
    weights = None
    if form.startswith("fp32"):
        cfg.precision = FST_FP32_BULK
        cfg.learned_c = 1
        tr = Trainer(seed=5)
        weights = tr.weights()
        tr.close()
    d = Dims(cfg)
    _say("case full size: z shift invariance,", form)
    s = 13  # not a multiple of the k-chunk
    base = _run(cfg, var, 2, weights=weights)
    shifted = _run(cfg, _roll_z(var[:d.nvt], d, s), 2, weights=weights)
    sl = d.interior()
    assert_bitwise(shifted[sl], _roll_z(base, d, s)[sl], form + " shifted run vs shifted result")


def test_fullsize_species_mass_is_conserved(rt):
    cfg, var = rt
    d = Dims(cfg)
    _say("case full size: species mass over 4 steps")
    out = _run(cfg, var, 4)
    sl = d.interior()
    for s in range(cfg.ns):
        v = d.ndim + 1 + s
        m0 = var[v][sl[1:]].sum()
        m1 = out[v][sl[1:]].sum()
        assert abs(m1 - m0) <= 1e-12 * abs(m0), (s, (m1 - m0) / m0)


def test_fullsize_fp32_form_close_to_fp64(rt):
    cfg, var = rt
    d = Dims(cfg)
    _say("case full size: FP32 form vs FP64, 2 steps")
    a = _run(cfg, var, 2)
    c32 = W.rt3d_config(N)
    c32.precision = FST_FP32_BULK
    b = _run(c32, var, 2)
    sl = d.interior()
    ns = cfg.ns
    rho_a = sum(a[d.ndim + 1 + s][sl[1:]] for s in range(ns))
    rho_b = sum(b[d.ndim + 1 + s][sl[1:]] for s in range(ns))
    err = np.sqrt(((rho_b - rho_a) ** 2).sum() / (rho_a ** 2).sum())
    assert err < 1e-4, err
    for s in range(ns):
        ya, yb = a[d.ndim + 1 + s][sl[1:]] / rho_a, b[d.ndim + 1 + s][sl[1:]] / rho_b
        assert np.sqrt(((yb - ya) ** 2).sum() / (ya ** 2).sum()) < 1e-4


def test_fused_path_above_2_29_elements():
    """A grid whose nvt arrays hold more than 2^29 elements (1024 x 58 x 1024
    cells, 678,976,000 elements per array: beyond 32-bit byte offsets) runs
    on the fused stage kernels and equals the kernel sequence bit for bit."""
    from fiesta_amd.config import make_config
    cfg = make_config(ndim=3, ns=1, nci=1024, ncj=58, nck=1024, dx=1e-3, dy=1e-3, dz=1e-3, gamma=(1.4,),
                      M=(0.029,), mu=(1.8e-5,), visc=True, ceq=True, kap=10.0, eps=5.0, alpha=20.0, xper=True,
                      zper=True, bc=("outflow",) * 6, dt=1e-8, R=8.314462)
    d = Dims(cfg)
    assert d.N * d.nvt > 2 ** 29
    _say("fused above 2^29 elements: %d x %d" % (d.N, d.nvt))
    var = np.zeros(d.shape(d.nvt))
    k, j, i = np.meshgrid(np.arange(d.ngk), np.arange(d.ngj), np.arange(d.ngi), indexing="ij", sparse=True)
    rho = 1.0 + 0.1 * np.sin(0.05 * i) * np.cos(0.07 * j) * np.sin(0.03 * k)
    u = 10.0 * np.sin(0.02 * j) + 0 * i + 0 * k
    var[0] = rho * u
    var[3] = 1e5 / 0.4 + 0.5 * rho * u * u
    var[4] = rho
    var[5] = np.abs(np.sin(0.1 * i) + 0 * j + 0 * k)
    outs = []
    for fused in (True, False):
        m = Module(cfg)
        m.upload(FST_VAR, var)
        m.pre_sim()
        m.set_fused(fused)
        m.step(1)
        outs.append(m.download(FST_VAR)[d.interior()])
        m.close()
    assert_bitwise(outs[0], outs[1], "fused vs kernel sequence above 2^29 elements")
