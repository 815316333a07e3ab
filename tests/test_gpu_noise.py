"""GPU parity of postStep's noise filter (src/noise.hpp) with the oracle, bit
for bit: detection on the coarse lattice, the C_hat cut-off, and the smoothing
sweeps, in 3D and 2D, in both modes, and in gen2d's variant; also interleaved
with time steps on both 3D kernel paths."""
import sys

import numpy as np
import pytest

import oracle
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_VAR, FST_VARX, make_config
from fiesta_amd.runtime import Module
from helpers import assert_bitwise, cfg2d, cfg3d

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = {
    # dh chosen so that part of the blocks flag on the random states
    "3d_mode1_ceq_coff": (cfg3d, dict(nci=21, ncj=14, nck=11), dict(n_dh=2e-3, n_eta=0.05, n_coff=0.3, n_nt=2)),
    "3d_mode2_noceq": (cfg3d, dict(nci=16, ncj=9, nck=12, ceq=False), dict(n_dh=1e-2, n_eta=0.02, n_mode=2)),
    "3d_odd_sizes": (cfg3d, dict(nci=7, ncj=5, nck=9), dict(n_dh=0.0, n_eta=0.05)),
    "2d_mode1_ceq_coff": (cfg2d, dict(nci=33, ncj=20), dict(n_dh=1e-1, n_eta=0.05, n_coff=0.4, n_nt=3)),
    "2d_mode2": (cfg2d, dict(nci=16, ncj=11), dict(n_dh=5.0, n_eta=0.05, n_mode=2)),
}


def _noise_cfg(name):
    f, kw, nz = CASES[name]
    cfg = f(**kw)
    cfg.noise = 1
    cfg.n_dh, cfg.n_eta = nz["n_dh"], nz["n_eta"]
    cfg.n_coff = nz.get("n_coff", 0.0)
    cfg.n_nt = nz.get("n_nt", 1)
    cfg.n_mode = nz.get("n_mode", 1)
    return cfg


@pytest.mark.parametrize("name", sorted(CASES))
def test_post_step_noise_filter(name):
    cfg = _noise_cfg(name)
    _say("case noise", name)
    var = W.random_state(cfg, 61)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    m = Module(cfg)
    m.upload(FST_VAR, var)
    o.pre_sim(); m.pre_sim()
    o.post_step(); m.post_step()
    got, want = m.download(FST_VAR), o.get(FST_VAR)
    assert not np.array_equal(want, var), "the filter should act on this state"
    assert_bitwise(got, want, name + " var after postStep")
    assert_bitwise(m.download(FST_VARX), o.get(FST_VARX), name + " varx", nan_ok=True)  # 3D ghosts hold rho = 0
    m.close()


@pytest.mark.parametrize("fused", [True, False])
def test_steps_with_noise_filter_3d(fused):
    """Fiesta::step = preStep + rkAdvance + postStep, three times."""
    cfg = _noise_cfg("3d_mode1_ceq_coff")
    _say("case noise steps 3d", "fused" if fused else "reference sequence")
    var = W.random_state(cfg, 62)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    m = Module(cfg)
    m.upload(FST_VAR, var)
    o.pre_sim(); m.pre_sim()
    m.set_fused(fused)
    for s in range(3):
        o.step(1); m.step(1)
        o.post_step(); m.post_step()
        assert_bitwise(m.download(FST_VAR), o.get(FST_VAR), "var after step %d" % (s + 1))
    m.close()


def test_gen2d_noise_filter():
    cfg = make_config(ndim=2, ns=2, nci=18, ncj=13, nck=1, grid=1, gamma=(1.4, 1.67), M=(0.02896, 0.004),
                      mu=(0.0, 0.0), noise=True, n_dh=1.0, n_eta=0.05, n_mode=2, dt=2e-7)
    _say("case noise gen2d")
    var = W.random_state_all(cfg, 63, umag=20.0)
    nodes = W.terrain_grid(cfg)
    from fiesta_amd.config import FST_GRID
    o = oracle.Oracle(cfg)
    o.set(FST_GRID, nodes)
    o.set(FST_VAR, var)
    m = Module(cfg)
    m.upload(FST_GRID, nodes)
    m.upload(FST_VAR, var)
    o.pre_sim(); m.pre_sim()
    o.post_step(); m.post_step()
    want = o.get(FST_VAR)
    assert not np.array_equal(want, var)
    assert_bitwise(m.download(FST_VAR), want, "gen2d var after postStep")
    m.close()
