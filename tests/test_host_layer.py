"""The C++ host rk_func layer (fiesta_amd/host) and its fiesta_run driver.

CPU: the driver is built and rejects bad arguments before touching a device.
GPU: time steps through the C++ interface -- the module sequence preStep /
rkAdvance / postStep (src/fiesta.cpp:333-340, src/rk.cpp:28-72) and the fused
fst_step -- equal the oracle bit for bit in 2D and 3D.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from fiesta_amd import build as B
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_VAR, Dims
from helpers import cfg2d, cfg3d

RUN = B.RUN_BIN


def test_driver_built_and_checks_arguments(tmp_path):
    assert os.path.exists(RUN), "fiesta_run not built"
    r = subprocess.run([RUN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    cfgf = tmp_path / "c.bin"
    cfgf.write_bytes(b"\0" * 7)  # not an fst_config
    r = subprocess.run([RUN, str(cfgf), str(cfgf), str(tmp_path / "o.bin"), "1"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "fst_config" in r.stderr


def _run_driver(tmp_path, cfg, var, nsteps, mode):
    cfgf, stf, outf = tmp_path / "cfg.bin", tmp_path / "state.bin", tmp_path / "out.bin"
    cfgf.write_bytes(bytes(cfg))
    stf.write_bytes(np.ascontiguousarray(var, np.float64).tobytes())
    r = subprocess.run([RUN, str(cfgf), str(stf), str(outf), str(nsteps), mode], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["steps"] == nsteps and info["mode"] == mode
    return np.frombuffer(outf.read_bytes(), np.float64).reshape(var.shape)


@pytest.mark.gpu
@pytest.mark.parametrize("ndim,mode,rk,noise", [(3, "rk", 2, 0), (3, "fused", 2, 0), (2, "rk", 2, 0), (3, "rk", 3, 0),
                                               (3, "fused", 3, 0), (3, "rk", 2, 1), (3, "fused", 2, 1)])
def test_driver_steps_equal_oracle(tmp_path, ndim, mode, rk, noise):
    """Fiesta::step (src/fiesta.cpp:333-340) runs postStep after every step;
    with the noise filter on it changes var, so both modes must run it per step."""
    import oracle
    print("case: fiesta_run %dD %s rk%d noise %d" % (ndim, mode, rk, noise), flush=True)
    cfg = cfg3d(nci=24, ncj=13, nck=10) if ndim == 3 else cfg2d(nci=40, ncj=12)
    cfg.rk_order = rk
    if noise:
        cfg.noise, cfg.n_dh, cfg.n_eta, cfg.n_coff, cfg.n_nt = 1, 2e-3, 0.05, 0.3, 2
    var = W.random_state(cfg, 5)
    got = _run_driver(tmp_path, cfg, var, 3, mode)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.pre_sim()
    for _ in range(3):
        o.pre_step()
        for s in range(1, rk + 1):
            o.compute()
            o.rk_stage(s)
        o.post_step()
    want = o.get(FST_VAR)
    d = Dims(cfg)
    sl = d.interior()
    assert np.array_equal(got[sl], want[sl])


@pytest.mark.gpu
def test_driver_gen3d_module_equals_oracle(tmp_path):
    """gen3d_func through fiesta_run: grid upload, preSim metrics, rk sequence."""
    import oracle
    from fiesta_amd.config import FST_GRID, make_config
    print("case: fiesta_run gen3d rk", flush=True)
    cfg = make_config(ndim=3, ns=2, nci=18, ncj=11, nck=9, grid=1, gamma=(1.4, 1.67), M=(0.02896, 0.004),
                      mu=(0.0, 0.0), dt=2e-7)
    nodes = W.terrain_grid(cfg)
    var = W.random_state_all(cfg, seed=6, umag=20.0)
    gridf = tmp_path / "grid.bin"
    gridf.write_bytes(np.ascontiguousarray(nodes, np.float64).tobytes())
    cfgf, stf, outf = tmp_path / "cfg.bin", tmp_path / "state.bin", tmp_path / "out.bin"
    cfgf.write_bytes(bytes(cfg))
    stf.write_bytes(np.ascontiguousarray(var, np.float64).tobytes())
    r = subprocess.run([RUN, str(cfgf), str(stf), str(outf), "3", "rk", str(gridf)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.frombuffer(outf.read_bytes(), np.float64).reshape(var.shape)
    o = oracle.Oracle(cfg)
    o.set(FST_GRID, nodes)
    o.set(FST_VAR, var)
    o.pre_sim()
    o.step(3)
    assert np.array_equal(got, o.get(FST_VAR))

