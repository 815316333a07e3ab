"""statusCheck on the device (src/status.cpp:155-251): fst_status's per-variable
min/max over the interior cells equal numpy's over the same cells, bit for
bit, with the reference's reduction semantics (NaN skipped, infinities kept,
ghost cells ignored, an all-NaN variable reports the reduction identities),
and the C++ statusCheck prints the reference's table with its colour marks."""
import re
import subprocess
import sys

import numpy as np
import pytest

from fiesta_amd import build as B
from fiesta_amd import io
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_VAR, FST_VARX, Dims
from fiesta_amd.runtime import Module
from helpers import cfg2d, cfg3d

pytestmark = pytest.mark.gpu

DBL_MAX = np.finfo(np.float64).max


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _expect(a, d):
    """Interior min/max per variable, NaN skipped (Kokkos::Min/Max identities
    when a variable holds NaN only)."""
    sl = d.interior()[1:]
    mn, mx = [], []
    for v in range(a.shape[0]):
        x = a[v][sl].ravel()
        x = x[~np.isnan(x)]
        mn.append(x.min() if x.size else DBL_MAX)
        mx.append(x.max() if x.size else -DBL_MAX)
    return np.array(mn), np.array(mx)


@pytest.mark.parametrize("ndim", [3, 2])
def test_status_of_a_stepped_state(ndim):
    _say("case status: %dD after 2 steps" % ndim)
    cfg = cfg3d(nci=33, ncj=17, nck=12) if ndim == 3 else cfg2d(nci=70, ncj=23)
    d = Dims(cfg)
    m = Module(cfg)
    m.upload(FST_VAR, W.random_state(cfg, 31))
    m.pre_sim()
    m.step(2)
    m.post_step()
    for arr, n in ((FST_VAR, d.nvt), (FST_VARX, d.nvx)):
        mn, mx = m.status(arr)
        emn, emx = _expect(m.download(arr), d)
        assert np.array_equal(mn, emn) and np.array_equal(mx, emx), arr
    m.close()


def test_status_semantics_nan_inf_ghosts_and_identities():
    _say("case status: planted NaN, inf, ghosts")
    cfg = cfg3d(nci=20, ncj=9, nck=7)
    d = Dims(cfg)
    var = W.random_state(cfg, 32)
    g = d.ng
    var[:, :, :, :g] = 1e300          # ghost columns: never counted
    var[0, g + 2, g + 3, g + 4] = np.nan
    var[0, g + 1, g + 1, g + 1] = -np.inf
    var[1, g + 3, g + 2, g + 5] = np.inf
    var[2][d.interior()[1:]] = np.nan  # no value but NaN
    m = Module(cfg)
    m.upload(FST_VAR, var)
    mn, mx = m.status(FST_VAR)
    emn, emx = _expect(var, d)
    assert np.array_equal(mn, emn) and np.array_equal(mx, emx)
    assert mn[0] == -np.inf and mx[1] == np.inf and (mn[2], mx[2]) == (DBL_MAX, -DBL_MAX)
    # fewer variables than the array holds
    mn3, mx3 = m.status(FST_VAR, nvar=3)
    assert np.array_equal(mn3, mn[:3]) and np.array_equal(mx3, mx[:3])
    m.close()


def test_driver_prints_the_status_table(tmp_path):
    _say("case status: fiesta_run --status-color")
    cfg = cfg3d(nci=16, ncj=9, nck=8)
    d = Dims(cfg)
    var = W.random_state(cfg, 33)
    g = d.ng
    var[0, g + 1, g + 2, g + 3] = np.inf      # bad: red
    var[3, g + 2, g + 2, g + 2] = 1.0e17      # concern: yellow
    cfgf, stf = tmp_path / "c.bin", tmp_path / "s.bin"
    cfgf.write_bytes(bytes(cfg))
    stf.write_bytes(np.ascontiguousarray(var).tobytes())
    r = subprocess.run([B.RUN_BIN, str(cfgf), str(stf), str(tmp_path / "o.bin"), "0", "rk", "--status-color",
                        "--species", "Air,He"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = {}
    for line in r.stdout.splitlines():
        m = re.match(r"\s{8}(\S.{0,15}?)\s*(\x1b\[0;3\dm)\s*(\S+)\x1b\[0m(\x1b\[0;3\dm)\s*(\S+)\x1b\[0m$", line)
        if m:
            rows[m.group(1).strip()] = (m.group(2), float(m.group(3)), m.group(4), float(m.group(5)))
    names = io.var_names(cfg, ["Air", "He"])
    assert all(n in rows for n in names), rows.keys()
    red, yellow, magenta = "\x1b[0;31m", "\x1b[0;33m", "\x1b[0;35m"
    assert rows["X-Momentum"][2] == red and rows["X-Momentum"][3] == np.inf
    assert rows["Energy"][2] == yellow and rows["Energy"][3] == pytest.approx(1e17, rel=1e-2)
    assert rows["Air Density"][0] == magenta and rows["Air Density"][2] == magenta
    emn, emx = _expect(var, d)
    for v, n in enumerate(names):
        assert rows[n][1] == pytest.approx(emn[v], rel=6e-3, abs=1e-300)  # printed with %.2e
