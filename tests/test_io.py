"""The writers' host logic, no GPU: the extents blockWriter's constructors
compute (src/block.cpp:44-280) for one process, their argument checks, the
library's exports and the fiesta_run options checked before a device is used.

Known answers follow from the rules of a view (cells start, start + s, ...
up to the limit; a stride block and the node after the last cell stay inside
the domain), not from the code's formulas.
"""
import os
import re
import subprocess

import pytest

from fiesta_amd import build as B
from fiesta_amd import io
from fiesta_amd.config import make_config
from fiesta_amd.runtime import FiestaError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cfg(ndim=3, nci=10, ncj=8, nck=6):
    return make_config(ndim=ndim, ns=1, nci=nci, ncj=ncj, nck=nck, dx=0.1, dy=0.2, dz=0.3)


def test_host_library_exports_every_declared_function():
    src = open(os.path.join(ROOT, "include", "fiesta_io.h")).read()
    declared = sorted(set(re.findall(r"^\s*int\s+(fio_\w+)\s*\(", src, re.M)))
    assert declared == sorted(io.exported_symbols())
    L = io.lib()
    assert all(hasattr(L, n) for n in declared)


def test_whole_domain_extents():
    e = io.block_extents(_cfg(), io.block("r", "/tmp", whole=True, f64=True))
    assert e["g_ext"] == (10, 8, 6) and e["l_ext"] == (10, 8, 6)
    assert e["g_ext_g"] == (11, 9, 7) and e["l_ext_g"] == (11, 9, 7)  # nodes
    assert e["l_start"] == (0, 0, 0) and e["l_end"] == (9, 7, 5) and e["l_end_g"] == (10, 8, 6)
    assert e["origin"] == (0.0, 0.0, 0.0) and e["spacing"] == (0.1, 0.2, 0.3)
    assert not e["write_varx"]


def test_whole_domain_keeps_the_chunk_setting_views_drop_it():
    assert io.block_extents(_cfg(), io.block("r", "/tmp", chunk=True))["chunk"]
    v = io.block("v", "/tmp", whole=False, start=(0, 0, 0), end=(9, 7, 5), stride=(1, 1, 1), chunk=True)
    assert not io.block_extents(_cfg(), v)["chunk"]


def test_stride_two_view_with_averaging():
    """Cells 0..9 every 2 along x: 0, 2, 4, 6, 8 (the block of 8 reads 8 and 9,
    inside the 10 cells), nodes 0, 2, ..., 10; y, z whole."""
    v = io.block("v", "/tmp", whole=False, start=(0, 0, 0), end=(9, 7, 5), stride=(2, 1, 1), average=True)
    e = io.block_extents(_cfg(), v)
    assert e["g_ext"][0] == 5 and e["l_ext"][0] == 5 and e["l_end"][0] == 8
    assert e["g_ext_g"][0] == 6 and e["l_ext_g"][0] == 6 and e["l_end_g"][0] == 10
    assert e["g_ext"][1:] == (8, 6) and e["l_ext_g"][1:] == (9, 7)
    assert e["spacing"] == (0.2, 0.2, 0.3) and e["write_varx"]


def test_last_block_that_would_leave_the_domain_is_dropped():
    """9 cells, 0..8 every 2: the block of cell 8 would read cell 9, so the view
    ends at 6 (4 cells; nodes 0..8)."""
    cfg = _cfg(nci=9)
    v = io.block("v", "/tmp", whole=False, start=(0, 0, 0), end=(8, 7, 5), stride=(2, 1, 1))
    e = io.block_extents(cfg, v)
    assert e["g_ext"][0] == 4 and e["l_ext"][0] == 4 and e["l_end"][0] == 6
    assert e["l_end_g"][0] == 8 and e["l_ext_g"][0] == 5 and e["g_ext_g"][0] == 5


def test_offset_view_aligns_its_end_to_the_stride():
    """12 cells, 2..9 every 3: cells 2, 5, 8 (9 is not on the stride), nodes
    2, 5, 8, 11; origin 2 dx, spacing 3 dx."""
    cfg = _cfg(nci=12)
    v = io.block("v", "/tmp", whole=False, start=(2, 0, 0), end=(9, 7, 5), stride=(3, 1, 1))
    e = io.block_extents(cfg, v)
    assert e["l_start"][0] == 2 and e["l_end"][0] == 8 and e["g_ext"][0] == 3
    assert e["l_end_g"][0] == 11 and e["l_ext_g"][0] == 4
    assert e["origin"][0] == pytest.approx(0.2) and e["spacing"][0] == pytest.approx(0.3)


def test_single_plane_view():
    v = io.block("v", "/tmp", whole=False, start=(4, 0, 3), end=(4, 7, 3), stride=(1, 1, 1))
    e = io.block_extents(_cfg(), v)
    assert e["g_ext"] == (1, 8, 1) and e["l_ext_g"] == (2, 9, 2)


def test_2d_view():
    cfg = _cfg(ndim=2, nci=16, ncj=9)
    v = io.block("v", "/tmp", whole=False, start=(1, 0), end=(14, 8), stride=(4, 2))
    e = io.block_extents(cfg, v)
    # x: 1, 5, 9, 13 (13's block reads 13..16: cell 16 is outside -> drop 13)
    assert e["l_end"] == (9, 6) and e["g_ext"] == (3, 4) and e["l_ext_g"] == (4, 5)


@pytest.mark.parametrize("start,end,stride,what", [
    ((11, 0, 0), (11, 7, 5), (1, 1, 1), "Starting index"),
    ((0, 0, 0), (11, 7, 5), (1, 1, 1), "Ending index"),
    ((5, 0, 0), (4, 7, 5), (1, 1, 1), "exceeds ending index"),
    ((0, 0, 0), (9, 7, 5), (5, 1, 1), "exceeds 4"),
    ((0, 0, 0), (1, 7, 5), (2, 1, 1), "exceeds view dimensions"),
    ((9, 0, 0), (9, 7, 5), (2, 1, 1), "holds no cells"),
])
def test_invalid_views_are_rejected(start, end, stride, what):
    with pytest.raises(FiestaError, match=what):
        io.block_extents(_cfg(), io.block("v", "/tmp", whole=False, start=start, end=end, stride=stride))


def test_variable_names_follow_the_modules():
    c3 = make_config(ndim=3, ns=2, nci=4, ncj=4, nck=4, gamma=(1.4, 1.67), M=(0.029, 0.004), mu=(0, 0), ceq=True)
    assert io.var_names(c3, ["Air", "He"]) == ["X-Momentum", "Y-Momentum", "Z-Momentum", "Energy", "Air Density",
                                                "He Density", "C", "C_hat", "C_1", "C_2", "C_3"]
    c2 = make_config(ndim=2, ns=1, nci=4, ncj=4, ceq=True)
    assert io.var_names(c2, ["Air"])[-5:] == ["C", "C_hat", "Tau_1", "Tau_2", "Debug"]
    g2 = make_config(ndim=2, ns=1, nci=4, ncj=4, grid=1)
    assert io.var_names(g2, ["Air"]) == ["X-Momentum", "Y-Momentum", "Energy", "Density Air"]
    assert io.varx_names(g2)[-1] == "Density" and io.varx_names(c3)[-1] == "Total Density"


def test_driver_checks_io_options_before_using_a_device(tmp_path):
    run = B.RUN_BIN
    cfgf = tmp_path / "c.bin"
    cfgf.write_bytes(bytes(_cfg()))
    r = subprocess.run([run, str(cfgf), "-", str(tmp_path / "o.bin"), "1"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "--restart-in" in r.stderr
    r = subprocess.run([run, str(cfgf), "-", str(tmp_path / "o.bin"), "1", "--view", "a"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
