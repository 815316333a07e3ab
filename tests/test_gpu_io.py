"""Solution and restart files on the GPU: fiesta_run and the fio_* C-ABI write
FIESTA's HDF5 + XDMF files (src/block.cpp:282-454, src/h5.cpp, src/xdmf.cpp)
and read restarts back (src/reader.cpp:33-119).

The files are read with HDF5's own h5dump, a reader independent of the
library: every dataset equals the oracle's dataPack of the oracle's state
after the same steps, bit for bit (double restarts; float views, with and
without stride averaging), the attributes and the XDMF descriptor match, and
a run restarted from a restart file ends where the uninterrupted run ends.
"""
import os
import re
import subprocess
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle
from fiesta_amd import build as B
from fiesta_amd import io
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_GRID, FST_VAR, FST_VARX, Dims, make_config
from fiesta_amd.runtime import FiestaError, Module
from helpers import cfg2d, cfg3d

pytestmark = pytest.mark.gpu

H5DUMP = os.path.join(B.HDF5_DIR, "bin", "h5dump")
RUN = B.RUN_BIN


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _h5data(path, dset, dtype, tmp):
    out = os.path.join(str(tmp), "dump.bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", out, path], check=True, capture_output=True, timeout=60)
    return np.fromfile(out, dtype)


def _h5header(path):
    return subprocess.run([H5DUMP, "-A", path], check=True, capture_output=True, text=True,
                          timeout=60).stdout


def _datasets(header):
    """{path: (type, dims)} from h5dump's header listing."""
    out, group = {}, ""
    for line in header.splitlines():
        m = re.match(r'\s*GROUP "(.*)"', line)
        if m:
            group = "" if m.group(1) == "/" else "/" + m.group(1).strip("/")
        m = re.match(r'\s*DATASET "(.*)"', line)
        if m:
            name = group + "/" + m.group(1)
            out[name] = None
        m = re.match(r"\s*DATATYPE\s+(\S+)", line)
        if m and out and list(out.values())[-1] is None:
            last = list(out)[-1]
            out[last] = [m.group(1)]
        m = re.match(r"\s*DATASPACE\s+SIMPLE \{ \( ([\d, ]+) \)", line)
        if m and out and isinstance(out[list(out)[-1]], list) and len(out[list(out)[-1]]) == 1:
            out[list(out)[-1]].append(tuple(int(x) for x in m.group(1).split(",")))
    return {k: tuple(v) for k, v in out.items() if v}


def _attrs(header):
    """/Properties attributes: {name: value text}."""
    return {m.group(1): m.group(2).strip().strip('"')
            for m in re.finditer(r'ATTRIBUTE "(\w+)" \{.*?DATA \{\s*\(0\): (.*?)\n', header, re.S)}


def _run(args, tmp):
    r = subprocess.run([RUN] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    import json
    return json.loads(r.stdout.strip().splitlines()[-1])


def _inputs(tmp, cfg, var):
    cfgf, stf = tmp / "cfg.bin", tmp / "state.bin"
    cfgf.write_bytes(bytes(cfg))
    stf.write_bytes(np.ascontiguousarray(var, np.float64).tobytes())
    return cfgf, stf


def _oracle_steps(cfg, var, n, grid=None):
    o = oracle.Oracle(cfg)
    if grid is not None:
        o.set(FST_GRID, grid)
    o.set(FST_VAR, var)
    o.pre_sim()
    for _ in range(n):
        o.pre_step()
        for s in range(1, (3 if cfg.rk_order == 3 else 2) + 1):
            o.compute()
            o.rk_stage(s)
        o.post_step()
    return o


@pytest.fixture(scope="module", autouse=True)
def _h5dump_present():
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not installed")


@pytest.mark.parametrize("ndim", [3, 2])
def test_driver_writes_restart_and_views_equal_to_the_oracle_pack(tmp_path, ndim):
    _say("case io: fiesta_run restart + views %dD" % ndim)
    if ndim == 3:
        cfg = cfg3d(nci=20, ncj=13, nck=10)
        box = ("2,1,3", "19,13,9", "2,3,2")   # deck's 1-based start/limit, strides
        pts = ("1,1,1", "20,13,10", "3,1,4")
    else:
        cfg = cfg2d(nci=30, ncj=14)
        box = ("2,1", "29,14", "4,3")
        pts = ("1,3", "30,14", "2,1")
    var = W.random_state(cfg, 21)
    cfgf, stf = _inputs(tmp_path, cfg, var)
    out = tmp_path / "out.bin"
    d = tmp_path / "io"
    info = _run([cfgf, stf, out, 2, "rk", "--species", "Air,Helium", "--title", "io test", "--tend", 1000,
                 "--restart-out", d, "rst", "--view", "full", d, 1, "--view-box", "box", d, 1, *box,
                 "--view-box", "pts", d, 0, *pts], tmp_path)
    names = ["rst", "full", "box", "pts"]
    assert [os.path.basename(f) for f in info["files"]] == ["%s-0002.h5" % n for n in names]
    o = _oracle_steps(cfg, var, 2)
    got = np.frombuffer(out.read_bytes(), np.float64).reshape(var.shape)
    dd = Dims(cfg)
    sl = dd.interior()
    assert np.array_equal(got[sl], o.get(FST_VAR)[sl])
    vn, vxn = io.var_names(cfg, ["Air", "Helium"]), io.varx_names(cfg)
    nc = [cfg.nci, cfg.ncj, cfg.nck][:ndim]
    for name, avg, spec in (("rst", False, None), ("full", True, None), ("box", True, box), ("pts", False, pts)):
        path = str(d / ("%s-0002.h5" % name))
        head = _h5header(path)
        dsets = _datasets(head)
        at = _attrs(head)
        assert at["time_index"] == "2" and float(at["time"]) == pytest.approx(2 * cfg.dt, rel=1e-5)
        assert at["title"] == "io test" and at["metadata"] == "none" and at["restart_file_version"] == "2"
        if spec is None:
            start, end, stride = [0] * ndim, [n - 1 for n in nc], [1] * ndim
            whole = True
        else:
            start = [int(x) - 1 for x in spec[0].split(",")]
            end = [int(x) - 1 for x in spec[1].split(",")]
            stride = [int(x) for x in spec[2].split(",")]
            whole = False
        e = io.block_extents(cfg, io.block(name, str(d), whole=whole, start=start, end=end, stride=stride))
        f64 = name == "rst"
        arrays = [(FST_VAR, v, n) for v, n in enumerate(vn)]
        if not whole:
            arrays += [(FST_VARX, v, n) for v, n in enumerate(vxn)]
        expect_names = {"/Solution/" + n for _, _, n in arrays}
        assert set(k for k in dsets if k.startswith("/Solution/")) == expect_names, name
        for arr, v, n in arrays:
            typ, dims = dsets["/Solution/" + n]
            assert dims == tuple(reversed(e["g_ext"])), (name, n, dims)
            assert typ == ("H5T_IEEE_F64LE" if f64 else "H5T_IEEE_F32LE")
            want = o.pack(arr, v, e["l_start"], e["l_end"], [1] * ndim if whole else stride, average=avg,
                          f32=not f64)
            data = _h5data(path, "/Solution/" + n, np.float64 if f64 else np.float32, tmp_path)
            assert np.array_equal(data.reshape(want.shape), want, equal_nan=True), (name, n)
        # XDMF: mesh, origin and spacing of the view, one attribute per dataset
        x = ET.parse(str(d / ("%s-0002.xmf" % name))).getroot()
        topo = x.find(".//Topology")
        assert topo.get("TopologyType") == "%dDCORECTMesh" % ndim
        assert tuple(int(t) for t in topo.get("Dimensions").split()) == tuple(g + 1 for g in reversed(e["g_ext"]))
        geo = {i.get("Name"): [float(t) for t in i.text.split()] for i in x.find(".//Geometry")}
        np.testing.assert_allclose(geo["Origin"], list(reversed(e["origin"])), rtol=1e-6)
        np.testing.assert_allclose(geo["Spacing"], list(reversed(e["spacing"])), rtol=1e-6)
        attrs = [a.get("Name") for a in x.iter("Attribute")]
        want_attrs = ["Momentum"] + vn[ndim:] + (["Velocity"] + vxn[ndim:] if not whole else [])
        assert attrs == want_attrs, name
        for item in x.iter("DataItem"):
            if item.get("Format") == "HDF":
                fname, dpath = item.text.strip().split(":")
                assert fname == "%s-0002.h5" % name and dpath in dsets


@pytest.mark.parametrize("ndim,mode", [(3, "fused"), (3, "rk"), (2, "rk")])
def test_restarted_run_ends_where_the_uninterrupted_run_ends(tmp_path, ndim, mode):
    """4 steps, against 2 steps + restart file + 2 steps from it (src/reader.cpp:
    var interior, time_index and time from the file; ghosts refilled by the
    boundary conditions)."""
    _say("case io: restart round trip %dD %s" % (ndim, mode))
    cfg = cfg3d(nci=18, ncj=11, nck=9) if ndim == 3 else cfg2d(nci=24, ncj=10)
    var = W.random_state(cfg, 22)
    cfgf, stf = _inputs(tmp_path, cfg, var)
    a, b, c = tmp_path / "a.bin", tmp_path / "b.bin", tmp_path / "c.bin"
    _run([cfgf, stf, a, 4, mode], tmp_path)
    ib = _run([cfgf, stf, b, 2, mode, "--restart-out", tmp_path / "r", "chk", "--tend", 100], tmp_path)
    rfile = ib["files"][0]
    assert os.path.basename(rfile) == "chk-002.h5"
    ic = _run([cfgf, "-", c, 2, mode, "--restart-in", rfile], tmp_path)
    assert ic["tstart"] == 2 and ic["time"] == pytest.approx(4 * cfg.dt, rel=1e-15)
    d = Dims(cfg)
    sl = d.interior()
    va = np.frombuffer(a.read_bytes(), np.float64).reshape(var.shape)
    vc = np.frombuffer(c.read_bytes(), np.float64).reshape(var.shape)
    assert np.array_equal(va[sl], vc[sl])


def test_generalized_grid_restart_carries_the_grid(tmp_path):
    """gen2d: the restart file holds /Grid/Dimension0..1 (nodes); reading it
    into a fresh module restores grid and var interior, ghosts zero."""
    _say("case io: gen2d restart grid")
    cfg = make_config(ndim=2, ns=1, nci=20, ncj=12, nck=1, grid=1, gamma=(1.4,), M=(0.02896,), mu=(0.0,), dt=2e-7)
    nodes = W.terrain_grid(cfg)
    var = W.random_state_all(cfg, 23, umag=20.0)
    m = Module(cfg)
    m.upload(FST_GRID, nodes)
    m.upload(FST_VAR, var)
    path = io.write(m, io.block("g", str(tmp_path), f64=True, title="grid"), species=["Air"], tdx=7, time=1.5)
    m.close()
    head = _h5header(path)
    ds = _datasets(head)
    assert ds["/Grid/Dimension0"] == ("H5T_IEEE_F64LE", (13, 21)) and "/Grid/Dimension2" not in ds
    gx = _h5data(path, "/Grid/Dimension1", np.float64, tmp_path).reshape(13, 21)
    assert np.array_equal(gx, nodes[1, 0])
    m2 = Module(cfg)
    tdx, t, title = io.read_restart(m2, path, species=["Air"])
    assert (tdx, t, title) == (7, 1.5, "grid")
    got = m2.download(FST_VAR)
    d = Dims(cfg)
    sl = d.interior()
    assert np.array_equal(got[sl], var[sl])
    ghost = np.ones(got.shape, bool)
    ghost[sl] = False
    assert not got[ghost].any()
    g2 = m2.download(FST_GRID)
    assert np.array_equal(g2[:2], nodes[:2])
    # reset: index and time stay with the caller
    assert io.read_restart(m2, path, species=["Air"], reset=True)[:2] == (0, 0.0)
    m2.close()


def test_restart_reader_rejects_other_extents_and_missing_files(tmp_path):
    _say("case io: restart reader errors")
    small = cfg3d(nci=8, ncj=6, nck=5)
    m = Module(small)
    m.upload(FST_VAR, W.random_state(small, 24))
    path = io.write(m, io.block("s", str(tmp_path), f64=True, append_step=False))
    assert os.path.basename(path) == "s.h5"
    m.close()
    big = cfg3d(nci=9, ncj=6, nck=5)
    m = Module(big)
    with pytest.raises(FiestaError, match="Extents of restart file different"):
        io.read_restart(m, path)
    with pytest.raises(FiestaError, match="does not exist"):
        io.read_restart(m, str(tmp_path / "none.h5"))
    m.close()
