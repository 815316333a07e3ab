"""GPU output pack (fst_pack: blockWriter::dataPack on the device) against the
oracle's dataPack, bit for bit, for var, varx and the grid nodes, strided and
averaged, float and double; boxes that leave the array are rejected."""
import sys

import numpy as np
import pytest

import oracle
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_GRID, FST_VAR, FST_VARX, make_config
from fiesta_amd.runtime import FiestaError, Module
from helpers import cfg2d, cfg3d

pytestmark = pytest.mark.gpu

BOXES_3D = [((0, 0, 0), (20, 12, 10), (1, 1, 1), False), ((1, 2, 0), (19, 11, 9), (2, 3, 2), False),
            ((0, 0, 0), (19, 11, 7), (4, 2, 4), True), ((3, 0, 1), (3, 11, 10), (1, 1, 1), True)]


def _pair(cfg, seed):
    var = W.random_state(cfg, seed)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    m = Module(cfg)
    m.upload(FST_VAR, var)
    o.pre_sim(); m.pre_sim()  # varx
    return o, m


@pytest.mark.parametrize("f32", [True, False])
def test_pack_3d(f32):
    cfg = cfg3d(nci=21, ncj=13, nck=11)
    print("case pack 3d f32=%s" % f32, flush=True)
    o, m = _pair(cfg, 71)
    for start, end, stride, avg in BOXES_3D:
        for arr, vn in ((FST_VAR, 0), (FST_VAR, 8), (FST_VARX, 4)):
            got = m.pack(arr, vn, start, end, stride, average=avg, f32=f32)
            want = o.pack(arr, vn, start, end, stride, average=avg, f32=f32)
            assert got.dtype == want.dtype and np.array_equal(got, want), (start, end, stride, avg, arr, vn)
    m.close()


def test_pack_2d_and_grid():
    print("case pack 2d + grid", flush=True)
    cfg = cfg2d(nci=30, ncj=14)
    o, m = _pair(cfg, 72)
    for start, end, stride, avg in [((0, 0), (29, 13), (1, 1), False), ((1, 1), (28, 12), (3, 2), True)]:
        got = m.pack(FST_VAR, 3, start, end, stride, average=avg, f32=True)
        assert np.array_equal(got, o.pack(FST_VAR, 3, start, end, stride, average=avg, f32=True))
    m.close()
    g = make_config(ndim=3, ns=1, nci=9, ncj=7, nck=5, grid=1, gamma=(1.4,), M=(0.02896,), mu=(0.0,), dt=1e-7)
    nodes = W.terrain_grid(g)
    o = oracle.Oracle(g)
    o.set(FST_GRID, nodes)
    m = Module(g)
    m.upload(FST_GRID, nodes)
    got = m.pack(FST_GRID, 1, (0, 0, 0), (9, 7, 5), (2, 2, 1), f32=False)
    assert np.array_equal(got, o.pack(FST_GRID, 1, (0, 0, 0), (9, 7, 5), (2, 2, 1), f32=False))
    m.close()


def test_pack_rejects_boxes_outside_the_array():
    cfg = cfg3d(nci=8, ncj=8, nck=8)
    m = Module(cfg)
    with pytest.raises(FiestaError):
        m.pack(FST_VAR, 0, (0, 0, 0), (20, 7, 7), (1, 1, 1))   # past the ghosts
    with pytest.raises(FiestaError):
        m.pack(FST_VAR, 99, (0, 0, 0), (7, 7, 7), (1, 1, 1))   # no such variable
    with pytest.raises(FiestaError):
        m.pack(FST_VAR, 0, (5, 0, 0), (2, 7, 7), (1, 1, 1))    # end before start
    m.close()
