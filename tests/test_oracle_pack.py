"""The oracle's output pack (blockWriter::dataPack, src/block.cpp:404-454)
against numpy slicing: stride-1 blocks are the interior values, strided
blocks every stride-th value, averaged blocks the mean of each stride block
(exact on a linear field: the value at the block centre), float output the
binary32 rounding of the same."""
import numpy as np
import pytest

import oracle
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_GRID, FST_VAR, Dims, make_config
from helpers import cfg2d, cfg3d


def _oracle(cfg, var):
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    return o


@pytest.mark.parametrize("stride", [(1, 1, 1), (2, 3, 2), (4, 1, 3)])
def test_3d_strided_pack_is_a_slice(stride):
    cfg = cfg3d(nci=13, ncj=10, nck=9)
    d = Dims(cfg)
    var = W.random_state(cfg, 5)
    o = _oracle(cfg, var)
    start, end = (1, 2, 0), (12, 9, 8)
    got = o.pack(FST_VAR, 4, start, end, stride, f32=False)
    g = d.ng
    want = var[4][g + start[2]:g + end[2] + 1:stride[2], g + start[1]:g + end[1] + 1:stride[1],
                  g + start[0]:g + end[0] + 1:stride[0]]
    assert np.array_equal(got, want)
    assert np.array_equal(o.pack(FST_VAR, 4, start, end, stride, f32=True), want.astype(np.float32))


def test_3d_average_of_a_linear_field_is_the_block_centre():
    cfg = cfg3d(nci=12, ncj=8, nck=8)
    d = Dims(cfg)
    var = W.random_state(cfg, 6)
    k, j, i = np.meshgrid(np.arange(d.ngk), np.arange(d.ngj), np.arange(d.ngi), indexing="ij")
    var[0] = 0.5 * i + 0.25 * j + 2.0 * k  # exact in binary
    o = _oracle(cfg, var)
    st = (2, 2, 4)
    got = o.pack(FST_VAR, 0, (0, 0, 0), (10, 6, 4), st, average=True, f32=False)
    g = d.ng
    ii, jj, kk = np.arange(0, 11, 2), np.arange(0, 7, 2), np.arange(0, 5, 4)
    ci = g + ii + (st[0] - 1) / 2.0
    cj = g + jj + (st[1] - 1) / 2.0
    ck = g + kk + (st[2] - 1) / 2.0
    want = 0.5 * ci[None, None, :] + 0.25 * cj[None, :, None] + 2.0 * ck[:, None, None]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_2d_pack_and_average():
    cfg = cfg2d(nci=16, ncj=8)
    d = Dims(cfg)
    var = W.random_state(cfg, 7)
    o = _oracle(cfg, var)
    got = o.pack(FST_VAR, 2, (0, 1), (15, 7), (3, 2), f32=False)
    g = d.ng
    assert np.array_equal(got, var[2][0, g + 1:g + 8:2, g:g + 16:3])
    avg = o.pack(FST_VAR, 2, (0, 0), (14, 6), (2, 2), average=True, f32=False)
    blk = var[2][0, g:g + 8, g:g + 16].reshape(4, 2, 8, 2)
    np.testing.assert_allclose(avg, blk.mean(axis=(1, 3)), rtol=1e-15)


def test_grid_nodes_pack():
    cfg = make_config(ndim=3, ns=1, nci=6, ncj=5, nck=4, grid=1, gamma=(1.4,), M=(0.02896,), mu=(0.0,), dt=1e-7)
    d = Dims(cfg)
    nodes = W.terrain_grid(cfg)
    o = oracle.Oracle(cfg)
    o.set(FST_GRID, nodes)
    got = o.pack(FST_GRID, 2, (0, 0, 0), (6, 5, 4), (1, 1, 1), f32=False)
    assert np.array_equal(got, nodes[2])
