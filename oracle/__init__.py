"""TEST INFRASTRUCTURE ONLY -- Python handle on the oracle (liborc.so).

The oracle is a CPU restatement of FIESTA's RHS/RK path (fiesta_oracle.cpp).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it,
and only as the checker; the product never imports this package.
"""
import ctypes as C
import os

import numpy as np

from fiesta_amd.config import (Dims, FST_VAR, FST_TMP1, FST_DVAR, FST_VARX, FST_P, FST_RHO, FST_T,
                               FST_VEL, FST_GRADRHO, FST_GRID, FST_METRICS)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_void_p]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_size.restype = C.c_longlong
        L.orc_size.argtypes = [C.c_void_p, C.c_int]
        for n in ("orc_set", "orc_get"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_int, C.c_void_p]
            getattr(L, n).restype = C.c_int
        for n in ("orc_apply_bcs", "orc_compute", "orc_pre_sim", "orc_pre_step", "orc_post_step"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.orc_rk_stage.argtypes = [C.c_void_p, C.c_int]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_maxima.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        for n in ("orc_weno3d", "orc_weno2d"):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [C.c_double] * 5
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        L.orc_set_learned.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]
        L.orc_set_learned.restype = C.c_int
        L.orc_features.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.orc_features.restype = C.c_int
        L.orc_pack.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                               C.c_void_p, C.c_size_t]
        L.orc_pack.restype = C.c_int
        _LIB = L
    return _LIB


_NVARS = {FST_VAR: "nvt", FST_TMP1: "nvt", FST_DVAR: "nvt", FST_VARX: "nvx", FST_P: 1, FST_RHO: 1,
          FST_T: 1, FST_VEL: "ndim", FST_GRADRHO: None}


class Oracle:
    """One module instance (cart2d/cart3d/gen2d/gen3d semantics) on the CPU."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.dims = Dims(cfg)
        self._cfg_keep = cfg
        self.h = lib().orc_create(C.byref(cfg))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_destroy(h)
            self.h = None

    def _nvar(self, array):
        n = lib().orc_size(self.h, array)
        if n < 0:
            raise ValueError("unknown array %d" % array)
        return n // self.dims.N

    def set(self, array, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == lib().orc_size(self.h, array), (array, a.shape)
        assert lib().orc_set(self.h, array, a.ctypes.data) == 0

    def get(self, array):
        if array == FST_GRID:
            shape = self.dims.grid_shape()
        elif array == FST_METRICS:
            shape = self.dims.metrics_shape()
        else:
            shape = self.dims.shape(self._nvar(array))
        out = np.empty(shape, dtype=np.float64)
        assert lib().orc_get(self.h, array, out.ctypes.data) == 0
        return out

    def apply_bcs(self):
        lib().orc_apply_bcs(self.h)

    def compute(self):
        lib().orc_compute(self.h)

    def rk_stage(self, s):
        lib().orc_rk_stage(self.h, s)

    def pre_sim(self):
        lib().orc_pre_sim(self.h)

    def pre_step(self):
        lib().orc_pre_step(self.h)

    def post_step(self):
        lib().orc_post_step(self.h)

    def step(self, n=1):
        lib().orc_step(self.h, n)

    def maxima(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        lib().orc_maxima(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def set_learned(self, chat, direction, max_chat):
        """Learned form: C_hat [N], direction [3][N] (any float dtype; used as
        doubles) and max C_hat for the following compute() calls."""
        ch = np.ascontiguousarray(chat, dtype=np.float64).ravel()
        dr = np.ascontiguousarray(direction, dtype=np.float64).ravel()
        assert ch.size == self.dims.N and dr.size == 3 * self.dims.N
        assert lib().orc_set_learned(self.h, ch.ctypes.data, dr.ctypes.data, float(max_chat)) == 0

    def pack(self, array, vn, start, end, stride, average=False, f32=True):
        """blockWriter::dataPack (src/block.cpp:404-454); packed block shaped
        (k, j, i) (2D: (j, i))."""
        from fiesta_amd.runtime import pack_args
        shape, args = pack_args(self.cfg.ndim, start, end, stride)
        out = np.empty(int(np.prod(shape)), np.float32 if f32 else np.float64)
        rc = lib().orc_pack(self.h, array, vn, *[a.ctypes.data for a in args], int(average), int(f32),
                            out.ctypes.data, out.size)
        if rc:
            raise ValueError("orc_pack failed (%d)" % rc)
        return out.reshape(shape)

    def features(self):
        """Primitives and learned-C features of the current var:
        F [3][N], M [N], direction [3][N], G [N] (float32, (k,j,i) planes)."""
        d = self.dims
        F = np.empty(d.shape(3), np.float32)
        M = np.empty(d.shape(1), np.float32)
        Dr = np.empty(d.shape(3), np.float32)
        G = np.empty(d.shape(1), np.float32)
        assert lib().orc_features(self.h, F.ctypes.data, M.ctypes.data, Dr.ctypes.data, G.ctypes.data) == 0
        return F, M[0], Dr, G[0]


def weno3d(*f):
    return lib().orc_weno3d(*[float(x) for x in f])


def weno2d(*f):
    return lib().orc_weno2d(*[float(x) for x in f])


def set_threads(n):
    return lib().orc_set_threads(int(n))
