"""ctypes binding of the C-ABI in include/fiesta_hip.h (libfiesta_hip.so).

`Module` mirrors FIESTA's `rk_func` (src/rkfunction.hpp:36-69): compute(),
pre_step(), post_step(), pre_sim(), post_sim(), plus the array accessors the
reference's writers use (var, varx, dvar).  There is no CPU fallback: without
the built library every call raises.
"""
import ctypes as C
import os

import numpy as np

from .config import (Dims, FST_VAR, FST_TMP1, FST_DVAR, FST_VARX, FST_P, FST_RHO, FST_T, FST_VEL,
                     FST_GRADRHO, FST_GRID, FST_METRICS)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfiesta_hip.so")
_LIB = None

_EXPORTS = {
    "fst_abi_version": ([], C.c_int),
    "fst_config_init": ([C.c_void_p], None),
    "fst_config_dims": ([C.c_void_p, C.c_void_p], C.c_int),
    "fst_config_check": ([C.c_void_p, C.c_char_p, C.c_size_t], C.c_int),
    "fst_create": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "fst_destroy": ([C.c_void_p], C.c_int),
    "fst_last_error": ([C.c_void_p], C.c_char_p),
    "fst_get_dims": ([C.c_void_p, C.c_void_p], C.c_int),
    "fst_upload": ([C.c_void_p, C.c_int, C.c_void_p, C.c_size_t], C.c_int),
    "fst_download": ([C.c_void_p, C.c_int, C.c_void_p, C.c_size_t], C.c_int),
    "fst_pre_sim": ([C.c_void_p], C.c_int),
    "fst_pre_step": ([C.c_void_p], C.c_int),
    "fst_apply_bcs": ([C.c_void_p], C.c_int),
    "fst_compute": ([C.c_void_p], C.c_int),
    "fst_rk_stage": ([C.c_void_p, C.c_int], C.c_int),
    "fst_post_step": ([C.c_void_p], C.c_int),
    "fst_step": ([C.c_void_p, C.c_int], C.c_int),
    "fst_step_static": ([C.c_void_p, C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "fst_last_maxima": ([C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "fst_sync": ([C.c_void_p], C.c_int),
    "fst_set_fused": ([C.c_void_p, C.c_int], C.c_int),
    "fst_last_step_ms": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "fst_profile_enable": ([C.c_void_p, C.c_int], C.c_int),
    "fst_profile_read": ([C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.c_char_p, C.c_size_t],
                         C.c_int),
    "fst_copy_probe": ([C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "fst_status": ([C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "fst_pack": ([C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                  C.c_size_t], C.c_int),
    "fst_profile_phases": ([C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "fst_mlp_nparams": ([], C.c_int),
    "fst_mlp_nparams_nd": ([C.c_int], C.c_int),
    "fst_set_mlp_weights": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "fst_learned_params": ([C.c_void_p, C.c_float, C.c_float], C.c_int),
    "fst_learned_compute": ([C.c_void_p], C.c_int),
    "fst_learned_stats": ([C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)],
                          C.c_int),
    "fst_learned_fields": ([C.c_void_p] + [C.c_void_p] * 4 + [C.c_size_t], C.c_int),
    "fst_mlp_create": ([C.c_uint64, C.POINTER(C.c_void_p)], C.c_int),
    "fst_mlp_create_nd": ([C.c_int, C.c_uint64, C.POINTER(C.c_void_p)], C.c_int),
    "fst_mlp_ndim": ([C.c_void_p], C.c_int),
    "fst_mlp_destroy": ([C.c_void_p], C.c_int),
    "fst_mlp_last_error": ([C.c_void_p], C.c_char_p),
    "fst_mlp_set_weights": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "fst_mlp_get_weights": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int),
    "fst_mlp_harvest": ([C.c_void_p, C.c_void_p, C.c_uint32, C.c_int64, C.POINTER(C.c_int64)], C.c_int),
    "fst_mlp_rows": ([C.c_void_p, C.POINTER(C.c_int64)], C.c_int),
    "fst_mlp_read_rows": ([C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p], C.c_int),
    "fst_mlp_train": ([C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_uint64, C.c_void_p], C.c_int),
    "fst_mlp_forward_backward": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                  C.c_void_p], C.c_int),
}


class FiestaError(RuntimeError):
    pass


def lib():
    """Load libfiesta_hip.so; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FiestaError("libfiesta_hip.so is not built: run __graft_entry__.build() "
                              "or `python -m fiesta_amd.build` (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _EXPORTS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_EXPORTS)


class Module:
    """A cart2d/cart3d/gen2d/gen3d module instance on the GPU (one HIP stream)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._cfg_keep = cfg
        self.dims = Dims(cfg)
        h = C.c_void_p()
        rc = lib().fst_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise FiestaError("fst_create failed (code %d)" % rc)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fst_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = lib().fst_last_error(self.h)
            raise FiestaError("%s failed (code %d): %s" % (what, rc, msg.decode() if msg else ""))

    def _nvar(self, array):
        d = self.dims
        return {FST_VAR: d.nvt, FST_TMP1: d.nvt, FST_DVAR: d.nvt, FST_VARX: d.nvx, FST_P: 1, FST_RHO: 1,
                FST_T: 1, FST_VEL: d.ndim, FST_GRADRHO: 5 if d.ndim == 3 else 4}[array]

    def upload(self, array, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._chk(lib().fst_upload(self.h, array, a.ctypes.data, a.size), "fst_upload")

    def download(self, array):
        if array == FST_GRID:
            shape = self.dims.grid_shape()
        elif array == FST_METRICS:
            shape = self.dims.metrics_shape()
        else:
            shape = self.dims.shape(self._nvar(array))
        out = np.empty(shape, dtype=np.float64)
        self._chk(lib().fst_download(self.h, array, out.ctypes.data, out.size), "fst_download")
        return out

    def pre_sim(self):
        self._chk(lib().fst_pre_sim(self.h), "fst_pre_sim")

    def pre_step(self):
        self._chk(lib().fst_pre_step(self.h), "fst_pre_step")

    def apply_bcs(self):
        self._chk(lib().fst_apply_bcs(self.h), "fst_apply_bcs")

    def compute(self):
        self._chk(lib().fst_compute(self.h), "fst_compute")

    def rk_stage(self, stage):
        self._chk(lib().fst_rk_stage(self.h, stage), "fst_rk_stage")

    def post_step(self):
        self._chk(lib().fst_post_step(self.h), "fst_post_step")

    def post_sim(self):
        pass

    def status(self, array=FST_VAR, nvar=None):
        """statusCheck's per-variable (min, max) over the interior cells
        (src/status.cpp:155-231), reduced on the device."""
        d = Dims(self.cfg)
        if nvar is None:
            nvar = d.nvt if array == FST_VAR else d.ndim + 3
        mn, mx = np.empty(nvar), np.empty(nvar)
        self._chk(lib().fst_status(self.h, array, mn.ctypes.data, mx.ctypes.data, int(nvar)), "fst_status")
        return mn, mx

    def pack(self, array, vn, start, end, stride, average=False, f32=True):
        """blockWriter::dataPack on the device (fst_pack); returns the packed
        block shaped (k, j, i) (2D: (j, i))."""
        shape, args = pack_args(self.cfg.ndim, start, end, stride)
        out = np.empty(int(np.prod(shape)), np.float32 if f32 else np.float64)
        self._chk(lib().fst_pack(self.h, array, vn, *[a.ctypes.data for a in args], int(average), int(f32),
                                 out.ctypes.data, out.size), "fst_pack")
        return out.reshape(shape)

    def step(self, n=1):
        self._chk(lib().fst_step(self.h, n), "fst_step")

    def set_fused(self, on):
        self._chk(lib().fst_set_fused(self.h, int(on)), "fst_set_fused")

    def sync(self):
        self._chk(lib().fst_sync(self.h), "fst_sync")

    def last_step_ms(self):
        v = C.c_double()
        self._chk(lib().fst_last_step_ms(self.h, C.byref(v)), "fst_last_step_ms")
        return v.value

    def maxima(self):
        a, b = C.c_double(), C.c_double()
        self._chk(lib().fst_last_maxima(self.h, C.byref(a), C.byref(b)), "fst_last_maxima")
        return a.value, b.value


    def profile(self, on=True):
        self._chk(lib().fst_profile_enable(self.h, int(on)), "fst_profile_enable")

    def profile_read(self):
        n, ms = C.c_int64(), C.c_double()
        name = C.create_string_buffer(160)
        self._chk(lib().fst_profile_read(self.h, C.byref(n), C.byref(ms), name, 160), "fst_profile_read")
        return n.value, ms.value, name.value.decode()

    def profile_phases(self):
        """Per-phase ms sums of the units read by the last profile_read()."""
        ms = np.zeros(8)
        n = C.c_int()
        self._chk(lib().fst_profile_phases(self.h, ms.ctypes.data, 8, C.byref(n)), "fst_profile_phases")
        return list(ms[:n.value])

    def copy_probe(self, nbytes, reps=10):
        ms = C.c_double()
        self._chk(lib().fst_copy_probe(self.h, nbytes, reps, C.byref(ms)), "fst_copy_probe")
        return ms.value

    # -- learned C-indicator (cfg.learned_c = 1) --------------------------------

    def set_mlp_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._chk(lib().fst_set_mlp_weights(self.h, w.ctypes.data, w.size), "fst_set_mlp_weights")

    def learned_params(self, tau=1e-3, tol=0.05):
        self._chk(lib().fst_learned_params(self.h, float(tau), float(tol)), "fst_learned_params")

    def learned_compute(self):
        self._chk(lib().fst_learned_compute(self.h), "fst_learned_compute")

    def learned_stats(self):
        a, f, m = C.c_int64(), C.c_int64(), C.c_double()
        self._chk(lib().fst_learned_stats(self.h, C.byref(a), C.byref(f), C.byref(m)), "fst_learned_stats")
        return a.value, f.value, m.value

    def learned_fields(self):
        """C_hat (k,j,i), direction (3,k,j,i), features (3,k,j,i), G (k,j,i), float32."""
        d = self.dims
        ch = np.empty(d.shape(1), np.float32)
        dr = np.empty(d.shape(3), np.float32)
        ft = np.empty(d.shape(3), np.float32)
        g = np.empty(d.shape(1), np.float32)
        self._chk(lib().fst_learned_fields(self.h, ch.ctypes.data, dr.ctypes.data, ft.ctypes.data, g.ctypes.data,
                                           d.N), "fst_learned_fields")
        return ch[0], dr, ft, g[0]


def mlp_nparams(ndim=3):
    return lib().fst_mlp_nparams() if ndim == 3 else lib().fst_mlp_nparams_nd(int(ndim))


class Trainer:
    """Dataset + network of the learned C-indicator, trained on the GPU."""

    def __init__(self, seed=1, ndim=3):
        h = C.c_void_p()
        rc = lib().fst_mlp_create_nd(int(ndim), int(seed), C.byref(h))
        if rc != 0:
            raise FiestaError("fst_mlp_create_nd failed (code %d)" % rc)
        self.h = h
        self.ndim = lib().fst_mlp_ndim(h)
        self.nparams = mlp_nparams(self.ndim)
        self.nin = 375 if self.ndim == 3 else 75

    def close(self):
        if getattr(self, "h", None):
            lib().fst_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = lib().fst_mlp_last_error(self.h)
            raise FiestaError("%s failed (code %d): %s" % (what, rc, msg.decode() if msg else ""))

    def weights(self):
        w = np.empty(self.nparams, np.float32)
        self._chk(lib().fst_mlp_get_weights(self.h, w.ctypes.data, w.size), "fst_mlp_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._chk(lib().fst_mlp_set_weights(self.h, w.ctypes.data, w.size), "fst_mlp_set_weights")

    def harvest(self, module, stride=0, max_rows=1 << 30):
        n = C.c_int64()
        self._chk(lib().fst_mlp_harvest(self.h, module.h, int(stride), int(max_rows), C.byref(n)), "fst_mlp_harvest")
        return n.value

    def rows(self):
        n = C.c_int64()
        self._chk(lib().fst_mlp_rows(self.h, C.byref(n)), "fst_mlp_rows")
        return n.value

    def read_rows(self, r0, n):
        X = np.empty((n, self.nin), np.float32)
        T = np.empty(n, np.float32)
        self._chk(lib().fst_mlp_read_rows(self.h, int(r0), int(n), X.ctypes.data, T.ctypes.data), "fst_mlp_read_rows")
        return X, T

    def train(self, epochs, batch=1024, lr=1e-3, seed=7):
        loss = np.empty(epochs, np.float32)
        self._chk(lib().fst_mlp_train(self.h, int(epochs), int(batch), float(lr), int(seed), loss.ctypes.data),
                  "fst_mlp_train")
        return loss

    def forward_backward(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        M = T.size
        if X.shape != (M, self.nin):
            raise ValueError("X must be (rows, %d) for the %dD network" % (self.nin, self.ndim))
        Y = np.empty(M, np.float32)
        G = np.empty(self.nparams, np.float32)
        L = np.empty(1, np.float32)
        self._chk(lib().fst_mlp_forward_backward(self.h, X.ctypes.data, T.ctypes.data, M, Y.ctypes.data,
                                                 G.ctypes.data, L.ctypes.data), "fst_mlp_forward_backward")
        return Y, G, float(L[0])


def pack_args(ndim, start, end, stride):
    """(k, j, i) shape of a packed block and its (start, end, stride) int64
    triples, padded to three directions."""
    args = [np.zeros(3, np.int64) for _ in range(3)]
    for a, v in zip(args, (start, end, stride)):
        a[:ndim] = v[:ndim]
    args[2][ndim:] = 1
    # an empty or inverted box gives extent 0; the library rejects it
    ext = [max((int(args[1][a]) - int(args[0][a])) // max(int(args[2][a]), 1) + 1, 0) for a in range(ndim)]
    return tuple(reversed(ext)), args


def rk_advance(f):
    """rkAdvance (src/rk.cpp:28-72): two compute() calls, two stage updates
    (three of each with the SSP-RK3 option, rk_order = 3)."""
    for s in range(1, (3 if f.cfg.rk_order == 3 else 2) + 1):
        f.compute()
        f.rk_stage(s)
