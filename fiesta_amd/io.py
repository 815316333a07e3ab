"""ctypes binding of include/fiesta_io.h (libfiesta_host.so): FIESTA's solution
and restart files (HDF5 + XDMF) written from and read into a `Module`.

`var_names` / `varx_names` give the dataset names the reference's modules use
(src/cart3d.cpp:98-121, src/cart2d.cpp:41-85, src/gen3d.cpp:67-86,
src/gen2d.cpp:37-78).
"""
import ctypes as C
import os

from .config import Dims
from .runtime import FiestaError, lib as hip_lib

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfiesta_host.so")
_LIB = None

RESTART_VERSION = 2  # FIESTA_RESTART_VERSION, src/fiesta.hpp:31


class FioBlock(C.Structure):
    _fields_ = [("name", C.c_char_p), ("path", C.c_char_p), ("whole", C.c_int), ("start", C.c_int64 * 3),
                ("end", C.c_int64 * 3), ("stride", C.c_int64 * 3), ("average", C.c_int), ("append_step", C.c_int),
                ("f64", C.c_int), ("chunk", C.c_int), ("compress", C.c_int), ("tend", C.c_int64),
                ("title", C.c_char_p), ("metadata", C.c_char_p)]


class FioExtents(C.Structure):
    _fields_ = [("g_ext", C.c_int64 * 3), ("g_ext_g", C.c_int64 * 3), ("l_start", C.c_int64 * 3),
                ("l_end", C.c_int64 * 3), ("l_end_g", C.c_int64 * 3), ("l_ext", C.c_int64 * 3),
                ("l_ext_g", C.c_int64 * 3), ("origin", C.c_double * 3), ("spacing", C.c_double * 3),
                ("write_varx", C.c_int), ("chunk", C.c_int)]


_EXPORTS = {
    "fio_block_extents": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t], C.c_int),
    "fio_write": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                   C.c_double, C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t], C.c_int),
    "fio_read_restart": ([C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                          C.POINTER(C.c_double), C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t], C.c_int),
}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FiestaError("libfiesta_host.so is not built: run __graft_entry__.build()")
        hip_lib()  # the host library links libfiesta_hip.so
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _EXPORTS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_EXPORTS)


def var_names(cfg, species=None):
    d = Dims(cfg)
    sp = list(species or [])
    sname = [sp[s] if s < len(sp) else "Species %d" % s for s in range(cfg.ns)]
    if cfg.grid == 1:
        base = ["X-Momentum", "Y-Momentum"] + (["Z-Momentum"] if cfg.ndim == 3 else []) + ["Energy"]
        return base + ["Density " + s for s in sname]
    if cfg.ndim == 3:
        names = ["X-Momentum", "Y-Momentum", "Z-Momentum", "Energy"] + [s + " Density" for s in sname]
        return names + (["C", "C_hat", "C_1", "C_2", "C_3"] if d.nvt > d.nv else [])
    names = ["X-Momentum", "Y-Momentum", "Energy"] + [s + " Density" for s in sname]
    return names + (["C", "C_hat", "Tau_1", "Tau_2", "Debug"] if d.nvt > d.nv else [])


def varx_names(cfg):
    vel = ["X-Velocity", "Y-Velocity"] + (["Z-Velocity"] if cfg.ndim == 3 else [])
    return vel + ["Pressure", "Temperature", "Density" if cfg.grid == 1 else "Total Density"]


def block(name, path, whole=True, start=(0, 0, 0), end=(0, 0, 0), stride=(1, 1, 1), average=False,
          append_step=True, f64=False, chunk=False, compress=False, tend=1, title="", metadata="none"):
    b = FioBlock()
    b.name, b.path = name.encode(), path.encode()
    b.whole = int(whole)
    for a in range(3):
        b.start[a] = start[a] if a < len(start) else 0
        b.end[a] = end[a] if a < len(end) else 0
        b.stride[a] = stride[a] if a < len(stride) else 1
    b.average, b.append_step, b.f64 = int(average), int(append_step), int(f64)
    b.chunk, b.compress, b.tend = int(chunk), int(compress), int(tend)
    b.title, b.metadata = title.encode(), metadata.encode()
    return b


def block_extents(cfg, b):
    """The extents a writer uses; raises FiestaError on an invalid view."""
    out = FioExtents()
    err = C.create_string_buffer(512)
    rc = lib().fio_block_extents(C.byref(cfg), C.byref(b), C.byref(out), err, len(err))
    if rc:
        raise FiestaError("fio_block_extents: %s (code %d)" % (err.value.decode(), rc))
    nd = cfg.ndim
    r = {k: tuple(getattr(out, k)[:nd]) for k in ("g_ext", "g_ext_g", "l_start", "l_end", "l_end_g", "l_ext",
                                                  "l_ext_g", "origin", "spacing")}
    r["write_varx"], r["chunk"] = bool(out.write_varx), bool(out.chunk)
    return r


def _names(names):
    arr = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    return arr, len(names)


def write(module, b, species=None, tdx=0, time=0.0):
    """blockWriter::write of `module`'s current device state; returns the .h5 path."""
    cfg = module.cfg
    vn, nv = _names(var_names(cfg, species))
    vx, nx = _names(varx_names(cfg))
    path = C.create_string_buffer(4096)
    err = C.create_string_buffer(512)
    rc = lib().fio_write(module.h, C.byref(cfg), C.byref(b), vn, nv, vx, nx, int(tdx), float(time), path, len(path),
                         err, len(err))
    if rc:
        raise FiestaError("fio_write: %s (code %d)" % (err.value.decode(), rc))
    return path.value.decode()


def read_restart(module, fname, species=None, reset=False):
    """readRestart into `module`; returns (time_index, time, title)."""
    cfg = module.cfg
    vn, nv = _names(var_names(cfg, species))
    tdx, tm = C.c_int(0), C.c_double(0.0)
    title = C.create_string_buffer(1024)
    err = C.create_string_buffer(512)
    rc = lib().fio_read_restart(module.h, C.byref(cfg), fname.encode(), vn, nv, int(reset), C.byref(tdx),
                                C.byref(tm), title, len(title), err, len(err))
    if rc:
        raise FiestaError("fio_read_restart: %s (code %d)" % (err.value.decode(), rc))
    return tdx.value, tm.value, title.value.decode()
