"""ctypes handle on the Lua input deck reader (include/fiesta_deck.h, in
libfiesta_host.so): executeConfiguration, loadGrid, loadInitialConditions
and luaReader::call (src/input.cpp:130-557, src/luaReader.cpp:36-345) on a
FIESTA deck, without a Lua installation.  Host code only (no GPU)."""
import ctypes as C
import os

import numpy as np

from fiesta_amd.config import Dims, FstConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class FdkRun(C.Structure):
    _fields_ = [
        ("title", C.c_char * 256), ("metadata", C.c_char * 256),
        ("nt", C.c_int64), ("tstart", C.c_int64), ("tend", C.c_int64), ("tinterval", C.c_int32),
        ("time0", C.c_double), ("out_freq", C.c_int64), ("stat_freq", C.c_int64), ("write_freq", C.c_int64),
        ("restart", C.c_int32), ("restart_reset", C.c_int32), ("auto_restart", C.c_int32),
        ("restart_freq", C.c_int64), ("restart_time_remaining", C.c_int64),
        ("restart_name", C.c_char * 512), ("restart_path", C.c_char * 512), ("auto_restart_name", C.c_char * 256),
        ("hdf5_chunk", C.c_int32), ("hdf5_compress", C.c_int32), ("mpi_type", C.c_char * 32),
        ("procs", C.c_int32 * 3), ("terrain_name", C.c_char * 256), ("tdx", C.c_double), ("tdy", C.c_double),
        ("h", C.c_double), ("beta", C.c_double), ("betae", C.c_double), ("st", C.c_int64),
        ("n_ioviews", C.c_int32),
    ]


class FdkIoview(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("path", C.c_char * 512), ("frequency", C.c_int64),
                ("average", C.c_int32), ("whole", C.c_int32), ("start", C.c_int64 * 3), ("limit", C.c_int64 * 3),
                ("stride", C.c_int64 * 3)]


class DeckError(RuntimeError):
    def __init__(self, msg, code):
        super().__init__(msg)
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libfiesta_host.so")
        if not os.path.exists(path):
            raise RuntimeError("libfiesta_host.so is not built: run __graft_entry__.build()")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.fdk_open.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.fdk_close.argtypes = [vp]
        L.fdk_last_error.argtypes = [vp]
        L.fdk_last_error.restype = C.c_char_p
        L.fdk_read_config.argtypes = [vp, vp, vp]
        L.fdk_species_name.argtypes = [vp, C.c_int, C.c_char_p, C.c_size_t]
        L.fdk_get_ioview.argtypes = [vp, C.c_int, vp]
        L.fdk_grid.argtypes = [vp, vp, vp, C.c_size_t, C.c_int]
        L.fdk_initial_conditions.argtypes = [vp, vp, vp, vp, C.c_size_t, C.c_int]
        L.fdk_call.argtypes = [vp, C.c_char_p, C.c_int, vp, C.POINTER(C.c_double)]
        L.fdk_get.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_char_p, C.c_size_t]
        for n in ("fdk_open", "fdk_close", "fdk_read_config", "fdk_species_name", "fdk_get_ioview", "fdk_grid",
                  "fdk_initial_conditions", "fdk_call", "fdk_get"):
            getattr(L, n).restype = C.c_int
        _LIB = L
    return _LIB


class Deck:
    """One FIESTA input deck (the script has run once)."""

    def __init__(self, path):
        self.path = path
        h = C.c_void_p()
        rc = lib().fdk_open(os.fsencode(path), C.byref(h))
        if rc:
            raise DeckError("fdk_open failed (code %d) for %s" % (rc, path), rc)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fdk_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            raise DeckError("%s failed (code %d): %s" % (what, rc, lib().fdk_last_error(self.h).decode()), rc)

    def config(self):
        """(FstConfig, run settings dict) as executeConfiguration sets them."""
        cfg, run = FstConfig(), FdkRun()
        self._chk(lib().fdk_read_config(self.h, C.byref(cfg), C.byref(run)), "fdk_read_config")
        r = {}
        for name, _ in FdkRun._fields_:
            v = getattr(run, name)
            r[name] = v.decode() if isinstance(v, bytes) else (list(v) if name == "procs" else v)
        self._cfg = cfg
        return cfg, r

    def species(self):
        out = []
        for s in range(self._cfg.ns):
            buf = C.create_string_buffer(256)
            self._chk(lib().fdk_species_name(self.h, s, buf, 256), "fdk_species_name")
            out.append(buf.value.decode())
        return out

    def ioviews(self, n):
        out = []
        for i in range(n):
            v = FdkIoview()
            self._chk(lib().fdk_get_ioview(self.h, i, C.byref(v)), "fdk_get_ioview")
            out.append(dict(name=v.name.decode(), path=v.path.decode(), frequency=v.frequency, average=v.average,
                            whole=v.whole, start=list(v.start), limit=list(v.limit), stride=list(v.stride)))
        return out

    def grid(self, cfg, threads=0):
        """Node coordinates (3, nk, nj, ni) from fiesta.initialize_grid."""
        d = Dims(cfg)
        g = np.zeros(d.grid_shape(), np.float64)
        self._chk(lib().fdk_grid(self.h, C.byref(cfg), g.ctypes.data, g.size, int(threads)), "fdk_grid")
        return g

    def initial_conditions(self, cfg, grid=None, threads=0):
        """var (nvt, ngk, ngj, ngi) from fiesta.initial_conditions."""
        d = Dims(cfg)
        var = np.zeros(d.shape(d.nvt), np.float64)
        gp = None if grid is None else np.ascontiguousarray(grid, np.float64).ctypes.data
        self._chk(lib().fdk_initial_conditions(self.h, C.byref(cfg), gp, var.ctypes.data, var.size, int(threads)),
                  "fdk_initial_conditions")
        return var

    def call(self, fname, *args):
        a = (C.c_double * max(len(args), 1))(*[float(x) for x in args])
        out = C.c_double()
        self._chk(lib().fdk_call(self.h, fname.encode(), len(args), a, C.byref(out)), "fdk_call")
        return out.value

    def get(self, key):
        """Value below the root table at a dotted key: None, bool, float, str,
        or the strings 'table' / 'function'."""
        kind, num = C.c_int(), C.c_double()
        buf = C.create_string_buffer(4096)
        self._chk(lib().fdk_get(self.h, key.encode(), C.byref(kind), C.byref(num), buf, 4096), "fdk_get")
        return [None, bool(num.value), num.value, buf.value.decode(), "table", "function"][kind.value]
