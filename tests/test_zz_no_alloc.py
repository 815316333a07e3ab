"""fst_step_static, the fused step without device allocation, against
fst_step on a context (the existing entry point): bit for bit, on two calls
in a row that share the static scratch array, the second on a smaller grid.
The caller's state lives in device memory it allocates itself."""
import ctypes as C

import numpy as np
import pytest

from fiesta_amd import workloads as W
from fiesta_amd.config import FST_VAR
from fiesta_amd.runtime import Module, lib
from helpers import assert_bitwise, cfg3d

pytestmark = pytest.mark.gpu


def _with_context(cfg, var, nsteps):
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.pre_sim()
    m.set_fused(True)
    m.step(nsteps)
    out = m.download(FST_VAR)
    m.close()
    return out


def _hip():
    """The HIP runtime libfiesta_hip.so is linked against (torch bundles a
    runtime of its own, which cannot share the device with this one)."""
    lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _static(cfg, var, nsteps):
    h = _hip()
    a = np.ascontiguousarray(var, np.float64)
    dev = C.c_void_p()
    assert h.hipMalloc(C.byref(dev), a.nbytes) == 0
    try:
        assert h.hipMemcpy(dev, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        rc = lib().fst_step_static(C.byref(cfg), dev, nsteps, None)
        assert rc == 0, rc
        assert h.hipDeviceSynchronize() == 0
        out = np.empty_like(a)
        assert h.hipMemcpy(out.ctypes.data, dev, a.nbytes, 2) == 0  # hipMemcpyDeviceToHost
    finally:
        h.hipFree(dev)
    return out


@pytest.mark.parametrize("rk,precision,ceq", [(2, 0, True), (3, 0, True), (2, 1, True), (2, 1, False)])
def test_static_step_equals_context_step_on_two_calls(rk, precision, ceq):
    for n, (nci, ncj, nck) in enumerate([(24, 13, 10), (12, 9, 7)]):
        print("case: fst_step_static rk%d precision %d ceq %d call %d grid %dx%dx%d"
              % (rk, precision, ceq, n, nci, ncj, nck), flush=True)
        cfg = cfg3d(nci=nci, ncj=ncj, nck=nck, ceq=ceq)
        cfg.rk_order = rk
        cfg.precision = precision
        var = W.random_state(cfg, 11 + n)
        want = _with_context(cfg, var, 2)
        got = _static(cfg, var, 2)
        assert_bitwise(got, want, "fst_step_static call %d vs fst_step" % n)
