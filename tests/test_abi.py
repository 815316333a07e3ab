"""C-ABI checks that need no GPU: the library loads, exports every function
include/fiesta_hip.h declares, and its host-only entry points (defaults,
derived sizes, argument validation) behave as src/input.cpp:130-443 does."""
import ctypes as C
import os
import re

import pytest

from fiesta_amd import runtime
from fiesta_amd.config import Dims, FstConfig, make_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for f in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if f.endswith(".h"):
            src = open(os.path.join(ROOT, "include", f)).read()
            names += re.findall(r"^\s*(?:int|void|const char \*)\s+(fst_\w+)\s*\(", src, re.M)
    return sorted(set(names))


class FstDims(C.Structure):
    _fields_ = [("ngi", C.c_int32), ("ngj", C.c_int32), ("ngk", C.c_int32), ("nv", C.c_int32),
                ("nvt", C.c_int32), ("nvx", C.c_int32), ("ncells_ghost", C.c_int64), ("ncells", C.c_int64)]


def test_library_exports_every_declared_symbol():
    lib = runtime.lib()
    declared = _declared()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared) <= set(runtime.exported_symbols()) | {"fst_config_init"}


def test_abi_version_and_struct_size():
    lib = runtime.lib()
    assert lib.fst_abi_version() == 7
    # fst_config layout: the ctypes mirror and the C struct must agree in size
    c = FstConfig()
    lib.fst_config_init(C.byref(c))
    assert (c.ndim, c.ng, c.ns, c.scheme) == (2, 3, 1, 1)
    assert c.R == 8.314462 and c.kap == 1.23456 and c.gAccel == 9.81


@pytest.mark.parametrize("ndim,ns,ceq", [(2, 1, 0), (2, 3, 1), (3, 2, 1), (3, 1, 0)])
def test_derived_sizes_match_input_cpp(ndim, ns, ceq):
    cfg = make_config(ndim=ndim, ns=ns, nci=17, ncj=9, nck=5, gamma=(1.4,) * ns, M=(0.029,) * ns,
                      mu=(1e-5,) * ns, ceq=bool(ceq))
    d = FstDims()
    assert runtime.lib().fst_config_dims(C.byref(cfg), C.byref(d)) == 0
    py = Dims(cfg)
    assert (d.ngi, d.ngj, d.ngk) == (py.ngi, py.ngj, py.ngk) == (23, 15, 11 if ndim == 3 else 1)
    assert d.nv == ndim + 1 + ns and d.nvt == d.nv + 5 * ceq and d.nvx == ndim + 3
    assert d.ncells == 17 * 9 * (5 if ndim == 3 else 1)


@pytest.mark.parametrize("field,value", [("ndim", 4), ("ns", 0), ("ns", 9), ("ng", 2), ("scheme", 7),
                                         ("bcL", 5), ("nci", 0), ("rk_order", 4), ("rk_order", 1)])
def test_invalid_configurations_are_rejected(field, value):
    cfg = make_config(ndim=3, ns=1, nci=8, ncj=8, nck=8)
    setattr(cfg, field, value)
    d = FstDims()
    assert runtime.lib().fst_config_dims(C.byref(cfg), C.byref(d)) == 1  # FST_ERR_ARG


def test_noise_filter_needs_nonnegative_sweeps():
    cfg = make_config(ndim=3, ns=1, nci=8, ncj=8, nck=8, noise=True, n_nt=-1)
    assert runtime.lib().fst_config_dims(C.byref(cfg), C.byref(FstDims())) == 1
    cfg.n_nt = 0
    assert runtime.lib().fst_config_dims(C.byref(cfg), C.byref(FstDims())) == 0


def test_periodic_direction_needs_ng_cells():
    cfg = make_config(ndim=3, ns=1, nci=2, ncj=8, nck=8, xper=True)
    assert runtime.lib().fst_config_dims(C.byref(cfg), C.byref(FstDims())) == 1


def test_null_arguments_are_errors_not_crashes():
    lib = runtime.lib()
    assert lib.fst_config_dims(None, None) == 1
    assert lib.fst_create(None, None) == 1
    assert lib.fst_destroy(None) == 1
    assert lib.fst_step(None, 1) == 1


def test_learned_form_sizes_and_validation():
    lib = runtime.lib()
    cfg = make_config(ndim=3, ns=2, nci=12, ncj=10, nck=8, gamma=(1.4, 1.67), M=(0.029, 0.004),
                      mu=(1e-5, 1e-5), ceq=True, learned_c=True)
    d = FstDims()
    assert lib.fst_config_dims(C.byref(cfg), C.byref(d)) == 0
    assert d.nvt == d.nv == 6 and Dims(cfg).nvt == 6  # no C-equation variables in the learned form
    for field, value in (("ceq", 0), ("noise", 1)):
        bad = make_config(ndim=3, ns=2, nci=12, ncj=10, nck=8, gamma=(1.4, 1.67), M=(0.029, 0.004),
                          mu=(1e-5, 1e-5), ceq=True, learned_c=True)
        setattr(bad, field, value)
        assert lib.fst_config_dims(C.byref(bad), C.byref(FstDims())) == 1
    # 2D: the cart2d sequence in FP64, 5x5 stencils, no C-equation variables
    cfg2 = make_config(ndim=2, ns=3, nci=16, ncj=12, gamma=(1.4, 1.67, 1.092), M=(0.029, 0.004, 0.146),
                       mu=(1e-5,) * 3, ceq=True, learned_c=True)
    d2 = FstDims()
    assert lib.fst_config_dims(C.byref(cfg2), C.byref(d2)) == 0
    assert d2.nvt == d2.nv == 6 and Dims(cfg2).nvt == 6
    cfg2.precision = 1
    assert lib.fst_config_dims(C.byref(cfg2), C.byref(FstDims())) == 1
    # parameter vector: W1[IN][64], b1, W2[64][64], b2, W3[64][64], b3, w4[64], b4
    rest = 64 + 2 * (64 * 64 + 64) + 64 + 1
    assert lib.fst_mlp_nparams() == 375 * 64 + rest
    assert lib.fst_mlp_nparams_nd(3) == 375 * 64 + rest
    assert lib.fst_mlp_nparams_nd(2) == 75 * 64 + rest
    assert lib.fst_mlp_nparams_nd(1) == -1 and lib.fst_mlp_nparams_nd(4) == -1


def test_learned_entry_points_reject_null_arguments():
    lib = runtime.lib()
    assert lib.fst_set_mlp_weights(None, None, 0) == 1
    assert lib.fst_learned_params(None, 1e-3, 0.05) == 1
    assert lib.fst_learned_compute(None) == 1
    assert lib.fst_learned_stats(None, None, None, None) == 1
    assert lib.fst_learned_fields(None, None, None, None, None, 0) == 1
    assert lib.fst_mlp_create(1, None) == 1
    assert lib.fst_mlp_create_nd(2, 1, None) == 1
    h = C.c_void_p()
    assert lib.fst_mlp_create_nd(4, 1, C.byref(h)) == 1 and not h.value  # checked before any device call
    assert lib.fst_mlp_ndim(None) == -1
    assert lib.fst_mlp_destroy(None) == 1
    assert lib.fst_mlp_train(None, 1, 16, 1e-3, 1, None) == 1
    assert lib.fst_mlp_harvest(None, None, 0, 0, None) == 1


@pytest.mark.parametrize("field,value", [("precision", 3), ("precision", -1)])
def test_unknown_precision_is_rejected(field, value):
    cfg = make_config(ndim=3, ns=1, nci=8, ncj=8, nck=8)
    setattr(cfg, field, value)
    assert runtime.lib().fst_config_dims(C.byref(cfg), C.byref(FstDims())) == 1


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("case", ["2d", "ns4", "gen3d", "over_cap"])
def test_fp32_and_fast64_forms_need_the_fused_3d_path(case, precision):
    """The FP32 form exists only in the fused 3D stage kernels (cartesian and
    gen3d), the fast FP64 form only in the cartesian ones; elsewhere the FP64
    sequence would run without notice, so the configuration is refused."""
    if case == "gen3d" and precision == 1:
        pytest.skip("the fused gen3d stage has an FP32 form (test_fp32_form_on_generalized_grids)")
    kw = dict(ndim=3, ns=2, nci=8, ncj=8, nck=8, gamma=(1.4, 1.67), M=(0.029, 0.004), mu=(1e-5, 1e-5), ceq=True)
    if case == "2d":
        kw["ndim"] = 2
    elif case == "ns4":
        kw.update(ns=4, gamma=(1.4,) * 4, M=(0.029,) * 4, mu=(1e-5,) * 4)
    elif case == "gen3d":
        kw["grid"] = 1
    else:
        kw.update(nci=740, ncj=740, nck=740)  # 746^3 cells x 11 planes > 2^32
    lib = runtime.lib()
    ok = make_config(**kw)
    if case == "over_cap":
        assert lib.fst_config_dims(C.byref(ok), C.byref(FstDims())) == 0  # FP64 runs on the sequence kernels
    cfg = make_config(precision=precision, **kw)
    assert lib.fst_config_dims(C.byref(cfg), C.byref(FstDims())) == 1
    good = make_config(precision=precision, **dict(kw, ndim=3, grid=0, nci=8, ncj=8, nck=8, ns=2, gamma=(1.4, 1.67),
                                           M=(0.029, 0.004), mu=(1e-5, 1e-5)))
    assert lib.fst_config_dims(C.byref(good), C.byref(FstDims())) == 0


def test_static_step_refuses_before_touching_a_device():
    """fst_step_static checks the configuration and the scratch size on the
    host, before any HIP call (so these run without a GPU)."""
    lib = runtime.lib()
    kw = dict(ndim=3, ns=2, gamma=(1.4, 1.67), M=(0.029, 0.004), mu=(1e-5, 1e-5), ceq=True, visc=True)
    dummy = C.c_void_p(0x1000)
    big = make_config(nci=100, ncj=100, nck=100, **kw)  # 106^3 cells x 264 B > 256 MiB
    assert lib.fst_step_static(C.byref(big), dummy, 1, None) == 1
    for bad in (make_config(nci=8, ncj=8, nck=8, learned_c=True, **kw),
                make_config(**dict(kw, ndim=2), nci=8, ncj=8),
                make_config(nci=8, ncj=8, nck=8, grid=1, **kw)):
        assert lib.fst_step_static(C.byref(bad), dummy, 1, None) == 1
    ok = make_config(nci=8, ncj=8, nck=8, **kw)
    assert lib.fst_step_static(C.byref(ok), None, 1, None) == 1  # null state
    assert lib.fst_step_static(C.byref(ok), dummy, -1, None) == 1
    assert lib.fst_step_static(None, dummy, 1, None) == 1


def test_fp32_form_on_generalized_grids():
    """precision 1 on grid 1: accepted where the fused gen3d stage runs (3D,
    WENO5, ns <= 3, arrays below 2^32 elements, with or without gen_ext),
    refused elsewhere; no learned C-indicator and no fast FP64 form there."""
    lib = runtime.lib()
    kw = dict(ndim=3, ns=3, nci=8, ncj=8, nck=8, grid=1, gamma=(1.4, 1.67, 1.09), M=(0.029, 0.004, 0.146),
              mu=(1e-5,) * 3, precision=1)
    assert lib.fst_config_dims(C.byref(make_config(**kw)), C.byref(FstDims())) == 0
    ext = dict(kw, gen_ext=True, buoyancy=True, visc=True, bc=("outflow",) * 6)
    assert lib.fst_config_dims(C.byref(make_config(**ext)), C.byref(FstDims())) == 0
    for bad in (dict(kw, ndim=2),
                dict(kw, ns=4, gamma=(1.4,) * 4, M=(0.029,) * 4, mu=(1e-5,) * 4),
                dict(kw, nci=1700, ncj=1700, nck=400),  # 1706 x 1706 x 406 cells x 7 planes > 2^32
                dict(kw, precision=2),
                dict(kw, learned_c=True, ceq=True)):
        assert lib.fst_config_dims(C.byref(make_config(**bad)), C.byref(FstDims())) == 1


def test_fast64_form_has_no_learned_variant():
    kw = dict(ndim=3, ns=2, nci=8, ncj=8, nck=8, gamma=(1.4, 1.67), M=(0.029, 0.004), mu=(1e-5, 1e-5), ceq=True)
    lib = runtime.lib()
    assert lib.fst_config_dims(C.byref(make_config(precision=2, **kw)), C.byref(FstDims())) == 0
    assert lib.fst_config_dims(C.byref(make_config(precision=2, learned_c=True, **kw)), C.byref(FstDims())) == 1


def test_gen_ext_validation():
    lib = runtime.lib()
    ok = make_config(ndim=3, ns=1, nci=8, ncj=8, nck=8, grid=1, gen_ext=True, buoyancy=True,
                     bc=("reflective", "outflow", "noslip", "outflow", "outflow", "outflow"))
    assert lib.fst_config_dims(C.byref(ok), C.byref(FstDims())) == 0
    # only on generalized grids
    cart = make_config(ndim=3, ns=1, nci=8, ncj=8, nck=8, gen_ext=True)
    assert lib.fst_config_dims(C.byref(cart), C.byref(FstDims())) == 1
    # the hydrostatic fill needs a physical cell size
    hyd = make_config(ndim=2, ns=1, nci=8, ncj=8, grid=1, gen_ext=True,
                      bc=("outflow", "outflow", "hydrostatic", "outflow", "outflow", "outflow"))
    assert lib.fst_config_dims(C.byref(hyd), C.byref(FstDims())) == 1
    hyd.gen_ext = 0
    assert lib.fst_config_dims(C.byref(hyd), C.byref(FstDims())) == 0
