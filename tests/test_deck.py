"""The Lua input deck reader (include/fiesta_deck.h), on the CPU.

- the interpreter against values Lua's semantics fix (operator precedence and
  associativity, floor division and modulo signs, closures, loops, generic
  for, methods, string.format, coercions, multiple results);
- errors: syntax and run-time errors of the script, missing required keys,
  bad values, with the reference's messages where it has them;
- executeConfiguration (src/input.cpp:130-443): every key of a deck mapped
  onto fst_config and the run settings, the defaults, the nt / tend rule,
  the 2D defaults, ioviews (getIOBlock, src/luaReader.cpp:208-312);
- loadInitialConditions / loadGrid (src/input.cpp:445-557): cell centres,
  variables v < nv only, generalized-grid node averages, against the same
  closed forms evaluated with numpy; single- and multi-threaded evaluation
  agree;
- the reference's own sample decks in the current format, when the
  reference tree is present.
"""
import os
import textwrap

import numpy as np
import pytest

from fiesta_amd.config import BC, Dims, FST_FP64_FAST
from fiesta_amd.deck import Deck, DeckError


def _write(tmp_path, text, name="fiesta.lua"):
    p = tmp_path / name
    p.write_text(textwrap.dedent(text))
    return str(p)


SEMANTICS = """
fiesta.title = "semantics"
local a, b = 3, 4
local t = {10, 20, 30, n = "v", [5] = 50}
local function fact(n) if n <= 1 then return 1 else return n * fact(n - 1) end end
local function counter() local k = 0; return function() k = k + 1; return k end end
local ct = counter(); ct(); ct()
local sum = 0; for i = 1, 10 do sum = sum + i end
local sum2 = 0; for i = 10, 1, -2 do sum2 = sum2 + i end
local w = 0; while w < 5 do w = w + 1; if w == 3 then break end end
local r = 0; repeat local z = r + 1; r = z until z >= 4
local keys = 0; for k, v in pairs({x = 1, y = 2, 3}) do keys = keys + 1 end
local isum = 0; for i, v in ipairs({5, 6, 7}) do isum = isum + i * v end
local obj = {v = 5}; function obj:get(x) return self.v + x end
local m1, m2 = (function() return 1, 2 end)()
local packed = {(function() return 1, 2, 3 end)()}
local nested = {list = {1, 2, {3, 4}}}
--[[ a long
     comment ]]
local long = [[line1
line2]]
fiesta.r = {
  c = a^2 + b^2, d = 2^3^2, e = -2^2, f = 7 // 2, g = -7 % 3, h = 7 % -3, k = 1 + 2 * 3 - 4 / 2,
  s = "x" .. 1 .. "y", len = #t, tn = t.n, t5 = t[5], t1 = t[1.0],
  cmp = (1 < 2) and "yes" or "no", nilor = nil or 7, andf = false and 1, notnil = not nil,
  fa = fact(10), cnt = ct(), sum = sum, sum2 = sum2, w = w, r = r, keys = keys, isum = isum,
  fmt = string.format("%05.2f|%d|%s|%e|%x", 3.14159, 42, "ok", 1234.5, 255),
  mx = math.max(3, 9, 2), mn = math.min(3, 9, 2), pi = math.pi, tanh1 = math.tanh(1), sq = math.sqrt(2),
  absn = math.abs(-3.5), fl = math.floor(-2.5), ce = math.ceil(2.1), ex = math.exp(1), lg = math.log(8, 2),
  mval = obj:get(2), m1 = m1, m2 = m2, npacked = #packed, deep = nested.list[3][2],
  ty = type(nested), strnum = "10" + 5, tostr = tostring(12) .. tostring(true), tonum = tonumber("0x10"),
  longlen = #long, concat_num = 1.5 .. "", eqt = ({} == {}), streq = ("ab" == "a" .. "b"), lt = "a" < "b",
}
function fiesta.f3(x, y, z, v) return x * 100 + y * 10 + z + v / 10 end
function fiesta.reads_global(x, y, z, v) return fiesta.r.c * x end
"""


def test_interpreter_semantics(tmp_path):
    d = Deck(_write(tmp_path, SEMANTICS))
    g = lambda k: d.get("r." + k)  # noqa: E731
    assert g("c") == 25 and g("d") == 512 and g("e") == -4 and g("k") == 5
    assert g("f") == 3 and g("g") == 2 and g("h") == -2          # floor division, modulo sign of the divisor
    assert g("s") == "x1y" and g("len") == 3 and g("tn") == "v" and g("t5") == 50 and g("t1") == 10
    assert g("cmp") == "yes" and g("nilor") == 7 and g("andf") is False and g("notnil") is True
    assert g("fa") == 3628800 and g("cnt") == 3                   # recursion, closures with their own upvalue
    assert g("sum") == 55 and g("sum2") == 30 and g("w") == 3 and g("r") == 4
    assert g("keys") == 3 and g("isum") == 38
    assert g("fmt") == "03.14|42|ok|1.234500e+03|ff"
    assert g("mx") == 9 and g("mn") == 2 and g("pi") == np.pi and g("tanh1") == np.tanh(1.0)
    assert g("sq") == np.sqrt(2.0) and g("absn") == 3.5 and g("fl") == -3 and g("ce") == 3
    assert g("ex") == np.exp(1.0) and g("lg") == pytest.approx(3.0, abs=1e-15)
    assert g("mval") == 7 and g("m1") == 1 and g("m2") == 2 and g("npacked") == 3 and g("deep") == 4
    assert g("ty") == "table" and g("strnum") == 15 and g("tostr") == "12true" and g("tonum") == 16
    assert g("longlen") == 11 and g("concat_num") == "1.5" and g("eqt") is False and g("streq") is True
    assert g("lt") is True
    assert d.call("f3", 1, 2, 3, 4) == 100 + 20 + 3 + 0.4
    assert d.call("reads_global", 2, 0, 0, 0) == 50
    assert d.get("r.missing") is None and d.get("f3") == "function" and d.get("r") == "table"


@pytest.mark.parametrize("src,code,msg", [
    ("fiesta.title = \n", 3, "unexpected symbol"),
    ("local x = (1 + 2\n", 3, "')' expected"),
    ("nothing()\n", 3, "attempt to call a nil value"),
    ("local t = nil; local y = t.x\n", 3, "attempt to index a nil value"),
    ("local y = {} + 1\n", 3, "arithmetic on a table value"),
    ("goto done\n", 3, "goto and labels are not supported"),
    ("error('custom failure')\n", 3, "custom failure"),
    ("local s = 'unfinished\n", 3, "unfinished string"),
])
def test_script_errors(tmp_path, src, code, msg):
    with pytest.raises(DeckError) as e:
        Deck(_write(tmp_path, src))
    assert e.value.code == code


def test_script_error_messages_name_the_line(tmp_path):
    from fiesta_amd.deck import lib
    import ctypes as C
    h = C.c_void_p()
    p = _write(tmp_path, "fiesta.title = 'x'\n\nlocal y = nil + 1\n")
    assert lib().fdk_open(p.encode(), C.byref(h)) == 3


BASE = """
fiesta.title = "base"
fiesta.species = {{name = "Air", gamma = 1.4, M = 0.02896, mu = 1.8e-5}}
fiesta.ioviews = {{name = "sol", path = "./sol", frequency = 10}}
fiesta.time.dt = 1e-6
fiesta.time.nt = 20
fiesta.grid.ndim = 3
fiesta.grid.ni = {8, 6, 4}
fiesta.grid.dx = {0.1, 0.2, 0.3}
fiesta.mpi.procs = {1, 1, 1}
fiesta.bc.xperiodic = true
fiesta.bc.ymin = "hydrostatic"
fiesta.bc.ymax = "reflective"
fiesta.bc.zmin = "noslip"
fiesta.bc.zmax = "outflow"
function fiesta.initial_conditions(x, y, z, v) return 0 end
"""


def test_config_mapping_and_defaults(tmp_path):
    d = Deck(_write(tmp_path, BASE))
    c, r = d.config()
    assert (c.ndim, c.nci, c.ncj, c.nck, c.ns, c.ng) == (3, 8, 6, 4, 1, 3)
    assert (c.dx, c.dy, c.dz, c.dt, c.R) == (0.1, 0.2, 0.3, 1e-6, 8.314462)
    assert (c.xPer, c.yPer, c.zPer) == (1, 0, 0)
    assert (c.bcB, c.bcT, c.bcH, c.bcF) == (BC["hydrostatic"], BC["reflective"], BC["noslip"], BC["outflow"])
    assert (c.visc, c.ceq, c.buoyancy, c.noise, c.scheme, c.grid) == (0, 0, 0, 0, 1, 0)
    assert (c.gamma[0], c.M[0], c.mu[0]) == (1.4, 0.02896, 1.8e-5)
    assert c.precision == 0 and c.rk_order == 2 and c.learned_c == 0
    assert r["title"] == "base" and r["metadata"] == "none" and r["nt"] == 20 and r["tend"] == 20
    assert r["tstart"] == 0 and r["tinterval"] == 1 and r["stat_freq"] == 0 and r["restart"] == 0
    assert r["restart_name"] == "restart-0000000.h5" and r["restart_path"] == "." and r["mpi_type"] == "host"
    assert d.species() == ["Air"]
    v = d.ioviews(r["n_ioviews"])
    assert v == [dict(name="sol", path="./sol", frequency=10, average=1, whole=1, start=[0, 0, 0],
                      limit=[7, 5, 3], stride=[1, 1, 1])]


def test_config_physics_keys(tmp_path):
    deck = BASE.replace('fiesta.time.nt = 20', 'fiesta.time.tend = 50\nfiesta.time.start_index = 10') + """
fiesta.viscosity.enabled = true
fiesta.buoyancy.enabled = true
fiesta.buoyancy.rho_reference = 0.5
fiesta.ceq.enabled = true
fiesta.ceq.epsilon = 5; fiesta.ceq.alpha = 20; fiesta.ceq.beta = 0; fiesta.ceq.betae = 1
fiesta.noise.enabled = true
fiesta.noise.dh = 1e-5; fiesta.noise.eta = 1e-3; fiesta.noise.coff = 0.05; fiesta.noise.mode = 2
fiesta.precision = "fp64_fast"
fiesta.time.integrator = "ssp-rk3"
fiesta.status.frequency = 5
fiesta.ioviews = {{name = "sol", frequency = 10},
                  {name = "box", path = "./box", frequency = 2, average = 0,
                   start = {2, 1, 1}, limit = {7, 5, 3}, stride = {2, 2, 1}}}
fiesta.species = {{name = "air", gamma = 1.4, M = 0.02896, mu = 1.8e-5},
                  {name = "he", gamma = 1.67, M = 0.004, mu = 2e-5}}
"""
    d = Deck(_write(tmp_path, deck))
    c, r = d.config()
    assert (c.visc, c.buoyancy, c.gAccel, c.rhoRef) == (1, 1, 9.81, 0.5)  # acceleration defaults to 9.81
    assert (c.ceq, c.kap, c.eps, c.alpha) == (1, 1.23456, 5.0, 20.0)       # kappa defaults to 1.23456
    assert (r["beta"], r["betae"], r["st"]) == (0.0, 1.0, 0)
    assert (c.noise, c.n_dh, c.n_eta, c.n_coff, c.n_nt, c.n_mode) == (1, 1e-5, 1e-3, 0.05, 1, 2)
    assert c.scheme == 1 and c.precision == FST_FP64_FAST and c.rk_order == 3 and c.ns == 2
    assert (r["tstart"], r["tend"], r["nt"], r["tinterval"], r["stat_freq"]) == (10, 50, 40, 0, 5)
    assert d.species() == ["air", "he"]
    v = d.ioviews(2)
    assert v[0]["path"] == "./" and v[0]["whole"] == 1
    assert v[1] == dict(name="box", path="./box", frequency=2, average=0, whole=0, start=[1, 0, 0],
                        limit=[6, 4, 2], stride=[2, 2, 1])


@pytest.mark.parametrize("scheme,num", [("weno5", 1), ("centered4", 2), ("quick", 3), ("unknown", 1)])
def test_2d_defaults_and_schemes(tmp_path, scheme, num):
    deck = BASE.replace("fiesta.grid.ndim = 3", "fiesta.grid.ndim = 2").replace("{8, 6, 4}", "{8, 6}")
    deck = deck.replace("{0.1, 0.2, 0.3}", "{0.1, 0.2}").replace("{1, 1, 1}", "{1, 1}")
    deck += 'fiesta.advection_scheme = "%s"\n' % scheme
    c, _ = Deck(_write(tmp_path, deck)).config()
    assert (c.ndim, c.nck, c.dz, c.zPer, c.bcH, c.bcF) == (2, 1, 0.1, 0, 0, 0)  # dz := dx (src/input.cpp:352)
    assert c.scheme == num  # unknown names keep weno5 (src/input.cpp:305-311)


def test_unsupported_combination_names_the_reason(tmp_path):
    d = Deck(_write(tmp_path, BASE + 'fiesta.advection_scheme = "quick"\n'))
    with pytest.raises(DeckError, match="3D supports weno5"):
        d.config()


@pytest.mark.parametrize("edit,code,msg", [
    (("fiesta.time.dt = 1e-6\n", ""), 2, "fiesta.time.dt"),
    (("fiesta.time.nt = 20\n", ""), 2, "fiesta.time.nt' or 'fiesta.time.tend"),
    (('fiesta.bc.ymin = "hydrostatic"', 'fiesta.bc.ymin = "sticky"'), 1, "invalid boundary condition type"),
    (("fiesta.mpi.procs = {1, 1, 1}\n", ""), 2, "fiesta.mpi.procs"),
    (('fiesta.ioviews = {{name = "sol", path = "./sol", frequency = 10}}\n', ""), 2, "fiesta.ioviews"),
    (('fiesta.species = {{name = "Air", gamma = 1.4, M = 0.02896, mu = 1.8e-5}}\n', ""), 2, "fiesta.species"),
    (("fiesta.grid.ni = {8, 6, 4}", "fiesta.grid.ni = {8, 6}"), 1, "needs 3 entries"),
    (('fiesta.title = "base"', 'fiesta.title = nil'), 2, "fiesta.title"),
])
def test_config_errors(tmp_path, edit, code, msg):
    deck = BASE.replace(*edit)
    assert deck != BASE
    d = Deck(_write(tmp_path, deck))
    with pytest.raises(DeckError) as e:
        d.config()
    assert e.value.code == code and msg in str(e.value), str(e.value)


IC = BASE + """
local rho0, p0, g = 1.2, 1.0e5, 1.4
function fiesta.initial_conditions(x, y, z, v)
  local rho = rho0 * (1 + 0.1 * math.sin(2 * math.pi * x) * math.cos(y))
  local u = 10 * z
  if v == 0 then return rho * u end
  if v == 1 then return -rho * x end
  if v == 2 then return 0 end
  if v == 3 then return p0 / (g - 1) + 0.5 * rho * (u * u + x * x) end
  if v == 4 then return rho end
  return 99  -- never asked: the C variables and v >= nv stay zero
end
"""


@pytest.mark.parametrize("ceq", [False, True])
def test_initial_conditions_at_cell_centres(tmp_path, ceq):
    deck = IC + ("\nfiesta.ceq.enabled = true; fiesta.ceq.epsilon = 1; fiesta.ceq.alpha = 1;"
                 " fiesta.ceq.beta = 0; fiesta.ceq.betae = 0\n" if ceq else "")
    d = Deck(_write(tmp_path, deck))
    cfg, _ = d.config()
    dm = Dims(cfg)
    var = d.initial_conditions(cfg, threads=1)
    assert np.array_equal(var, d.initial_conditions(cfg, threads=4))
    x = (np.arange(cfg.nci) * cfg.dx + 0.5 * cfg.dx)[None, None, :]
    y = (np.arange(cfg.ncj) * cfg.dy + 0.5 * cfg.dy)[None, :, None]
    z = (np.arange(cfg.nck) * cfg.dz + 0.5 * cfg.dz)[:, None, None]
    rho = 1.2 * (1 + 0.1 * np.sin(2 * np.pi * x) * np.cos(y))
    u = 10 * z
    want = [rho * u, -rho * x + 0 * z, 0 * rho, 1.0e5 / 0.4 + 0.5 * rho * (u * u + x * x), rho + 0 * z]
    sl = dm.interior()[1:]
    for v in range(dm.nv):
        np.testing.assert_allclose(var[v][sl], np.broadcast_to(want[v], var[v][sl].shape), rtol=1e-15, atol=1e-12)
    assert dm.nvt == dm.nv + (5 if ceq else 0)
    assert not var[dm.nv:].any()                        # C-equation variables zero
    ghost = np.ones(var.shape[1:], bool)
    ghost[sl] = False
    assert not var[:, ghost].any()                      # ghosts zero


GEN = """
fiesta.title = "gen"
fiesta.species = {{name = "Air", gamma = 1.4, M = 0.02896, mu = 0}}
fiesta.ioviews = {{name = "sol", frequency = 0}}
fiesta.time.dt = 1e-7
fiesta.time.nt = 2
fiesta.grid.type = "generalized"
fiesta.grid.ndim = 2
fiesta.grid.ni = {10, 6}
fiesta.grid.dx = {1, 1}
fiesta.mpi.procs = {1, 1}
fiesta.bc.xmin = "outflow"; fiesta.bc.xmax = "outflow"; fiesta.bc.ymin = "reflective"; fiesta.bc.ymax = "outflow"
local Lx, Ly, ni, nj = 10.0, 6.0, 10, 6
local function xn(i, j) return i * Lx / ni end
function fiesta.initialize_grid(i, j, k, v)
  if v == 0 then return xn(i, j) end
  local dh = 0.3 * math.sin(xn(i, j) * 2.0)
  return dh + (Ly - dh) * (j * Ly / nj) / Ly
end
function fiesta.initial_conditions(x, y, z, v)
  if v == 2 then return 2.5e5 + x end
  if v == 3 then return 1.2 + 0.01 * y end
  return 0
end
"""


def test_generalized_grid_and_node_average_centres(tmp_path):
    d = Deck(_write(tmp_path, GEN))
    cfg, _ = d.config()
    assert (cfg.grid, cfg.dx, cfg.dy) == (1, 1.0, 1.0)
    g = d.grid(cfg, threads=3)
    i = np.arange(11)[None, :]
    j = np.arange(7)[:, None]
    xn = i * 10.0 / 10 + 0 * j
    dh = 0.3 * np.sin(xn * 2.0)
    yn = dh + (6.0 - dh) * (j * 6.0 / 6) / 6.0
    np.testing.assert_allclose(g[0, 0], xn, rtol=0, atol=1e-15)
    np.testing.assert_allclose(g[1, 0], yn, rtol=0, atol=1e-15)
    assert not g[2].any()
    var = d.initial_conditions(cfg, grid=g)
    xc = (xn[:-1, :-1] + xn[1:, :-1] + xn[:-1, 1:] + xn[1:, 1:]) / 4.0
    yc = (yn[:-1, :-1] + yn[1:, :-1] + yn[:-1, 1:] + yn[1:, 1:]) / 4.0
    dm = Dims(cfg)
    sl = dm.interior()[1:]
    np.testing.assert_allclose(var[2][sl][0], 2.5e5 + xc, rtol=1e-15)
    np.testing.assert_allclose(var[3][sl][0], 1.2 + 0.01 * yc, rtol=1e-15)
    with pytest.raises(DeckError):
        d.initial_conditions(cfg, grid=None)  # a generalized grid needs its nodes


def test_callback_must_return_a_number(tmp_path):
    d = Deck(_write(tmp_path, BASE.replace("return 0 end", "return 'abc' end")))
    cfg, _ = d.config()
    with pytest.raises(DeckError, match="should return a number"):
        d.initial_conditions(cfg)


REF_SAMPLES = "/root/reference/samples"


@pytest.mark.skipif(not os.path.isdir(REF_SAMPLES), reason="the reference tree is not present")
@pytest.mark.parametrize("name", ["rt3d", "3D_Expansion", "boundary_layer", "2D_Expansion", "2dubble", "3dubble",
                                  "rt2d"])
def test_reference_sample_decks(name):
    """The reference's sample decks in the current format configure, and
    their initial conditions evaluate, without change."""
    d = Deck(os.path.join(REF_SAMPLES, name, "fiesta.lua"))
    cfg, run = d.config()
    assert run["title"] and cfg.ns >= 1 and run["n_ioviews"] >= 1
    if cfg.nci * cfg.ncj * cfg.nck <= 300000:
        var = d.initial_conditions(cfg)
        dm = Dims(cfg)
        assert np.isfinite(var).all() and var[dm.ndim + 1][dm.interior()[1:]].min() > 0  # a positive density
    if name == "rt3d":
        # the deck sets fiesta.R, but executeConfiguration reads eos.R (src/input.cpp:163): the default holds
        assert (cfg.ndim, cfg.nci, cfg.ncj, cfg.nck, cfg.R, cfg.gAccel) == (3, 10, 100, 100, 8.314462, 0.1)
        assert (cfg.xPer, cfg.bcB, cfg.bcH, cfg.ceq, cfg.kap, cfg.eps, cfg.alpha) == (1, 3, 1, 1, 10.0, 5.0, 20.0)
        assert (cfg.noise, cfg.n_dh, cfg.n_eta, cfg.n_coff) == (1, 1e-5, 1e-3, 0.05)
        assert run["nt"] == 10000 and run["stat_freq"] == 0  # the deck sets status_frequency, not status.frequency


def test_host_library_exports_every_deck_function():
    import re
    import subprocess
    from fiesta_amd import build as B
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = open(os.path.join(root, "include", "fiesta_deck.h")).read()
    declared = sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(fdk_\w+)\s*\(", src, re.M)))
    assert len(declared) == 10, declared
    nm = subprocess.run(["nm", "-D", "--defined-only", B.HOST_LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s+(\w+)$", nm, re.M))
    assert not [n for n in declared if n not in exported]
    assert os.path.exists(B.MAIN_BIN)  # the deck driver is built beside the library
