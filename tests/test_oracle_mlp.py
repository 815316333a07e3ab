"""Pins of the learned C-indicator's CPU oracle (oracle/mlp.py, features3d).

- the float64 gradient of the network loss against central finite
  differences (a dropped term, a transposed operand or a wrong activation
  derivative fails it);
- GELU (tanh form) identities: gelu(x) - gelu(-x) = x, gelu(0) = 0, and the
  derivative against finite differences;
- bf16 rounding on exactly representable values and ties (round to even);
- the stencil order of the network inputs;
- the features on fields whose derivatives the difference formulas
  reproduce exactly (linear fields);
- the learned compute path: fed with the plain method's own C variables it
  reproduces the plain RHS of the flow variables bit for bit.
"""
import numpy as np

import oracle
from oracle import mlp
from fiesta_amd.config import FST_DVAR, FST_RHO, FST_VAR, Dims, make_config


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, scale, mlp.NPARAM)
    w[mlp.OFF_W1:mlp.OFF_B1] *= 0.2
    return w


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = _params(1)
    X = rng.random((6, mlp.NIN)) * 0.5
    T = rng.random(6)
    loss, g, _ = mlp.loss_grad(w, X, T)
    idx = np.concatenate([rng.choice(mlp.OFF_B1, 60, replace=False),
                          np.arange(mlp.OFF_B1, mlp.OFF_B1 + 8), rng.choice(np.arange(mlp.OFF_W2, mlp.OFF_B2), 30),
                          np.arange(mlp.OFF_B2, mlp.OFF_B2 + 8), rng.choice(np.arange(mlp.OFF_W3, mlp.OFF_B3), 30),
                          np.arange(mlp.OFF_B3, mlp.OFF_B3 + 8), np.arange(mlp.OFF_W4, mlp.NPARAM)])
    h = 1e-6
    for q in idx:
        wp, wm = w.copy(), w.copy()
        wp[q] += h
        wm[q] -= h
        fd = (mlp.loss_grad(wp, X, T)[0] - mlp.loss_grad(wm, X, T)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])


def test_gelu_identities():
    x = np.linspace(-6, 6, 241)
    assert np.allclose(mlp.gelu(x) - mlp.gelu(-x), x, rtol=0, atol=1e-14)
    assert mlp.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mlp.gelu(np.array([20.0]))[0] - 20.0) < 1e-12
    h = 1e-6
    fd = (mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h)
    assert np.allclose(fd, mlp.gelu_grad(x), rtol=0, atol=1e-8)


def test_bf16_rounding():
    exact = np.array([0.0, 1.0, -2.5, 0.15625, 3.0 * 2.0 ** -20], np.float32)
    assert np.array_equal(mlp.bf16(exact), exact.astype(np.float64))
    one = np.float32(1.0)
    tie_even = np.float32(1.0 + 2.0 ** -8)           # halfway between 1 and 1 + 2^-7
    tie_odd = np.float32(1.0 + 3.0 * 2.0 ** -8)      # halfway between 1 + 2^-7 and 1 + 2^-6
    above = np.float32(1.0 + 2.0 ** -8 + 2.0 ** -20)
    r = mlp.bf16(np.array([tie_even, tie_odd, above]))
    assert r[0] == one and r[1] == 1.0 + 2.0 ** -6 and r[2] == 1.0 + 2.0 ** -7


def test_stencil_order():
    cfg = make_config(ndim=3, ns=1, nci=6, ncj=6, nck=6)
    d = Dims(cfg)
    F = np.zeros((3, d.ngk, d.ngj, d.ngi))
    k, j, i = np.meshgrid(np.arange(d.ngk), np.arange(d.ngj), np.arange(d.ngi), indexing="ij")
    for c in range(3):
        F[c] = c * 1e6 + i + 100 * j + 10000 * k
    cell = 4 + d.ngi * (5 + d.ngj * 6)
    X = mlp.stencils(F, np.array([cell]), d)[0]
    # input q = ch*125 + (dz+2)*25 + (dy+2)*5 + (dx+2)
    for ch, dz, dy, dx in ((0, -2, -2, -2), (1, 0, 1, -1), (2, 2, 2, 2), (0, 1, -2, 2)):
        q = ch * 125 + (dz + 2) * 25 + (dy + 2) * 5 + (dx + 2)
        assert X[q] == ch * 1e6 + (4 + dx) + 100 * (5 + dy) + 10000 * (6 + dz)


def _cfg(**kw):
    base = dict(ndim=3, ns=2, nci=10, ncj=9, nck=8, dx=0.01, dy=0.012, dz=0.011, gamma=(1.4, 1.67),
                M=(0.02896, 0.004), mu=(1.8e-5, 2.0e-5), visc=True, ceq=True, kap=10.0, eps=5.0, alpha=20.0,
                bc=("outflow",) * 6)
    base.update(kw)
    return make_config(**base)


def _state(cfg, rho, u, p, Y0):
    """Conserved variables from primitive fields (k,j,i) arrays."""
    d = Dims(cfg)
    Ru = cfg.R
    var = np.zeros(d.shape(d.nvt))
    cp = Y0 * cfg.gamma[0] * (Ru / cfg.M[0]) / (cfg.gamma[0] - 1) + (1 - Y0) * cfg.gamma[1] * (Ru / cfg.M[1]) / (cfg.gamma[1] - 1)
    cv = Y0 * (Ru / cfg.M[0]) / (cfg.gamma[0] - 1) + (1 - Y0) * (Ru / cfg.M[1]) / (cfg.gamma[1] - 1)
    gam = cp / cv
    for a in range(3):
        var[a] = rho * u[a]
    var[3] = p / (gam - 1) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    var[4] = rho * Y0
    var[5] = rho * (1 - Y0)
    return var, gam


def test_features_on_linear_fields():
    cfg = _cfg(learned_c=True)
    d = Dims(cfg)
    k, j, i = np.meshgrid(np.arange(d.ngk), np.arange(d.ngj), np.arange(d.ngi), indexing="ij")
    x, y, z = i * cfg.dx, j * cfg.dy, k * cfg.dz
    a = 3.0
    rho = 1.0 + a * x                       # linear density
    p = np.full_like(rho, 1.0e5)
    u = [-40.0 * x, np.zeros_like(x), np.zeros_like(x)]   # uniform compression du/dx = -40
    var, gam = _state(cfg, rho, u, p, np.ones_like(rho))
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    F, M, Dr, G = o.features()
    hmin = min(cfg.dx, cfg.dy, cfg.dz)
    sl = (slice(2, -2),) * 3
    c = np.sqrt(gam * p / rho)
    assert np.allclose(F[0][sl], (a * hmin / rho)[sl], rtol=1e-6, atol=0)
    assert np.allclose(F[1][sl], 0.0, atol=1e-7)
    assert np.allclose(F[2][sl], (-40.0 * hmin / c)[sl], rtol=1e-6, atol=0)
    assert np.allclose(Dr[0][sl], a, rtol=1e-6) and np.allclose(Dr[1][sl], 0, atol=1e-6)
    assert np.allclose(M[sl], np.maximum(np.abs(F[0]), np.abs(F[2]))[sl], rtol=0, atol=0)
    # zero outside the difference stencils' reach
    assert np.all(F[:, 0] == 0) and np.all(Dr[:, :, 1] == 0) and np.all(G[:, :, -2] == 0)
    # contact switch: specific internal energy falls where density rises (p
    # uniform) -> dnednr < 0 -> G = |grad rho| h / rho
    cfg2 = _cfg(learned_c=True)
    var2, _ = _state(cfg2, rho, [np.zeros_like(x)] * 3, p, np.ones_like(rho))
    o2 = oracle.Oracle(cfg2)
    o2.set(FST_VAR, var2)
    _, _, Dr2, G2 = o2.features()
    assert np.allclose(G2[sl], (np.abs(Dr2[0]) * hmin / rho)[sl], rtol=1e-6)


def test_learned_compute_reproduces_plain_terms():
    """With C_hat, n := the plain method's C variables (after the ghost fill)
    and max C_hat := theirs, the learned compute equals the plain compute on
    the flow variables bit for bit (only the relaxation solve is replaced)."""
    cfg = _cfg(bc=("outflow", "reflective", "hydrostatic", "outflow", "noslip", "outflow"), buoyancy=True,
               g=9.81)
    d = Dims(cfg)
    rng = np.random.default_rng(11)
    k, j, i = np.meshgrid(np.arange(d.ngk), np.arange(d.ngj), np.arange(d.ngi), indexing="ij")
    rho = 1.0 + 0.3 * np.tanh((i - 8) / 1.5) + 0.01 * rng.random(i.shape)
    u = [5 * rng.standard_normal(i.shape) for _ in range(3)]
    p = 1e5 * (1 + 0.05 * rng.random(i.shape))
    var, _ = _state(cfg, rho, u, p, 0.5 + 0.4 * np.tanh((j - 6) / 2.0))
    var[6:] = 50 * rng.random(var[6:].shape)  # C variables
    plain = oracle.Oracle(cfg)
    plain.set(FST_VAR, var)
    plain.compute()
    want = plain.get(FST_DVAR)[:d.nv]
    cvars = plain.get(FST_VAR)  # after the ghost fill compute() applied
    _, maxCh, _ = plain.maxima()
    lcfg = _cfg(bc=("outflow", "reflective", "hydrostatic", "outflow", "noslip", "outflow"), buoyancy=True,
                g=9.81, learned_c=True)
    lrn = oracle.Oracle(lcfg)
    lrn.set(FST_VAR, var[:d.nv])
    lrn.set_learned(cvars[d.nv + 1], cvars[d.nv + 2:d.nv + 5], maxCh)
    lrn.compute()
    got = lrn.get(FST_DVAR)
    assert got.shape[0] == d.nv
    assert np.array_equal(got, want)
    # and a different C_hat changes the momentum RHS (the term is live)
    lrn.set_learned(2 * cvars[d.nv + 1], cvars[d.nv + 2:d.nv + 5], maxCh)
    lrn.compute()
    assert not np.array_equal(lrn.get(FST_DVAR)[:3], want[:3])


def test_residual_check_rests_on_the_steady_relaxation():
    """A field that satisfies the resting C-equation relaxation exactly
    (G = C - a lap(C), G built from C) passes the check, and the fallback
    value of every cell of it is the field itself (the local solution of the
    resting equation); a spike planted at one cell fails there."""
    from fiesta_amd.config import Dims, make_config
    cfg = make_config(ndim=3, ns=1, nci=10, ncj=9, nck=8, dx=0.01, dy=0.012, dz=0.009, kap=10.0, eps=5.0)
    d = Dims(cfg)
    rng = np.random.default_rng(4)
    C = rng.uniform(0.5, 2.0, d.shape(1)[1:])
    h = min(cfg.dx, cfg.dy, cfg.dz)
    dd = (cfg.dx, cfg.dy, cfg.dz)
    a = cfg.kap * cfg.eps * sum(x * x for x in dd)
    lap = np.zeros_like(C)
    for ax, hh in zip((2, 1, 0), dd):
        lap += (np.roll(C, 1, ax) - 2 * C + np.roll(C, -1, ax)) / hh
    rho = rng.uniform(1.0, 1.5, C.shape)
    G = (C - a * lap) * h / rho            # dimensionless source G1
    g = d.ng
    inner = np.zeros(C.shape, bool)
    inner[g:-g, g:-g, g:-g] = True
    cells = np.flatnonzero(inner)
    C32 = C.astype(np.float32)
    G = np.where(inner, G, 0.0)
    # float32 storage of C and G leaves a residual of a few ulps of the terms
    v, fb, rn, thr = mlp.residual_check(C32, G.astype(np.float32), rho, cells, d, h, cfg.kap, cfg.eps, dd, 1e-4)
    assert not fb.any()
    v0, fb0, _, _ = mlp.residual_check(C32, G.astype(np.float32), rho, cells, d, h, cfg.kap, cfg.eps, dd, 0.0)
    assert fb0.all()
    np.testing.assert_allclose(v0, C32.ravel()[cells], rtol=2e-5)
    spike = C32.copy()
    c0 = cells[len(cells) // 2]
    spike.ravel()[c0] *= 1.5
    _, fbs, _, _ = mlp.residual_check(spike, G.astype(np.float32), rho, cells, d, h, cfg.kap, cfg.eps, dd, 1e-3)
    assert fbs[list(cells).index(c0)]


# ---------------------------------------------------------------- 2D network and fields


def test_2d_layout_and_gradient():
    """The 2D network: 75 inputs, the same block order; its float64 gradient
    against central finite differences."""
    L = mlp.layout(2)
    assert L["NIN"] == 75 and L["NPARAM"] == mlp.NPARAM2 == 75 * 64 + 64 + 2 * (64 * 64 + 64) + 64 + 1
    rng = np.random.default_rng(8)
    w = rng.normal(0.0, 0.3, mlp.NPARAM2)
    w[L["OFF_W1"]:L["OFF_B1"]] *= 0.4
    X = rng.random((5, 75)) * 0.5
    T = rng.random(5)
    loss, g, _ = mlp.loss_grad(w, X, T)
    idx = np.concatenate([rng.choice(L["OFF_B1"], 40, replace=False), np.arange(L["OFF_B1"], L["OFF_B1"] + 4),
                          rng.choice(np.arange(L["OFF_W2"], L["OFF_B2"]), 20),
                          rng.choice(np.arange(L["OFF_W3"], L["OFF_B3"]), 20), np.arange(L["OFF_W4"], L["NPARAM"])])
    h = 1e-6
    for q in idx:
        wp, wm = w.copy(), w.copy()
        wp[q] += h
        wm[q] -= h
        fd = (mlp.loss_grad(wp, X, T)[0] - mlp.loss_grad(wm, X, T)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])


def test_stencil_order_2d():
    cfg = make_config(ndim=2, ns=1, nci=7, ncj=6)
    d = Dims(cfg)
    F = np.zeros((3, 1, d.ngj, d.ngi))
    j, i = np.meshgrid(np.arange(d.ngj), np.arange(d.ngi), indexing="ij")
    for c in range(3):
        F[c, 0] = c * 1e6 + i + 100 * j
    cell = 5 + d.ngi * 4
    X = mlp.stencils(F, np.array([cell]), d)[0]
    assert X.size == 75
    # input q = ch*25 + (dy+2)*5 + (dx+2)
    for ch, dy, dx in ((0, -2, -2), (1, 1, -1), (2, 2, 2), (0, -2, 2), (2, 0, 0)):
        q = ch * 25 + (dy + 2) * 5 + (dx + 2)
        assert X[q] == ch * 1e6 + (5 + dx) + 100 * (4 + dy)


def test_active_cells_2d_brute_force():
    cfg = make_config(ndim=2, ns=1, nci=9, ncj=7)
    d = Dims(cfg)
    rng = np.random.default_rng(2)
    M = np.where(rng.random((1, d.ngj, d.ngi)) < 0.08, 1.0, 0.0).astype(np.float32)
    got = mlp.active_cells(M, d, 0.5)
    want = []
    g = d.ng
    for jj in range(g, g + d.ncj):
        for ii in range(g, g + d.nci):
            if M[0, jj - 1:jj + 2, ii - 1:ii + 2].max() > 0.5:
                want.append(ii + d.ngi * jj)
    assert list(got) == want and len(want) > 0


def _cfg2(**kw):
    base = dict(ndim=2, ns=2, nci=12, ncj=10, dx=0.01, dy=0.013, gamma=(1.4, 1.67), M=(0.02896, 0.004),
                mu=(1.8e-5, 2.0e-5), visc=True, ceq=True, kap=10.0, eps=5.0, alpha=20.0, bc=("outflow",) * 6)
    base.update(kw)
    return make_config(**base)


def _state2d(cfg, rho, u, p, Y0):
    d = Dims(cfg)
    Ru = cfg.R
    var = np.zeros(d.shape(d.nvt))
    cp = Y0 * cfg.gamma[0] * (Ru / cfg.M[0]) / (cfg.gamma[0] - 1) + (1 - Y0) * cfg.gamma[1] * (Ru / cfg.M[1]) / (cfg.gamma[1] - 1)
    cv = Y0 * (Ru / cfg.M[0]) / (cfg.gamma[0] - 1) + (1 - Y0) * (Ru / cfg.M[1]) / (cfg.gamma[1] - 1)
    gam = cp / cv
    var[0] = rho * u[0]
    var[1] = rho * u[1]
    var[2] = p / (gam - 1) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2)
    var[3] = rho * Y0
    var[4] = rho * (1 - Y0)
    return var, gam


def test_features2d_on_linear_fields():
    """Linear density along y, compression along y: the difference formulas
    are exact, so |grad rho|, the tangent (-d rho/dy, d rho/dx) and div u
    have closed forms; a swapped direction or velocity index fails here."""
    cfg = _cfg2(learned_c=True)
    d = Dims(cfg)
    j, i = np.meshgrid(np.arange(d.ngj), np.arange(d.ngi), indexing="ij")
    x, y = (i * cfg.dx)[None], (j * cfg.dy)[None]
    a, b = 2.0, 5.0
    rho = 1.0 + a * x + b * y
    p = np.full_like(rho, 1.0e5)
    u = [np.zeros_like(x), -30.0 * y]  # dv/dy = -30, du/dx = 0
    var, gam = _state2d(cfg, rho, u, p, np.full_like(rho, 0.3))
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    F, M, Dr, G = o.features()
    h = min(cfg.dx, cfg.dy)
    sl = (slice(None), slice(2, -2), slice(2, -2))
    c = np.sqrt(gam * p / rho)
    assert np.allclose(F[0][sl], (np.hypot(a, b) * h / rho)[sl], rtol=1e-6)
    assert np.allclose(F[1][sl], 0.0, atol=1e-7)
    assert np.allclose(F[2][sl], (-30.0 * h / c)[sl], rtol=1e-6)
    assert np.allclose(Dr[0][sl], -b, rtol=1e-6) and np.allclose(Dr[1][sl], a, rtol=1e-6)
    assert np.all(Dr[2] == 0)
    # zero outside the stencils' reach
    assert np.all(F[:, 0, 0] == 0) and np.all(Dr[:, 0, :, 1] == 0) and np.all(G[0, -2] == 0)
    # at rest with uniform p: nrg = E / rho falls where rho rises -> G = |grad rho| h / rho
    var2, _ = _state2d(cfg, rho, [np.zeros_like(x)] * 2, p, np.full_like(rho, 0.3))
    o2 = oracle.Oracle(cfg)
    o2.set(FST_VAR, var2)
    _, _, _, G2 = o2.features()
    assert np.allclose(G2[sl], (np.hypot(a, b) * h / rho)[sl], rtol=1e-6)
    # rho rising with nrg rising (p growing faster than rho) -> no contact source
    var3, _ = _state2d(cfg, rho, [np.zeros_like(x)] * 2, p * rho ** 2, np.full_like(rho, 0.3))
    o3 = oracle.Oracle(cfg)
    o3.set(FST_VAR, var3)
    assert np.all(o3.features()[3] == 0)


def test_learned_compute_2d_reproduces_plain_terms():
    """2D: with C_hat and the tangent := the plain method's var(nv+1..nv+3)
    and max C_hat := theirs, the learned compute equals the plain compute on
    the flow variables bit for bit; the C_hat term is live."""
    cfg = _cfg2(bc=("outflow", "reflective", "outflow", "outflow", "outflow", "outflow"))
    d = Dims(cfg)
    rng = np.random.default_rng(5)
    j, i = np.meshgrid(np.arange(d.ngj), np.arange(d.ngi), indexing="ij")
    rho = (1.0 + 0.3 * np.tanh((i - 8) / 1.5) + 0.01 * rng.random(i.shape))[None]
    u = [5 * rng.standard_normal(rho.shape) for _ in range(2)]
    p = 1e5 * (1 + 0.05 * rng.random(rho.shape))
    var, _ = _state2d(cfg, rho, u, p, 0.5 + 0.4 * np.tanh((j[None] - 6) / 2.0))
    var[d.nv:] = 50 * rng.standard_normal(var[d.nv:].shape)
    plain = oracle.Oracle(cfg)
    plain.set(FST_VAR, var)
    plain.apply_bcs()
    cvars = plain.get(FST_VAR)
    plain.compute()
    want = plain.get(FST_DVAR)[:d.nv]
    _, maxCh, _ = plain.maxima()
    lcfg = _cfg2(bc=("outflow", "reflective", "outflow", "outflow", "outflow", "outflow"), learned_c=True)
    lrn = oracle.Oracle(lcfg)
    lrn.set(FST_VAR, cvars[:d.nv])
    tang = np.concatenate([cvars[d.nv + 2:d.nv + 4], np.zeros_like(cvars[:1])])
    lrn.set_learned(cvars[d.nv + 1], tang, maxCh)
    lrn.compute()
    got = lrn.get(FST_DVAR)
    assert got.shape[0] == d.nv
    assert np.array_equal(got, want)
    lrn.set_learned(cvars[d.nv + 1], tang[[1, 0, 2]], maxCh)  # swapped tangent components
    lrn.compute()
    assert not np.array_equal(lrn.get(FST_DVAR)[:2], want[:2])


def test_residual_check_2d_rests_on_the_steady_relaxation():
    """2D (updateCeq2D, src/ceq.hpp:153-195): G built from C by the resting
    relaxation over x and y passes; the fallback of every cell is C itself;
    a z term (the 3D Laplacian) would not rest."""
    cfg = make_config(ndim=2, ns=1, nci=11, ncj=9, dx=0.01, dy=0.014, kap=10.0, eps=5.0)
    d = Dims(cfg)
    rng = np.random.default_rng(6)
    C = rng.uniform(0.5, 2.0, (1, d.ngj, d.ngi))
    h = min(cfg.dx, cfg.dy)
    dd = (cfg.dx, cfg.dy)
    a = cfg.kap * cfg.eps * (cfg.dx ** 2 + cfg.dy ** 2)
    lap = np.zeros_like(C)
    for ax, hh in zip((2, 1), dd):
        lap += (np.roll(C, 1, ax) - 2 * C + np.roll(C, -1, ax)) / hh
    rho = rng.uniform(1.0, 1.5, C.shape)
    G = (C - a * lap) * h / rho
    g = d.ng
    inner = np.zeros(C.shape, bool)
    inner[0, g:-g, g:-g] = True
    cells = np.flatnonzero(inner)
    C32 = C.astype(np.float32)
    G = np.where(inner, G, 0.0).astype(np.float32)
    v, fb, rn, thr = mlp.residual_check(C32, G, rho, cells, d, h, cfg.kap, cfg.eps, dd, 1e-4)
    assert not fb.any()
    v0, fb0, _, _ = mlp.residual_check(C32, G, rho, cells, d, h, cfg.kap, cfg.eps, dd, 0.0)
    assert fb0.all()
    np.testing.assert_allclose(v0, C32.ravel()[cells], rtol=2e-5)
    # the same G judged with a 3D spacing (an added dz^2 in a) does not rest
    _, fb3, _, _ = mlp.residual_check(C32, G, rho, cells, d, h, cfg.kap, cfg.eps * 1.5, dd, 1e-4)
    assert fb3.any()
