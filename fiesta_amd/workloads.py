"""Synthetic initial states for the RHS/RK path.

Initial conditions fill the interior only; ghosts stay zero until the first
ghost fill (src/input.cpp:455-511).  States are numpy arrays of shape
(nvt, ngk, ngj, ngi): the LayoutLeft (i,j,k,v) order of the device arrays.
Seeds use numpy's PCG64 so a workload is reproducible from its seed alone.
"""
import math

import numpy as np

from .config import Dims, make_config

AIR = dict(gamma=1.4, M=0.02896, mu=1.8e-5)
HOT = dict(gamma=1.67, M=0.004, mu=1.8e-5)
SF6 = dict(gamma=1.092, M=0.14606, mu=1.8e-5)
RU = 8.314462


def _mix(ns, specs):
    specs = list(specs)[:ns]
    return dict(gamma=tuple(s["gamma"] for s in specs), M=tuple(s["M"] for s in specs),
                mu=tuple(s["mu"] for s in specs))


def _cells(d):
    """Cell-centre coordinates (interior) on the unit box, shapes broadcastable to (k,j,i)."""
    x = (np.arange(d.nci) + 0.5) / d.nci
    y = (np.arange(d.ncj) + 0.5) / d.ncj
    z = (np.arange(d.nck) + 0.5) / d.nck
    return x[None, None, :], y[None, :, None], z[:, None, None]


def fill_conserved(cfg, rho_s, u, p, cvars=None, ghosts=False):
    """Pack primitive fields (interior, shape (k,j,i)) into a conserved state array.

    rho_s: list of species densities; u: list of ndim velocity components;
    p: pressure.  Energy uses the mixture gamma of src/secondary.hpp:128-141.
    ghosts=True: the fields cover every cell (shape (ngk,ngj,ngi)).
    """
    d = Dims(cfg)
    var = np.zeros(d.shape(d.nvt), dtype=np.float64)
    rho = sum(rho_s)
    Cp = 0.0
    Cv = 0.0
    for s in range(cfg.ns):
        g = cfg.gamma[s]
        Rs = cfg.R / cfg.M[s]
        Cp = Cp + (rho_s[s] / rho) * (g * Rs / (g - 1))
        Cv = Cv + (rho_s[s] / rho) * (Rs / (g - 1))
    gam = Cp / Cv
    ke = 0.5 * rho * sum(ui * ui for ui in u)
    E = p / (gam - 1) + ke
    sl = (slice(None),) * 3 if ghosts else d.interior()[1:]
    for a in range(cfg.ndim):
        var[(a,) + sl] = rho * u[a]
    var[(cfg.ndim,) + sl] = E
    for s in range(cfg.ns):
        var[(cfg.ndim + 1 + s,) + sl] = rho_s[s]
    if cvars is not None and d.nvt > d.nv:
        for n, cv in enumerate(cvars):
            var[(d.nv + n,) + sl] = cv
    return var


def random_state(cfg, seed=0, amp=0.2):
    """Physically valid random state with per-cell noise (exercises every branch)."""
    rng = np.random.default_rng(seed)
    d = Dims(cfg)
    sh = (d.nck, d.ncj, d.nci)
    base_rho = [1.2, 0.16, 6.0][: cfg.ns] + [1.0] * max(0, cfg.ns - 3)
    rho_s = [b * rng.uniform(0.3, 1.0, sh) for b in base_rho[: cfg.ns]]
    p = 1.0e5 * rng.uniform(1 - amp, 1 + amp, sh)
    u = [rng.normal(0.0, 60.0, sh) for _ in range(cfg.ndim)]
    cv = None
    if cfg.ceq:
        cv = [np.abs(rng.normal(0, 5.0, sh)), np.abs(rng.normal(0, 5.0, sh))] + \
             [rng.normal(0, 5.0, sh) for _ in range(3)]
    return fill_conserved(cfg, rho_s, u, p, cv)


def random_state_all(cfg, seed=0, amp=0.2, umag=60.0):
    """Random physical state on every cell, ghosts included.  The generalized
    grid modules never fill ghosts (src/gen3d.cpp:341-369), so their inputs
    carry ghost values from the caller."""
    rng = np.random.default_rng(seed)
    d = Dims(cfg)
    sh = (d.ngk, d.ngj, d.ngi)
    base_rho = [1.2, 0.16, 6.0][: cfg.ns] + [1.0] * max(0, cfg.ns - 3)
    rho_s = [b * rng.uniform(0.3, 1.0, sh) for b in base_rho[: cfg.ns]]
    p = 1.0e5 * rng.uniform(1 - amp, 1 + amp, sh)
    u = [rng.normal(0.0, umag, sh) for _ in range(cfg.ndim)]
    cv = None
    if cfg.ceq:
        cv = [np.abs(rng.normal(0, 5.0, sh)) for _ in range(5)]
    return fill_conserved(cfg, rho_s, u, p, cv, ghosts=True)


def grid_nodes(cfg, fn):
    """Node array (3, nk, nj, ni) of a generalized grid: fn(i, j, k) -> (x, y, z)
    on broadcast index arrays (node indices 0..nc)."""
    d = Dims(cfg)
    k, j, i = np.meshgrid(np.arange(d.nk, dtype=np.float64), np.arange(d.nj, dtype=np.float64),
                          np.arange(d.ni, dtype=np.float64), indexing="ij")
    x, y, z = fn(i, j, k)
    g = np.zeros(d.grid_shape())
    g[0], g[1] = x, y
    if cfg.ndim == 3:
        g[2] = z
    return g


def terrain_grid(cfg, L=1.0, H=1.0, amp=0.05, beta=2.0, vertical=2):
    """Terrain-following nodes: ground z_g = amp L sin(2 pi x/L) cos(2 pi y/L),
    vertical stretch z = z_g + (H - z_g)(1 + tanh(beta (xi - 1))/tanh beta),
    xi = k/nck (3D); in 2D the ground is y_g = amp L sin(2 pi x/L) under the
    same stretch in y.  3D with vertical=1: y is the vertical axis (the axis
    gravity acts along, src/buoyancy.hpp:22-47), the ground
    y_g = amp L sin(2 pi x/L) cos(2 pi z/L) and the stretch runs along j."""
    d = Dims(cfg)

    def fn(i, j, k):
        x = L * i / d.nci
        if cfg.ndim == 3 and vertical == 1:
            z = L * k / d.nck
            yg = amp * L * np.sin(2 * np.pi * x / L) * np.cos(2 * np.pi * z / L)
            eta = j / d.ncj
            y = yg + (H - yg) * (1.0 + np.tanh(beta * (eta - 1.0)) / np.tanh(beta))
            return x, y, z
        if cfg.ndim == 3:
            y = L * j / d.ncj
            zg = amp * L * np.sin(2 * np.pi * x / L) * np.cos(2 * np.pi * y / L)
            xi = k / d.nck
            z = zg + (H - zg) * (1.0 + np.tanh(beta * (xi - 1.0)) / np.tanh(beta))
            return x, y, z
        yg = amp * L * np.sin(2 * np.pi * x / L)
        eta = j / d.ncj
        y = yg + (H - yg) * (1.0 + np.tanh(beta * (eta - 1.0)) / np.tanh(beta))
        return x, y, None

    return grid_nodes(cfg, fn)


def gen3d_terrain_config(n=(512, 512, 256), ns=3, dt=None):
    """3D generalized-grid module over terrain, n = (nci, ncj, nck) cells,
    ns species (air, hot gas, SF6)."""
    return make_config(ndim=3, ns=ns, nci=n[0], ncj=n[1], nck=n[2], grid=1, R=RU,
                       **_mix(ns, [AIR, HOT, SF6]), dt=dt if dt is not None else 1e-7)


def cell_centres(cfg, nodes):
    """Interior cell centres (the mean of each cell's 4 (2D) or 8 (3D) nodes)
    of a node array (3, nk, nj, ni): a tuple of ndim arrays (k, j, i)."""
    out = []
    for a in range(cfg.ndim):
        g = nodes[a]
        if cfg.ndim == 3:
            c = (g[:-1, :-1, :-1] + g[:-1, :-1, 1:] + g[:-1, 1:, :-1] + g[:-1, 1:, 1:] +
                 g[1:, :-1, :-1] + g[1:, :-1, 1:] + g[1:, 1:, :-1] + g[1:, 1:, 1:]) * 0.125
        else:
            c = (g[:, :-1, :-1] + g[:, :-1, 1:] + g[:, 1:, :-1] + g[:, 1:, 1:]) * 0.25
        out.append(c)
    return tuple(out)


def terrain_expansion_config(n=(512, 256, 512), ns=3, rk=3, dt=None, visc=True):
    """Ideal expansion over terrain on a generalized grid (gen3d with the
    gen_ext option): n = (nci, ncj, nck) cells, y vertical (terrain_grid with
    vertical=1), air, hot gas and SF6, gravity 9.81 along -y, viscous (the
    gen_ext viscous terms in metric form, mu = 1.8e-5), periodic in x and z,
    a no-slip ground (ymin) and a no-slip lid (ymax), SSP-RK3.
    Closed in y and periodic in x and z, the box keeps every species' mass:
    the index-space no-slip fill mirrors every momentum component and the
    metric ghosts mirror the metrics, so the contravariant face velocity at
    both walls is zero.  (A reflective wall flips only the y momentum, which
    leaves a small face velocity where the grid lines meet it at an angle.)"""
    return make_config(ndim=3, ns=ns, nci=n[0], ncj=n[1], nck=n[2], grid=1, gen_ext=True, R=RU,
                       **_mix(ns, [AIR, HOT, SF6]), visc=visc, buoyancy=True, g=9.81, rhoRef=0.0, xper=True,
                       zper=True,
                       bc=("outflow", "outflow", "noslip", "noslip", "outflow", "outflow"), rk=rk,
                       dt=dt if dt is not None else 1e-7)


def terrain_expansion_state(cfg, nodes, T0=300.0, p0=1.0e5, r0=0.05, pr=50.0, blob=(0.5, 0.2, 0.5),
                            layer=0.08):
    """At rest: an isothermal air column in hydrostatic balance,
    p = p0 exp(-g y M_air / (R T0)); a Gaussian hot-gas blob (radius r0,
    pressure ratio pr) at height blob[1] above the ground under (blob[0],
    blob[2]); with 3 species, SF6 pooled within `layer` of the ground.
    Returns (var, cell_centres)."""
    x, y, z = cell_centres(cfg, nodes)
    yg = 0.05 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * z)  # the ground under each cell (terrain_grid)
    cx, hy, cz = blob
    r2 = (x - cx) ** 2 + (y - (0.05 * np.sin(2 * np.pi * cx) * np.cos(2 * np.pi * cz) + hy)) ** 2 + (z - cz) ** 2
    yh = np.exp(-r2 / r0 ** 2)
    ysf = np.zeros_like(yh) if cfg.ns < 3 else 0.5 * (1.0 - np.tanh((y - yg - layer) / (0.25 * layer)))
    ysf = ysf * (1.0 - yh)
    yair = 1.0 - yh - ysf
    ph = p0 * np.exp(-cfg.gAccel * y * cfg.M[0] / (cfg.R * T0))
    p = ph * (1.0 + (pr - 1.0) * yh)
    Ys = [yair, yh, ysf][: cfg.ns]
    Rm = sum(Ys[s] * cfg.R / cfg.M[s] for s in range(cfg.ns))
    rho = p / (Rm * T0)
    rho_s = [rho * Ys[s] for s in range(cfg.ns)]
    zero = np.zeros_like(rho)
    var = fill_conserved(cfg, rho_s, [zero, zero, zero], p)
    return var, (x, y, z)


def gen_cfl_dt(cfg, var, nodes, cfl):
    """dt = cfl * h_min / max(|u| + c) on a generalized grid, h_min the
    shortest cell edge of the node array."""
    h = np.inf
    for a in range(3):
        e = np.sqrt(sum(np.diff(nodes[q], axis=2 - a) ** 2 for q in range(3)))
        h = min(h, float(e.min()))
    d = Dims(cfg)
    v = var[d.interior()]
    rho_s = v[4: 4 + cfg.ns]
    rho = rho_s.sum(axis=0)
    Cp = sum(rho_s[s] / rho * (cfg.gamma[s] * cfg.R / cfg.M[s] / (cfg.gamma[s] - 1)) for s in range(cfg.ns))
    Cv = sum(rho_s[s] / rho * (cfg.R / cfg.M[s] / (cfg.gamma[s] - 1)) for s in range(cfg.ns))
    gam = Cp / Cv
    m2 = v[0] ** 2 + v[1] ** 2 + v[2] ** 2
    p = (gam - 1) * (v[3] - 0.5 / rho * m2)
    c = np.sqrt(gam * p / rho) + np.sqrt(m2) / rho
    return float(cfl * h / c.max())


def sod_config(nci=1000, ncj=4, dt=2e-4, ceq=True):
    """Sod shock tube, 2D Cartesian 1000x4 on [0,1]x[0,0.004], one ideal gas
    gamma=1.4, outflow x / periodic y, WENO5 + C-equation (kappa 10, eps 1,
    alpha 2 as the reference's 2dSod deck)."""
    return make_config(ndim=2, ns=1, nci=nci, ncj=ncj, dx=1.0 / nci, dy=1.0 / nci, R=RU,
                       gamma=(1.4,), M=(0.02897,), mu=(2.928e-5,), ceq=ceq, kap=10.0, eps=1.0,
                       alpha=2.0, yper=True, bc=("outflow", "outflow", "outflow", "outflow", "outflow", "outflow"),
                       dt=dt)


def sod_state(cfg, seed=1):
    """Seed 1: Sod's states (rho, p) = (1, 1) | (0.125, 0.1); another seed:
    a shock tube with left and right states drawn from it (rho_l, p_l in
    [0.5, 2], rho_r, p_r in [0.05, 0.3])."""
    d = Dims(cfg)
    x, y, z = _cells(d)
    sh = (1, d.ncj, d.nci)
    left = np.broadcast_to(x < 0.5, sh)
    rl, pl, rr, pr_ = 1.0, 1.0, 0.125, 0.1
    if seed != 1:
        rng = np.random.default_rng(seed)
        rl, pl = rng.uniform(0.5, 2.0, 2)
        rr, pr_ = rng.uniform(0.05, 0.3, 2)
    rho = np.where(left, rl, rr)
    p = np.where(left, pl, pr_)
    u = [np.zeros(sh), np.zeros(sh)]
    return fill_conserved(cfg, [rho], u, p, [np.zeros(sh)] * 5 if cfg.ceq else None)


def expansion2d_config(n=1024, ceq=True, visc=True, ns=2):
    """2D ideal expansion: 1 m box, n^2 cells, air + hot gas, Gaussian
    high-pressure blob (50 atm, radius 0.05) in 1 atm, outflow on all sides,
    dt from CFL 0.5 at t=0."""
    mix = _mix(ns, [AIR, HOT, SF6])
    cfg = make_config(ndim=2, ns=ns, nci=n, ncj=n, dx=1.0 / n, dy=1.0 / n, R=RU, visc=visc, ceq=ceq,
                      kap=10.0, eps=1.0, alpha=2.0, bc=("outflow",) * 6, dt=1.0, **mix)
    cfg.dt = cfl_dt(cfg, expansion2d_state(cfg), 0.5)
    return cfg


def expansion2d_state(cfg, T0=300.0, r0=0.05, seed=1):
    """Seed 1: the blob at the centre, 50 atm; another seed: the blob at a
    position drawn from it in [0.35, 0.65]^2 with a pressure ratio in
    [10, 100]."""
    d = Dims(cfg)
    x, y, _ = _cells(d)
    cx, cy, pr = 0.5, 0.5, 50.0
    if seed != 1:
        rng = np.random.default_rng(seed)
        cx, cy = rng.uniform(0.35, 0.65, 2)
        pr = rng.uniform(10.0, 100.0)
    r2 = (x - cx) ** 2 + (y - cy) ** 2
    yh = np.exp(-r2 / r0 ** 2)[None, :, :] * np.ones((1, 1, 1))
    p = 101325.0 * (1.0 + (pr - 1.0) * yh)
    Rm = (1 - yh) * cfg.R / cfg.M[0] + yh * cfg.R / cfg.M[1]
    rho = p / (Rm * T0)
    rho_s = [rho * (1 - yh), rho * yh]
    sh = rho.shape
    return fill_conserved(cfg, rho_s, [np.zeros(sh), np.zeros(sh)], p,
                          [np.zeros(sh)] * 5 if cfg.ceq else None)


def multiblast2d_config(n=2048, ceq=True, learned_c=False, visc=True, rk=3, seed=4, nblobs=8, cfl=0.5, kap=10.0,
                        eps=5.0, alpha=20.0):
    """2D multi-blast: 1 m box, n^2 cells, three species (air, hot gas, SF6):
    `nblobs` Gaussian high-pressure blobs of hot gas (radius 0.05, pressure
    ratio ~ U(10, 100)) at seeded uniform positions in [0.1, 0.9]^2, in air
    at 1 atm below a heavy SF6 layer (y > 0.7, tanh interface of width
    0.01); every cell keeps at least 2 % air (see multiblast2d_state);
    outflow on all sides, viscous, SSP-RK3 (rk=3) and dt from CFL `cfl` at
    t = 0.  The C-equation takes the 3D Rayleigh-Taylor deck's constants
    (kappa 10, eps 5, alpha 20, samples/rt3d/fiesta.lua:42-45): the plain 2D
    method's wave speed includes |rho u| in place of |u| (src/ceq.hpp:72-73),
    which reaches ~1e4 m/s in the shocked SF6, and with eps = 1 (the 2D Sod
    deck) its explicit relaxation, rate maxS dt / (eps |dx|) ~ 2.8 at CFL
    0.4, leaves SSP-RK3's stability interval and diverges."""
    mix = _mix(3, [AIR, HOT, SF6])
    cfg = make_config(ndim=2, ns=3, nci=n, ncj=n, dx=1.0 / n, dy=1.0 / n, R=RU, visc=visc, ceq=ceq,
                      learned_c=learned_c, kap=kap, eps=eps, alpha=alpha, bc=("outflow",) * 6, dt=1.0, rk=rk,
                      **mix)
    cfg.dt = cfl_dt(cfg, multiblast2d_state(cfg, seed, nblobs), cfl)
    return cfg


def multiblast2d_blobs(seed=4, nblobs=8):
    """(x, y, pressure ratio) of each blob, from the seed."""
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0.1, 0.9, (nblobs, 2))
    pr = rng.uniform(10.0, 100.0, nblobs)
    return [(float(xy[b, 0]), float(xy[b, 1]), float(pr[b])) for b in range(nblobs)]


def multiblast2d_state(cfg, seed=4, nblobs=8, T0=300.0, r0=0.05, p0=101325.0, air_min=0.02):
    """The blobs' hot gas and the SF6 layer share the cells with at least
    `air_min` air.  The plain 2D C-equation needs it: its wave speed
    (maxWaveSpeed2D, src/ceq.hpp:43-78) reads the species constants with
    stride 2, so species 1 enters Cv as gamma_1 / (mu_0 - 1) < 0 and Cv
    passes zero where species 1 exceeds about 99.8 %; there the wave speed,
    and with it the C-equation's explicit relaxation rate, becomes
    unbounded (6.7e5 m/s at 2048^2 with pure hot gas at the blob centres)
    and the plain run diverges within 25 steps."""
    d = Dims(cfg)
    x, y, _ = _cells(d)
    x = np.broadcast_to(x, (1, d.ncj, d.nci))
    y = np.broadcast_to(y, (1, d.ncj, d.nci))
    yhot = np.zeros((1, d.ncj, d.nci))
    excess = np.zeros_like(yhot)  # p / p0 - 1
    for bx, by, pr in multiblast2d_blobs(seed, nblobs):
        g = np.exp(-((x - bx) ** 2 + (y - by) ** 2) / r0 ** 2)
        yhot = np.maximum(yhot, g)
        excess = excess + (pr - 1.0) * g
    p = p0 * (1.0 + excess)
    yhot = (1.0 - air_min) * yhot
    ysf6 = (1.0 - air_min - yhot) * 0.5 * (1.0 + np.tanh((y - 0.7) / 0.01))
    yair = 1.0 - yhot - ysf6
    Ys = [yair, yhot, ysf6]
    Rm = sum(Ys[s] * cfg.R / cfg.M[s] for s in range(3))
    rho = p / (Rm * T0)
    rho_s = [rho * Ys[s] for s in range(3)]
    sh = rho.shape
    return fill_conserved(cfg, rho_s, [np.zeros(sh), np.zeros(sh)], p, [np.zeros(sh)] * 5 if cfg.ceq else None)


def rt3d_config(n=256, ceq=True, visc=True, seed=3, steps_dt=None, cfl=0.2):
    """3D Rayleigh-Taylor: 1 m cube, n^3 cells, heavy air (M=28.96 g/mol)
    over light gas (M=4 g/mol), gravity 9.81 along -y, x/z periodic, y
    hydrostatic, viscous (mu 1.8e-5), C-equation (kappa 10, eps 5, alpha 20
    as the reference's rt3d deck), dt = cfl * dx / c_light.

    cfl = 0.2 per direction: with RK2 midpoint and WENO5 the perturbed
    interface grows without bound at 0.5 per direction (0.5 * 3 in 3D) within
    ~80 steps, and the C-equation's diffusion (Laplacian divided by d, not d^2)
    has an RK2 eigenvalue of about 104 * dx * (cfl / 0.5), so coarse grids need
    the smaller step too."""
    mix = _mix(2, [AIR, HOT])
    cfg = make_config(ndim=3, ns=2, nci=n, ncj=n, nck=n, dx=1.0 / n, dy=1.0 / n, dz=1.0 / n, R=RU,
                      visc=visc, ceq=ceq, buoyancy=True, g=9.81, rhoRef=0.0, kap=10.0, eps=5.0, alpha=20.0,
                      xper=True, zper=True, bc=("outflow", "outflow", "hydrostatic", "hydrostatic",
                                                "outflow", "outflow"), dt=1.0, **mix)
    # c_max = sqrt(gamma_light * p / rho_light) at p ~ 1e5 Pa, T0 = 300 K
    c_light = math.sqrt(1.67 * RU / 0.004 * 300.0)
    cfg.dt = cfl * cfg.dx / c_light if steps_dt is None else steps_dt
    return cfg


def rt3d_state(cfg, seed=3, T0=300.0, p_top=1.0e5, nmodes=16):
    """Interface y = 0.5 + eta(x,z), eta = sum of 16 Fourier modes with
    amplitudes N(0, 1e-3), distinct wavenumbers in {1..8}^2, phases U[0,2pi)."""
    d = Dims(cfg)
    rng = np.random.default_rng(seed)
    x, y, z = _cells(d)
    ks = rng.permutation(64)[:nmodes]
    eta = np.zeros((d.nck, 1, d.nci))
    for kk in ks:
        kx, kz = 1 + kk % 8, 1 + kk // 8
        a = rng.normal(0.0, 1.0e-3)
        ph = rng.uniform(0.0, 2 * math.pi)
        eta = eta + a * np.cos(2 * math.pi * (kx * x + kz * z) + ph)
    h = 0.5 + eta
    g = cfg.gAccel
    rho_h = p_top / (cfg.R / cfg.M[0] * T0)
    rho_l = p_top / (cfg.R / cfg.M[1] * T0)
    heavy = y > h
    p = np.where(heavy, p_top + rho_h * g * (1.0 - y), p_top + rho_h * g * (1.0 - h) + rho_l * g * (h - y))
    sh = p.shape
    rh = np.where(heavy, rho_h, 0.0) * np.ones(sh)
    rl = np.where(heavy, 0.0, rho_l) * np.ones(sh)
    # trace amounts keep both species present everywhere (mass fractions in (0,1))
    rh = rh + 1e-6 * rho_h
    rl = rl + 1e-6 * rho_l
    zero = np.zeros(sh)
    return fill_conserved(cfg, [rh, rl], [zero, zero, zero], p, [zero] * 5 if cfg.ceq else None)


def cfl_dt(cfg, var, cfl):
    """dt = cfl * min(d) / max(|u| + c) over the interior of `var`."""
    d = Dims(cfg)
    sl = d.interior()
    v = var[sl]
    rho_s = v[cfg.ndim + 1: cfg.ndim + 1 + cfg.ns]
    rho = rho_s.sum(axis=0)
    Cp = 0.0
    Cv = 0.0
    for s in range(cfg.ns):
        g = cfg.gamma[s]
        Rs = cfg.R / cfg.M[s]
        Cp = Cp + rho_s[s] / rho * (g * Rs / (g - 1))
        Cv = Cv + rho_s[s] / rho * (Rs / (g - 1))
    gam = Cp / Cv
    m2 = sum(v[a] ** 2 for a in range(cfg.ndim))
    p = (gam - 1) * (v[cfg.ndim] - 0.5 / rho * m2)
    c = np.sqrt(gam * p / rho) + np.sqrt(m2) / rho
    dmin = min(cfg.dx, cfg.dy, cfg.dz if cfg.ndim == 3 else cfg.dy)
    return float(cfl * dmin / c.max())


def shear3d_config(n=48, nz=4, U=100.0, ceq=True, learned_c=False, precision=0, cfl=0.4):
    """3D shear layer (Kelvin-Helmholtz): unit square in x-y, n x n x nz cells,
    heavy air (y > 0.5) moving at +U over light gas moving at -U, equal
    pressure 1 bar and temperature 300 K, x/z periodic, y outflow, viscous,
    C-equation with the reference rt3d deck's constants (kappa 10, eps 5,
    alpha 20)."""
    mix = _mix(2, [AIR, HOT])
    cfg = make_config(ndim=3, ns=2, nci=n, ncj=n, nck=nz, dx=1.0 / n, dy=1.0 / n, dz=1.0 / n, R=RU,
                      visc=True, ceq=ceq, buoyancy=False, kap=10.0, eps=5.0, alpha=20.0, xper=True, zper=True,
                      bc=("outflow",) * 6, dt=1.0, precision=precision, learned_c=learned_c, **mix)
    c_light = math.sqrt(1.67 * RU / 0.004 * 300.0)
    cfg.dt = cfl * cfg.dx / (c_light + U)
    return cfg


def shear3d_state(cfg, seed=1, U=100.0, T0=300.0, p0=1.0e5, nmodes=4, amp=0.1):
    """Interface y = 0.5 with a tanh profile two cells wide; transverse
    velocity v = amp*U*sum_m a_m sin(2 pi k_m x + phi_m) exp(-((y-0.5)/0.08)^2)
    with k_m in {1..6}, a_m ~ U[0.5, 1], phi_m ~ U[0, 2 pi) from the seed."""
    d = Dims(cfg)
    rng = np.random.default_rng(seed)
    x, y, z = _cells(d)
    delta = 2.0 * cfg.dy
    s = 0.5 * (1.0 + np.tanh((y - 0.5) / delta))
    rho_h = p0 / (cfg.R / cfg.M[0] * T0)
    rho_l = p0 / (cfg.R / cfg.M[1] * T0)
    sh = np.broadcast(x, y, z).shape
    rh = (s * rho_h + 1e-6 * rho_h) * np.ones(sh)
    rl = ((1.0 - s) * rho_l + 1e-6 * rho_l) * np.ones(sh)
    u = U * np.tanh((y - 0.5) / delta) * np.ones(sh)
    v = np.zeros(sh)
    env = np.exp(-((y - 0.5) / 0.08) ** 2)
    for _ in range(nmodes):
        k = rng.integers(1, 7)
        a = rng.uniform(0.5, 1.0)
        ph = rng.uniform(0.0, 2 * math.pi)
        v = v + amp * U * a * np.sin(2 * math.pi * k * x + ph) * env
    w = np.zeros(sh)
    p = p0 * np.ones(sh)
    d_nvt_c = Dims(cfg).nvt > Dims(cfg).nv
    return fill_conserved(cfg, [rh, rl], [u, v, w], p, [np.zeros(sh)] * 5 if d_nvt_c else None)


def riemann_ideal(rho_l, p_l, rho_r, p_r, gamma):
    """Exact solution of the Riemann problem for one ideal gas at rest on both
    sides (Toro, Riemann Solvers and Numerical Methods, ch. 4): returns the
    star pressure p*, velocity u*, the star densities left and right of the
    contact, and the right shock's speed (p_l > p_r: a left rarefaction and a
    right shock)."""
    g = gamma
    a_l, a_r = math.sqrt(g * p_l / rho_l), math.sqrt(g * p_r / rho_r)

    def f(p, rho, pk, a):
        if p > pk:
            A, B = 2.0 / ((g + 1.0) * rho), (g - 1.0) / (g + 1.0) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2.0 * a / (g - 1.0) * ((p / pk) ** ((g - 1.0) / (2.0 * g)) - 1.0)

    lo, hi = min(p_l, p_r), max(p_l, p_r)
    for _ in range(200):  # bisection: f_l + f_r is increasing in p
        mid = 0.5 * (lo + hi)
        if f(mid, rho_l, p_l, a_l) + f(mid, rho_r, p_r, a_r) > 0.0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (f(ps, rho_r, p_r, a_r) - f(ps, rho_l, p_l, a_l))
    gm = (g - 1.0) / (g + 1.0)
    rho_sr = rho_r * (ps / p_r + gm) / (gm * ps / p_r + 1.0)
    rho_sl = rho_l * (ps / p_l) ** (1.0 / g)
    s_r = a_r * math.sqrt((g + 1.0) / (2.0 * g) * ps / p_r + (g - 1.0) / (2.0 * g))
    return dict(p=ps, u=us, rho_l=rho_sl, rho_r=rho_sr, s_shock=s_r)


def shock3d_config(pr, nci=200, ncj=4, nck=4, steps=300, x0=0.4, travel=0.3, ceq=True, learned_c=False,
                   precision=0, T0=300.0, p_r=1.0e5):
    """3D shock tube along x: air at rest, temperature T0 on both sides,
    pressure pr * p_r left of x0 and p_r right of it; 1 m long, nci x ncj x
    nck cubic cells, y/z periodic, x outflow, inviscid, C-equation with the
    reference rt3d deck's constants (kappa 10, eps 5, alpha 20).  dt is set
    so that the exact shock travels `travel` in `steps` steps, whatever pr
    is (CFL about 0.3 for pr <= 12)."""
    cfg = make_config(ndim=3, ns=1, nci=nci, ncj=ncj, nck=nck, dx=1.0 / nci, dy=1.0 / nci, dz=1.0 / nci, R=RU,
                      gamma=(AIR["gamma"],), M=(AIR["M"],), mu=(0.0,), visc=False, ceq=ceq, kap=10.0, eps=5.0,
                      alpha=20.0, yper=True, zper=True, bc=("outflow",) * 6, dt=1.0, precision=precision,
                      learned_c=learned_c)
    ex = shock3d_exact(cfg, pr, T0, p_r)
    cfg.dt = travel / (ex["s_shock"] * steps)
    return cfg


def shock3d_exact(cfg, pr, T0=300.0, p_r=1.0e5):
    rho_r = p_r / (cfg.R / cfg.M[0] * T0)
    return riemann_ideal(pr * rho_r, pr * p_r, rho_r, p_r, cfg.gamma[0])


def shock3d_state(cfg, pr, x0=0.4, T0=300.0, p_r=1.0e5):
    d = Dims(cfg)
    x, y, z = _cells(d)
    sh = (d.nck, d.ncj, d.nci)
    left = np.broadcast_to(x < x0, sh)
    p = np.where(left, pr * p_r, p_r)
    rho = p / (cfg.R / cfg.M[0] * T0)
    u = [np.zeros(sh)] * 3
    return fill_conserved(cfg, [rho], u, p, [np.zeros(sh)] * 5 if d.nvt > d.nv else None)


def rm3d_config(nci=128, ncj=32, nck=4, pr=5.0, ceq=True, learned_c=False, precision=0, cfl=0.3, T0=300.0,
                p_r=1.0e5):
    """3D Richtmyer-Meshkov: a shock in air (pressure ratio pr) runs into a
    corrugated air/SF6 interface.  x in [0, 1] with nci cells, cubic cells
    (y extent ncj/nci, z extent nck/nci), y/z periodic, x outflow, inviscid,
    C-equation with the reference rt3d deck's constants.  dt = cfl dx /
    (u* + a*), the fastest signal behind the incident shock."""
    mix = _mix(2, [AIR, SF6])
    cfg = make_config(ndim=3, ns=2, nci=nci, ncj=ncj, nck=nck, dx=1.0 / nci, dy=1.0 / nci, dz=1.0 / nci, R=RU,
                      visc=False, ceq=ceq, kap=10.0, eps=5.0, alpha=20.0, yper=True, zper=True,
                      bc=("outflow",) * 6, dt=1.0, precision=precision, learned_c=learned_c, **mix)
    ex = rm3d_incident(cfg, pr, T0, p_r)
    a_s = math.sqrt(cfg.gamma[0] * ex["p"] / ex["rho_r"])
    cfg.dt = cfl * cfg.dx / (ex["u"] + a_s)
    return cfg


def rm3d_incident(cfg, pr, T0=300.0, p_r=1.0e5):
    """The incident shock: the Riemann problem of the driver (pr p_r) and the
    driven air (p_r), both at T0."""
    rho_r = p_r / (cfg.R / cfg.M[0] * T0)
    return riemann_ideal(pr * rho_r, pr * p_r, rho_r, p_r, cfg.gamma[0])


def rm3d_state(cfg, seed=1, pr=5.0, x_shock=0.12, x_int=0.25, amp=0.02, T0=300.0, p_r=1.0e5):
    """Driver air (pr p_r) left of x_shock, air at p_r up to the interface
    x = x_int + eta(y, z), SF6 at p_r beyond; eta = amp * (sum of two modes
    along y and one along z, wavenumbers and phases from the seed), smoothed
    over one cell with tanh."""
    d = Dims(cfg)
    rng = np.random.default_rng(seed)
    x, _, _ = _cells(d)
    ly, lz = d.ncj / d.nci, d.nck / d.nci
    y = ((np.arange(d.ncj) + 0.5) / d.nci)[None, :, None]
    z = ((np.arange(d.nck) + 0.5) / d.nci)[:, None, None]
    eta = np.zeros((d.nck, d.ncj, 1))
    for _ in range(2):
        k = rng.integers(1, 4)
        eta = eta + rng.uniform(0.5, 1.0) * np.sin(2 * math.pi * k * y / ly + rng.uniform(0, 2 * math.pi))
    eta = eta + 0.3 * np.sin(2 * math.pi * z / lz + rng.uniform(0, 2 * math.pi))
    eta = amp * eta / 2.3
    sh = (d.nck, d.ncj, d.nci)
    s = 0.5 * (1.0 + np.tanh((x - x_int - eta) / cfg.dx)) * np.ones(sh)  # SF6 fraction
    drv = np.broadcast_to(x < x_shock, sh)
    p = np.where(drv, pr * p_r, p_r)
    rho_air = p / (cfg.R / cfg.M[0] * T0)
    rho_sf6 = p / (cfg.R / cfg.M[1] * T0)
    ra = (1.0 - s) * rho_air + 1e-6 * rho_air
    rs = s * rho_sf6 + 1e-6 * rho_sf6
    u = [np.zeros(sh)] * 3
    return fill_conserved(cfg, [ra, rs], u, p, [np.zeros(sh)] * 5 if d.nvt > d.nv else None)
