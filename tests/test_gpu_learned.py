"""GPU tests of the learned C-indicator, all through the C-ABI.

- forward/backward of the BF16-MFMA training kernels against the numpy
  oracle with the same BF16 rounding points (tolerances cover FP32
  accumulation order only), and against exact float64 (BF16-level tolerance);
- the feature pass equals the C oracle's bit for bit;
- inference (active cells, network, residual check, fallback, ghost layer)
  against the numpy oracle;
- a full FP64 learned step equals the oracle's learned step bit for bit when
  the oracle is given the same C_hat, direction and max C_hat;
- training on rows harvested from plain runs lowers the loss;
- the gain of the trained network: on unseen shear layers the learned run's
  density is closer to the plain C-equation run than a run without the term;
- shock-driven flow (Richtmyer-Meshkov) judged by the plain run: the learned
  run keeps the shock position and is closer on density, cell by cell and as
  profiles along the shock direction.
"""
import numpy as np
import pytest

import oracle
from oracle import mlp
from fiesta_amd import workloads as W
from fiesta_amd.config import FST_FP32_BULK, FST_RHO, FST_VAR, Dims
from fiesta_amd.runtime import Module, Trainer

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


BLOCKS = {"W1": (mlp.OFF_W1, mlp.OFF_B1), "b1": (mlp.OFF_B1, mlp.OFF_W2), "W2": (mlp.OFF_W2, mlp.OFF_B2),
          "b2": (mlp.OFF_B2, mlp.OFF_W3), "W3": (mlp.OFF_W3, mlp.OFF_B3), "b3": (mlp.OFF_B3, mlp.OFF_W4),
          "w4": (mlp.OFF_W4, mlp.OFF_B4), "b4": (mlp.OFF_B4, mlp.NPARAM)}


def test_forward_backward_matches_oracle():
    print("case: forward_backward M=300 seed=5", flush=True)
    tr = Trainer(seed=5)
    w = tr.weights()
    rng = np.random.default_rng(0)
    M = 300
    X = (rng.random((M, 375)) * rng.random((M, 1))).astype(np.float32)
    T = (0.2 * rng.random(M)).astype(np.float32)
    Y, G, L = tr.forward_backward(X, T)
    loss, g, y = mlp.loss_grad(w, X, T, bf=True)
    assert _rel(Y, y) < 1e-3, _rel(Y, y)
    assert abs(L - loss) <= 2e-3 * loss
    for name, (a, b) in BLOCKS.items():
        assert _rel(G[a:b], g[a:b]) < 5e-3, (name, _rel(G[a:b], g[a:b]))
    # the BF16 path approximates the exact gradient
    _, g64, y64 = mlp.loss_grad(w, X, T, bf=False)
    assert _rel(Y, y64) < 3e-2
    for name, (a, b) in BLOCKS.items():
        assert _rel(G[a:b], g64[a:b]) < 0.1, (name, _rel(G[a:b], g64[a:b]))
    tr.close()


def _learned_case(n=20, nz=6, seed=2, precision=0):
    cfg = W.shear3d_config(n, nz=nz, learned_c=True, precision=precision)
    var = W.shear3d_state(cfg, seed)
    rng = np.random.default_rng(seed)
    d = Dims(cfg)
    sl = d.interior()
    var[sl][0:3] += 0.5 * rng.standard_normal(var[sl][0:3].shape)
    return cfg, var


def test_features_equal_oracle():
    print("case: features shear 20x20x6", flush=True)
    cfg, var = _learned_case()
    tr = Trainer(seed=3)
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.set_mlp_weights(tr.weights())
    m.learned_compute()
    chat, dr, ft, g = m.learned_fields()
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.apply_bcs()
    F, M, Dr, G = o.features()
    assert np.array_equal(ft, F)
    assert np.array_equal(dr, Dr)
    assert np.array_equal(g, G)
    assert np.abs(F[0]).max() > 0.1  # the interface is resolved by the features
    m.close()
    tr.close()


def test_inference_matches_oracle():
    print("case: inference shear 20x20x6", flush=True)
    cfg, var = _learned_case()
    d = Dims(cfg)
    tr = Trainer(seed=9)
    w = tr.weights()
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.set_mlp_weights(w)
    tau, tol = 1e-3, 0.05
    m.learned_params(tau, tol)
    m.learned_compute()
    chat, dr, ft, g = m.learned_fields()
    active, nfb, maxch = m.learned_stats()
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.apply_bcs()
    F, M, Dr, G = o.features()
    rho = o.get(FST_RHO)[0]
    hmin = min(cfg.dx, cfg.dy, cfg.dz)
    want, y, fb, cells, rn, thr = mlp.chat_field(w, F, M, G, rho, d, hmin, cfg.kap, cfg.eps,
                                                 (cfg.dx, cfg.dy, cfg.dz), tau, tol, True, False, True)
    assert active == cells.size and active > 0
    # cells whose residual sits within the network's BF16-accumulation error of
    # the threshold may decide either way on the two sides
    near = np.abs(rn - thr) < 5e-3 * max(thr, 1e-30)
    near_cells = set(cells[near].tolist())
    assert abs(nfb - int(fb.sum())) <= int(near.sum())
    assert 0 < fb.sum() < fb.size  # both the network and the fallback are exercised (untrained weights)
    got = chat.ravel()
    ref = want.ravel()
    keep = np.ones(d.N, bool)
    # a near cell and its 6 neighbours (whose fallback reads it)
    s1, s2 = d.ngi, d.ngi * d.ngj
    for c in near_cells:
        for o in (0, 1, -1, s1, -s1, s2, -s2):
            keep[c + o] = False
    scale = np.abs(ref).max()
    err = np.abs(got - ref)[keep]
    assert np.all(err <= 2e-3 * np.abs(ref[keep]) + 1e-5 * scale), float(err.max())
    assert abs(maxch - got.max()) == 0.0 or np.isclose(maxch, got.max(), rtol=1e-7)
    # the fallback cells hold the plain C-equation's local resting value
    fb_cells = cells[fb & ~near]
    assert fb_cells.size > 0
    np.testing.assert_allclose(got[fb_cells], ref[fb_cells], rtol=2e-3, atol=1e-5 * scale)
    m.close()
    tr.close()


@pytest.mark.parametrize("rk", [2, 3])
@pytest.mark.parametrize("precision", [0])
def test_learned_steps_equal_oracle(precision, rk):
    print("case: learned steps shear 24x24x8 fp64, rk", rk, flush=True)
    cfg, var = _learned_case(n=24, nz=8, seed=4, precision=precision)
    cfg.rk_order = rk
    tr = Trainer(seed=11)
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.set_mlp_weights(tr.weights())
    m.pre_sim()
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.pre_sim()
    for s in range(3):
        m.step(1)
        chat, dr, _, _ = m.learned_fields()
        _, _, mch = m.learned_stats()
        o.set_learned(chat, dr, mch)
        o.step(1)
        got, want = m.download(FST_VAR), o.get(FST_VAR)
        d = Dims(cfg)
        sl = d.interior()
        assert np.array_equal(got[sl], want[sl]), ("step", s, int((got[sl] != want[sl]).sum()))
        assert mch > 0
    m.close()
    tr.close()


@pytest.mark.parametrize("rk", [2, 3])
def test_fp32_learned_close_to_fp64_learned(rk):
    print("case: fp32 learned vs fp64 learned shear 24x24x8, rk", rk, flush=True)
    outs = []
    tr = Trainer(seed=11)
    w = tr.weights()
    for prec in (0, FST_FP32_BULK):
        cfg, var = _learned_case(n=24, nz=8, seed=4, precision=prec)
        cfg.rk_order = rk
        m = Module(cfg)
        m.upload(FST_VAR, var)
        m.set_mlp_weights(w)
        m.pre_sim()
        m.step(5)
        outs.append((cfg, m.download(FST_VAR)))
        m.close()
    d = Dims(outs[0][0])
    sl = d.interior()
    a, b = outs[1][1][sl], outs[0][1][sl]
    rho_a, rho_b = a[4:6].sum(axis=0), b[4:6].sum(axis=0)
    assert _rel(rho_a, rho_b) < 1e-8
    tr.close()


def test_training_lowers_loss():
    print("case: harvest + train shear 32x32x4", flush=True)
    cfg = W.shear3d_config(32, nz=4)
    m = Module(cfg)
    m.upload(FST_VAR, W.shear3d_state(cfg, 1))
    m.pre_sim()
    tr = Trainer(seed=1)
    for _ in range(4):
        m.step(20)
        tr.harvest(m, stride=8)
    n = tr.rows()
    assert n > 1000
    X, T = tr.read_rows(0, n)
    assert np.all(np.isfinite(X)) and np.all(T >= -1e-3 * max(T.max(), 1e-30))
    _, _, before = tr.forward_backward(X, T)
    loss = tr.train(epochs=20, batch=256, lr=1e-3, seed=3)
    assert np.all(np.isfinite(loss))
    _, _, after = tr.forward_backward(X, T)
    print("dataset loss before %.4g after %.4g (target variance %.4g)" % (before, after, T.var()))
    assert after < 0.3 * before and after < 0.3 * T.var(), (before, after, T.var())
    m.close()
    tr.close()


def _shear_run(seed, steps, ceq=True, learned=False, w=None, trainer=None, n=48):
    cfg = W.shear3d_config(n, nz=4, ceq=ceq, learned_c=learned)
    m = Module(cfg)
    m.upload(FST_VAR, W.shear3d_state(cfg, seed))
    m.pre_sim()
    if learned:
        m.set_mlp_weights(w)
    done = 0
    while done < steps:
        m.step(20)
        done += 20
        if trainer is not None and done >= 40:
            trainer.harvest(m, stride=16)
    out = m.download(FST_VAR)
    m.close()
    d = Dims(cfg)
    return out[d.interior()][4:6].sum(axis=0)


def test_learned_gain_over_no_ceq():
    """What the trained network gains: on shear layers it has not seen, the
    density of the learned run is closer to the plain C-equation run than the
    density of a run without the C-equation term is."""
    print("case: shear gain 48x48x4, train seeds 1 2 5 6, eval 3 4", flush=True)
    tr = Trainer(seed=1)
    for s in (1, 2, 5, 6):
        _shear_run(s, 300, trainer=tr)
    tr.train(epochs=40, batch=512, lr=1e-3, seed=40)
    tr.train(epochs=20, batch=512, lr=3e-4, seed=20)
    w = tr.weights()
    for s in (3, 4):
        plain = _shear_run(s, 300)
        noc = _shear_run(s, 300, ceq=False)
        lrn = _shear_run(s, 300, learned=True, w=w)
        e_noc, e_lrn = _rel(noc, plain), _rel(lrn, plain)
        print("seed %d: density rel L2 vs plain C-equation: no C %.3e, learned C %.3e" % (s, e_noc, e_lrn))
        assert e_lrn < 0.6 * e_noc, (s, e_lrn, e_noc)
    tr.close()


def _mix_p(cfg, a):
    """Pressure of the two-species mixture (src/secondary.hpp:109-151)."""
    rho_s = [a[4], a[5]]
    rho = rho_s[0] + rho_s[1]
    cp = sum(r / rho * cfg.gamma[s] * cfg.R / cfg.M[s] / (cfg.gamma[s] - 1) for s, r in enumerate(rho_s))
    cv = sum(r / rho * cfg.R / cfg.M[s] / (cfg.gamma[s] - 1) for s, r in enumerate(rho_s))
    return (cp / cv - 1) * (a[3] - 0.5 * (a[0] ** 2 + a[1] ** 2 + a[2] ** 2) / rho)


def _shock_x(cfg, a, p_ahead=1.0e5):
    """Transmitted shock position per (z, y) row: the rightmost crossing of
    p = 1.5 p_ahead, interpolated linearly between cells."""
    p = _mix_p(cfg, a)
    nk, nj, ni = p.shape
    xs = np.empty((nk, nj))
    th = 1.5 * p_ahead
    for k in range(nk):
        for j in range(nj):
            r = p[k, j]
            i = np.nonzero(r >= th)[0].max()
            xs[k, j] = (i + 0.5 + (r[i] - th) / (r[i] - r[i + 1])) / ni
    return xs


RM_STEPS = 250


def _rm_run(seed, pr, ceq=True, learned=False, w=None, trainer=None, tol=None):
    cfg = W.rm3d_config(pr=pr, ceq=ceq, learned_c=learned)
    m = Module(cfg)
    m.upload(FST_VAR, W.rm3d_state(cfg, seed, pr=pr))
    m.pre_sim()
    if learned:
        m.set_mlp_weights(w)
        if tol is not None:
            m.learned_params(1e-3, tol)
    done = 0
    while done < RM_STEPS:
        m.step(25)
        done += 25
        if trainer is not None and done >= 50:
            trainer.harvest(m, stride=16)
    stats = m.learned_stats() if learned else None
    out = m.download(FST_VAR)
    m.close()
    return cfg, out[Dims(cfg).interior()], stats


def test_learned_shock_position_and_density_profiles():
    """The plain run judges the learned one on shock-driven flow.  A shock
    (pressure ratio pr) crosses a corrugated air/SF6 interface
    (Richtmyer-Meshkov, 128x32x4, 250 steps).  Trained on plain C-equation runs
    of four interface shapes at pr 3 and 6, the network runs two unseen
    shapes at the unseen ratios 4 and 5:
    - the transmitted shock sits where the plain run has it (mean over the
      rows within 0.05 cells, every row within 0.25 cells);
    - the density, cell by cell and as the y-z averaged profile along x, is
      closer to the plain run than a run without the C-equation term is."""
    print("case: RM 128x32x4 learned vs plain, train seeds 1 2 (pr 3) 5 6 (pr 6), eval 3 (pr 4) 4 (pr 5)",
          flush=True)
    tr = Trainer(seed=3)
    for s, pr in ((1, 3.0), (2, 3.0), (5, 6.0), (6, 6.0)):
        _rm_run(s, pr, trainer=tr)
    print("rows", tr.rows(), flush=True)
    tr.train(epochs=40, batch=512, lr=1e-3, seed=41)
    tr.train(epochs=20, batch=512, lr=3e-4, seed=21)
    w = tr.weights()
    tr.close()
    for s, pr in ((3, 4.0), (4, 5.0)):
        cfg, plain, _ = _rm_run(s, pr)
        _, noc, _ = _rm_run(s, pr, ceq=False)
        _, lrn, st = _rm_run(s, pr, learned=True, w=w)
        # the same network with the residual check switched off (no fallback)
        _, raw, st_raw = _rm_run(s, pr, learned=True, w=w, tol=1e30)
        assert st_raw[1] == 0 and np.isfinite(raw).all()
        e_raw = _rel(raw[4] + raw[5], plain[4] + plain[5])
        print("seed %d pr %g: density rel L2 vs plain without the fallback %.3e (fallback cells %d of %d active)"
              % (s, pr, e_raw, st[1], st[0]), flush=True)
        rho = {k: a[4] + a[5] for k, a in (("plain", plain), ("noc", noc), ("lrn", lrn))}
        prof = {k: r.mean(axis=(0, 1)) for k, r in rho.items()}
        xs = {k: _shock_x(cfg, a) for k, a in (("plain", plain), ("noc", noc), ("lrn", lrn))}
        e3 = {k: _rel(rho[k], rho["plain"]) for k in ("noc", "lrn")}
        e1 = {k: _rel(prof[k], prof["plain"]) for k in ("noc", "lrn")}
        dx_lrn = (xs["lrn"] - xs["plain"]) * cfg.nci
        print("seed %d pr %g: density rel L2 vs plain: no C %.3e, learned %.3e; profile: no C %.3e, learned %.3e; "
              "shock x plain %.5f, learned offset mean %.4f max %.4f cells; active %d fallback %d"
              % (s, pr, e3["noc"], e3["lrn"], e1["noc"], e1["lrn"], xs["plain"].mean(), dx_lrn.mean(),
                 np.abs(dx_lrn).max(), st[0], st[1]), flush=True)
        assert xs["plain"].mean() > 0.3  # the shock has crossed the interface
        assert abs(dx_lrn.mean()) < 0.05 and np.abs(dx_lrn).max() < 0.25
        assert e3["lrn"] < 0.6 * e3["noc"], (s, e3)
        assert e1["lrn"] < 0.6 * e1["noc"], (s, e1)


def test_harvest_rejects_contexts_without_the_fused_primitives():
    """The harvest runs the 32-bit-offset primitives pass of the fused
    cartesian path: generalized-grid and 2D contexts are refused with an
    error instead of harvesting cartesian features."""
    from fiesta_amd.config import make_config
    from fiesta_amd.runtime import FiestaError
    print("case: harvest rejection (gen3d, 2D)", flush=True)
    tr = Trainer(seed=1)
    for kw in (dict(ndim=3, grid=1, nck=6), dict(ndim=2)):
        cfg = make_config(ns=2, nci=12, ncj=10, gamma=(1.4, 1.67), M=(0.029, 0.004), mu=(1e-5, 1e-5), ceq=True,
                          **kw)
        m = Module(cfg)
        with pytest.raises(FiestaError, match="fst_mlp_harvest"):
            tr.harvest(m)
        m.close()
    tr.close()
