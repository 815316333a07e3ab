#!/usr/bin/env python3
"""Benchmark of the FIESTA RHS + RK2 path on one MI355X.

Workload: 3D Rayleigh-Taylor, 256^3 cells, 2 species (air over light gas),
viscous, C-equation, gravity, RK2 midpoint (the reference integrator), FP64.
A step is one full RK2 time step (two RHS evaluations + stage updates); the
metric counts interior cells x 2 stages / wall second (Gcell-stage/s).

Prints ONE JSON line.  `value` is the FP64 plain path (bitwise parity with
the oracle).  The fast FP64 form (fp64_fast: FMA contraction and
reciprocal-based division in the WENO5 and viscous face terms) is reported
beside it with its error after the configuration's 500 steps.  The roofline block covers the dominant kernel, timed live with
HIP events on the stream it runs on; algorithmic bytes are 20*nvt per
cell-stage (stage 1 reads U, writes U*; stage 2 reads U*, U^n, writes U^{n+1}).

Forms: fp64 (reported as `value`), fp32 (FP32 stencils, FP64 EOS/RK) and
fp32_learned (fp32 + the learned C-indicator in place of the C-equation's
relaxation solve: nvt = nv); a suffix _rk3 (fp64_rk3, fp32_learned_rk3) runs
the three-stage SSP-RK3 option in place of the reference's RK2 midpoint and
is compared with fp64_rk3.  The learned network is trained first, outside
the timed region, on rows harvested on the GPU from plain runs (a 64^3 RT
run and four 48x48x4 shear layers); each form's final state is compared with
the FP64 plain final state (density, pressure, mass fractions).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU legs (baseline and the 1000 x 4 Sod tube)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--path", choices=["fused", "reference"], default="fused")
    ap.add_argument("--copy-probe", action="store_true", help="also time a plain copy kernel")
    ap.add_argument("--forms", default="fp64,fp32,fp32_learned,fp64_fast,fp64_rk3,fp32_learned_rk3",
                    help="comma list of arithmetic forms to time: fp64 (reported as value), fp32, fp32_learned, "
                         "fp64_fast; suffix _rk3 runs the SSP-RK3 option instead of the reference's RK2 midpoint")
    ap.add_argument("--accuracy-steps", type=int, default=500,
                    help="steps of the untimed accuracy leg: each RK2 form against fp64 after this many steps "
                         "(0: skip); the Rayleigh-Taylor configuration runs 500")
    ap.add_argument("--gen3d-size", type=int, nargs=3, default=[512, 512, 256],
                    help="cells of the gen3d leg of --extras")
    ap.add_argument("--extras", action="store_true",
                    help="also time postStep's noise filter, the device-side output pack (fst_pack), the gen3d module and the 2D expansion on the "
                         "fp64 workload")
    ap.add_argument("--gen3d", action="store_true",
                    help="the gen3d module on the terrain grid (FP64 and FP32 forms); also part of --extras")
    ap.add_argument("--cart2d", action="store_true",
                    help="the 2D module on the expansion (fused stage and kernel sequence); also part of --extras")
    ap.add_argument("--multiblast", action="store_true",
                    help="the 2D learned C-indicator on the multi-blast (trained on 2D Sod and expansion runs), "
                         "against the plain 2D C-equation; also part of --extras")
    ap.add_argument("--terrain", action="store_true",
                    help="the 3D ideal expansion over terrain (gen3d with the gen_ext option: 3 species, gravity, "
                         "no-slip ground and lid, SSP-RK3); also part of --extras")
    ap.add_argument("--terrain-size", type=int, nargs=3, default=[512, 256, 512],
                    help="cells (i, j vertical, k) of the terrain expansion")
    ap.add_argument("--terrain-steps", type=int, default=300)
    ap.add_argument("--multiblast-size", type=int, default=2048)
    ap.add_argument("--multiblast-steps", type=int, default=400)
    ap.add_argument("--save-weights", default=None, help="write the trained network (.npy, float32)")
    ap.add_argument("--load-weights", default=None, help="use this network (.npy) instead of training one")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def train_learned():
    """Harvest (stencil, C_hat) rows from plain FP64 runs on the GPU and train
    the network with the library's own kernels.  Returns weights and a summary."""
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_VAR
    from fiesta_amd.runtime import Module, Trainer

    t0 = time.perf_counter()
    tr = Trainer(seed=1)
    runs = []
    cfg = W.rt3d_config(64)
    runs.append((cfg, W.rt3d_state(cfg), 200))
    for s in (1, 2, 5, 6):
        cfg = W.shear3d_config(48, nz=4)
        runs.append((cfg, W.shear3d_state(cfg, s), 300))
    for cfg, var, steps in runs:
        m = Module(cfg)
        m.upload(FST_VAR, var)
        m.pre_sim()
        for done in range(20, steps + 1, 20):
            m.step(20)
            if done >= 40:
                tr.harvest(m, stride=16)
        m.close()
    rows = tr.rows()
    t1 = time.perf_counter()
    l1 = tr.train(epochs=40, batch=512, lr=1e-3, seed=40)
    l2 = tr.train(epochs=20, batch=512, lr=3e-4, seed=20)
    t2 = time.perf_counter()
    w = tr.weights()
    tr.close()
    info = {"rows": int(rows), "harvest_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2),
            "epochs": 60, "batch": 512, "loss_first_epoch": float("%.4g" % l1[0]),
            "loss_last_epoch": float("%.4g" % l2[-1]),
            "data": "plain FP64 runs: RT 64^3 (200 steps) + 4 shear layers 48x48x4 (300 steps), "
                    "harvested every 20 steps from step 40"}
    return w, info


def train_learned2d(max_rows=2_000_000, c2_size=1024, c2_steps=2000, sod_steps=1000):
    """The 2D network (5x5 stencils) on rows harvested from plain FP64 cart2d
    runs with SSP-RK3: the Sod tube (1000 x 4, dt 2e-4, `sod_steps` steps;
    seed 1 Sod's states, seed 2 drawn ones), harvested every 20 steps, every
    4th smooth cell; and the ideal expansion (c2_size^2, CFL 0.5, `c2_steps`
    steps; seed 1 the centred 50 atm blob, seed 2 a drawn position and
    ratio), harvested every 100 steps from step 100, every 64th smooth cell;
    at most max_rows rows."""
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_VAR
    from fiesta_amd.runtime import Module, Trainer

    t0 = time.perf_counter()
    tr = Trainer(seed=1, ndim=2)
    plan = []
    for seed in (1, 2):
        cfg = W.sod_config(1000, 4, dt=2e-4)
        cfg.rk_order = 3
        plan.append((cfg, W.sod_state(cfg, seed), sod_steps, 20, 20, 4))
    for seed in (1, 2):
        cfg = W.expansion2d_config(c2_size)
        cfg.rk_order = 3
        plan.append((cfg, W.expansion2d_state(cfg, seed=seed), c2_steps, 100, 100, 64))
    per_run = []
    for cfg, var, steps, every, first, stride in plan:
        m = Module(cfg)
        m.upload(FST_VAR, var)
        m.pre_sim()
        n0 = tr.rows()
        for done in range(every, steps + 1, every):
            m.step(every)
            # only states whose plain C_hat is bounded: the 2D expansion's
            # plain C-equation diverges for a while where the hot gas is
            # nearly pure (its wave speed misreads the species constants)
            if done >= first and tr.rows() < max_rows and m.maxima()[1] < 1e6:
                tr.harvest(m, stride=stride, max_rows=max_rows - tr.rows())
        m.close()
        per_run.append(int(tr.rows() - n0))
    rows = tr.rows()
    t1 = time.perf_counter()
    l1 = tr.train(epochs=12, batch=1024, lr=1e-3, seed=40)
    l2 = tr.train(epochs=6, batch=1024, lr=3e-4, seed=20)
    t2 = time.perf_counter()
    w = tr.weights()
    tr.close()
    info = {"rows": int(rows), "rows_per_run": per_run, "harvest_s": round(t1 - t0, 2),
            "train_s": round(t2 - t1, 2), "epochs": 18, "batch": 1024,
            "loss_first_epoch": float("%.4g" % l1[0]), "loss_last_epoch": float("%.4g" % l2[-1]),
            "data": "plain FP64 cart2d SSP-RK3 runs: Sod 1000x4 seeds 1-2 (%d steps), expansion %d^2 seeds 1-2 "
                    "(%d steps)" % (sod_steps, c2_size, c2_steps)}
    return w, info


def _blast_rays(cfg, var, blobs, rmax=0.3):
    """Pressure and density along four rays (+-x, +-y) from each blob centre
    (nearest cells, one per cell width) and the shock position on each ray:
    the steepest pressure drop beyond the blob radius, in cells."""
    from fiesta_amd.config import Dims
    d = Dims(cfg)
    g = d.ng
    ns = cfg.ns
    v = var[:, 0]
    rho = v[3:3 + ns].sum(axis=0)
    Cp = 0.0
    Cv = 0.0
    for q in range(ns):
        Rs = cfg.R / cfg.M[q]
        Cp = Cp + v[3 + q] / rho * (cfg.gamma[q] * Rs / (cfg.gamma[q] - 1))
        Cv = Cv + v[3 + q] / rho * (Rs / (cfg.gamma[q] - 1))
    p = (Cp / Cv - 1) * (v[2] - 0.5 * (v[0] ** 2 + v[1] ** 2) / rho)
    n = d.nci
    out = []
    for bx, by, _ in blobs:
        ci, cj = int(bx * n), int(by * n)
        for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            lim = min(int(rmax * n), (n - 1 - ci) if dx > 0 else ci if dx < 0 else n,
                      (n - 1 - cj) if dy > 0 else cj if dy < 0 else n)
            r = np.arange(lim)
            ii, jj = g + ci + dx * r, g + cj + dy * r
            pr, rr = p[jj, ii], rho[jj, ii]
            dp = np.diff(pr)
            start = int(0.05 * n)
            k = start + int(np.argmin(dp[start:])) if dp.size > start + 1 else 0
            out.append((k, pr, rr))
    return out


def multiblast_leg(args, weights, train_info):
    """The 2D multi-blast (n^2 cells, 8 blobs, 3 species,
    viscous, SSP-RK3) with the plain 2D C-equation, the learned C-indicator
    (FP64 cart2d sequence) and no C-equation; the learned and C-equation-free
    runs judged by the plain run on density (field and ray profiles) and
    shock positions, and timed (device ms per step)."""
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_VAR, Dims
    from fiesta_amd.runtime import Module

    n, steps = args.multiblast_size, args.multiblast_steps
    blobs = W.multiblast2d_blobs(4, 8)
    res = {}
    for form in ("plain", "learned", "no_ceq"):
        cfg = W.multiblast2d_config(n, ceq=form != "no_ceq", learned_c=form == "learned")
        d = Dims(cfg)
        m = Module(cfg)
        m.upload(FST_VAR, W.multiblast2d_state(cfg))
        if form == "learned":
            m.set_mlp_weights(weights)
        m.pre_sim()
        m.step(steps)
        ms = m.last_step_ms()
        e = {"ms_per_step": round(ms / steps, 4), "ms_per_stage": round(ms / steps / 3, 4),
             "value": round(d.ncells * 3 * steps / (ms * 1e-3) / 1e9, 4), "nvt": d.nvt}
        if form == "learned":
            act, fb, mch = m.learned_stats()
            e["active_cells_last_step"] = int(act)
            e["fallback_cells_last_step"] = int(fb)
        res[form] = (cfg, m.download(FST_VAR), e)
        m.close()
        log("multiblast %s: %.3f ms/step" % (form, ms / steps))
    cfg_p, vp, ep = res["plain"]
    dp = Dims(cfg_p)
    sl = dp.interior()[1:]
    rho_p = vp[3:6][(slice(None),) + sl].sum(axis=0)
    mom_p = np.hypot(vp[0][sl], vp[1][sl])
    rays_p = _blast_rays(cfg_p, vp, blobs)
    out = {"workload": "2D multi-blast %d^2, 8 blobs (seed 4, p ratio U(10,100)), air / hot gas / SF6, viscous, "
                       "SSP-RK3, %d steps, dt %.4g" % (n, steps, cfg_p.dt),
           "plain_ceq": ep, "training": train_info}
    for form in ("learned", "no_ceq"):
        cfg, v, e = res[form]
        rho = v[3:6][(slice(None),) + sl].sum(axis=0)
        mom = np.hypot(v[0][sl], v[1][sl])
        rays = _blast_rays(cfg, v, blobs)
        dk = np.array([abs(a[0] - b[0]) for a, b in zip(rays, rays_p)])
        prof = np.concatenate([a[2] - b[2] for a, b in zip(rays, rays_p)])
        prof_ref = np.concatenate([b[2] for b in rays_p])
        e.update({"density_rel_l2_vs_plain": float("%.4g" % (np.linalg.norm(rho - rho_p) / np.linalg.norm(rho_p))),
                  "density_linf_rel_vs_plain": float("%.4g" % (np.abs(rho - rho_p).max() / np.abs(rho_p).max())),
                  "momentum_rel_l2_vs_plain": float("%.4g" % (np.linalg.norm(mom - mom_p) / np.linalg.norm(mom_p))),
                  "ray_density_profile_rel_l2_vs_plain": float("%.4g" % (np.linalg.norm(prof) /
                                                                         np.linalg.norm(prof_ref))),
                  "shock_position_cells_vs_plain": {"mean_abs": round(float(dk.mean()), 3), "max_abs": int(dk.max()),
                                                    "rays": int(dk.size)},
                  "speedup_vs_plain": round(e["value"] / ep["value"], 3)})
        out[form] = e
    l, z = out["learned"], out["no_ceq"]
    out["learned_error_over_no_ceq_error"] = {
        k: round(l[k] / z[k], 4) if z[k] else None
        for k in ("density_rel_l2_vs_plain", "momentum_rel_l2_vs_plain", "ray_density_profile_rel_l2_vs_plain")}
    return out


def gpu_leg(args, form="fp64", var=None, keep_state=False, weights=None):
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_FP32_BULK, FST_FP64_FAST, FST_VAR, Dims
    from fiesta_amd.runtime import Module

    cfg = W.rt3d_config(args.size)
    base, nst = (form[:-4], 3) if form.endswith("_rk3") else (form, 2)
    cfg.rk_order = nst
    if base.startswith("fp32"):
        cfg.precision = FST_FP32_BULK
    elif base == "fp64_fast":
        cfg.precision = FST_FP64_FAST
    learned = base.endswith("learned")
    if learned:
        cfg.learned_c = 1
    d = Dims(cfg)
    m = Module(cfg)
    m.upload(FST_VAR, var[:d.nvt])
    if learned:
        m.set_mlp_weights(weights)
    m.pre_sim()
    fused_ok = True
    if args.path == "reference":
        m.set_fused(False)
    else:
        try:
            m.set_fused(True)
        except Exception:
            fused_ok = False
            m.set_fused(False)
    m.step(args.warmup)
    m.sync()
    log("warmup done")
    m.profile(True)
    t0 = time.perf_counter()
    m.step(args.steps)
    m.sync()
    wall = time.perf_counter() - t0
    launches, kms, kname = m.profile_read()
    phases = m.profile_phases()
    m.profile(False)
    ev_ms = m.last_step_ms()
    cell_stages = d.ncells * nst * args.steps
    value = cell_stages / wall / 1e9
    res = dict(cfg=cfg, dims=d, wall=wall, ev_ms=ev_ms, value=value, launches=launches, kernel_ms=kms,
               kernel=kname, fused=fused_ok and args.path == "fused", phases=phases)
    if learned:
        a, nfb, mch = m.learned_stats()
        res["learned"] = {"active_cells": int(a), "fallback_cells": int(nfb), "max_chat": float("%.5g" % mch)}
    if keep_state:
        res["state"] = m.download(FST_VAR)
    if args.copy_probe or form == "fp64":
        nbytes = 2 << 30
        ms = m.copy_probe(nbytes, 10)
        res["copy_gbs"] = 2 * nbytes / (ms * 1e-3) / 1e9
    m.close()
    return res


def extras_leg(args, var):
    """postStep with the noise filter (mode 1: three momentum components, one
    sweep, C_hat cut-off on) and fst_pack of every variable (float32, full
    interior) on the fp64 workload; wall-clock ms per call after one warm-up."""
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_VAR, Dims
    from fiesta_amd.runtime import Module

    cfg = W.rt3d_config(args.size)
    cfg.noise, cfg.n_dh, cfg.n_eta, cfg.n_coff, cfg.n_nt, cfg.n_mode = 1, 1e-3, 0.05, 0.1, 1, 1
    d = Dims(cfg)
    m = Module(cfg)
    m.upload(FST_VAR, var[:d.nvt])
    m.pre_sim()
    m.post_step()
    m.sync()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        m.post_step()
    m.sync()
    post_ms = (time.perf_counter() - t0) * 1e3 / reps
    n = args.size
    box = ((0, 0, 0), (n - 1, n - 1, n - 1), (1, 1, 1))
    m.pack(FST_VAR, 0, *box)
    t0 = time.perf_counter()
    for v in range(d.nvt):
        m.pack(FST_VAR, v, *box)
    pack_ms = (time.perf_counter() - t0) * 1e3
    m.close()
    return {"post_step_with_noise_ms": round(post_ms, 3),
            "pack_all_vars_f32_ms": round(pack_ms, 3),
            "pack_bytes_to_host": int(d.nvt * d.ncells * 4),
            "note": "postStep = secondary variables + noise filter (3 variables, 1 sweep); pack = %d variables, "
                    "float32, whole interior, each copied to pageable host memory" % d.nvt}


def gen3d_leg(args):
    """The generalized-grid module (gen3d: EOS, contravariant velocity, WENO
    fluxes, metric pressure gradient; src/gen3d.cpp:341-369) on the 3D
    terrain workload, 512 x 512 x 256 cells, 3 species, on the fused gen3d
    stage: the advection kernel (fluxes of every variable and the metric
    pressure gradient in one launch), then one pass with the RK update and
    the EOS / contravariant velocity of the new state.  FP64 (the reference
    arithmetic) and the FP32 form (binary32 fluxes, float primitive planes;
    FP64 pressure gradient, EOS and RK accumulation).  Device ms per step from
    hipEvents over 3 steps after one warm-up step; the FP32 form's error
    against FP64 after 20 steps from the same state."""
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_FP32_BULK, FST_FP64, FST_GRID, FST_VAR, Dims
    from fiesta_amd.runtime import Module

    n = tuple(args.gen3d_size)
    cfg = W.gen3d_terrain_config(n=n, ns=3, dt=1e-7)
    d = Dims(cfg)
    nodes = W.terrain_grid(cfg)
    var = W.random_state_all(cfg, seed=5, umag=20.0)
    acc = 20
    ms, states = {}, {}
    for form, prec in (("fp64", FST_FP64), ("fp32", FST_FP32_BULK)):
        cfg.precision = prec
        m = Module(cfg)
        m.upload(FST_GRID, nodes)
        m.upload(FST_VAR, var)
        m.pre_sim()
        m.step(1)
        m.sync()
        steps = 3
        m.step(steps)
        ms[form] = m.last_step_ms() / steps
        m.step(acc - 1 - steps)
        states[form] = m.download(FST_VAR)
        m.close()
    cfg.precision = FST_FP64
    cs = d.ncells * 2
    balg = 20 * d.nvt          # U read/write of the RK2 stages
    bmet = balg + 9 * 8        # plus one read of the 9 metric terms per stage

    def line(ms_step):
        stage_ms = ms_step / 2
        return {"ms_per_stage": round(stage_ms, 4), "value": round(cs / (ms_step * 1e-3) / 1e9, 4),
                "unit": "Gcell-stage/s",
                "roofline": {"bound": "hbm", "algorithmic_bytes_per_cell_stage": balg,
                             "with_metrics_bytes_per_cell_stage": bmet,
                             "achieved_GBps": round(d.ncells * balg / (stage_ms * 1e-3) / 1e9, 1),
                             "frac": round(d.ncells * balg / (stage_ms * 1e-3) / 8.0e12, 4)}}

    out = {"workload": "gen3d terrain %dx%dx%d, 3 species, RK2, FP64 (fused gen3d stage: advection + metric "
                       "pressure gradient kernel, RK update + EOS kernel)" % n}
    out.update(line(ms["fp64"]))
    f32 = line(ms["fp32"])
    f32["speedup_vs_fp64"] = round(ms["fp64"] / ms["fp32"], 3)
    f32["error_vs_fp64_%d_steps" % acc] = form_errors(cfg, states["fp32"], states["fp64"], cfg)
    out["fp32"] = f32
    return out


def terrain_leg(args):
    """3D ideal expansion over terrain: gen3d (src/gen3d.cpp:341-369) with the
    gen_ext option (the cartesian ghost fill and buoyancy in index space and
    the viscous terms in metric form, which the reference's gen3d lacks),
    512 x 256 x 512 cells with y vertical, ground y_g = 0.05 sin(2 pi x)
    cos(2 pi z) and a tanh (beta = 2) vertical stretch, air + hot gas + SF6,
    a 50:1 hot-gas blob over a pooled SF6 layer, gravity, no-slip ground and
    lid, periodic x and z, SSP-RK3 at CFL 0.4.  FP64 and the FP32 form
    (binary32 fluxes and viscous faces) from the same state: device ms per
    step from hipEvents over the steps after one warm-up step; then each
    final state is checked (finite, every species' total the same to
    rounding: the closed box keeps mass) and the FP32 form's density,
    pressure and mass fractions are compared with the FP64 path's (equal to
    the oracle bit for bit) after the same args.terrain_steps steps."""
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_FP32_BULK, FST_FP64, FST_GRID, FST_VAR, Dims
    from fiesta_amd.runtime import Module

    n = tuple(args.terrain_size)
    cfg = W.terrain_expansion_config(n)
    nodes = W.terrain_grid(cfg, vertical=1)
    var, _ = W.terrain_expansion_state(cfg, nodes)
    cfg.dt = W.gen_cfl_dt(cfg, var, nodes, 0.4)
    d = Dims(cfg)
    sl = d.interior()
    mass0 = var[sl][4:].sum(axis=(1, 2, 3))
    steps = max(1, args.terrain_steps - 1)
    nst = 3
    cs = d.ncells * nst
    balg = 8 * 8 * d.nvt / 3.0     # SSP-RK3: U reads/writes of the three stages, per stage
    bmet = balg + 9 * 8            # plus one read of the 9 metric terms per stage
    res, states = {}, {}
    for form, prec in (("fp64", FST_FP64), ("fp32", FST_FP32_BULK)):
        cfg.precision = prec
        m = Module(cfg)
        m.upload(FST_GRID, nodes)
        m.upload(FST_VAR, var)
        m.pre_sim()
        m.step(1)
        m.sync()
        log("terrain: %s, %d timed steps" % (form, steps))
        m.step(steps)
        ms_step = m.last_step_ms() / steps
        v = m.download(FST_VAR)
        m.close()
        vi = v[sl]
        rho = vi[4:].sum(axis=0)
        mass1 = vi[4:].sum(axis=(1, 2, 3))
        umax = float((np.sqrt(vi[0] ** 2 + vi[1] ** 2 + vi[2] ** 2) / rho).max())
        del vi, rho
        states[form] = v
        stage_ms = ms_step / nst
        res[form] = {"ms_per_stage": round(stage_ms, 4), "value": round(cs / (ms_step * 1e-3) / 1e9, 4),
                     "unit": "Gcell-stage/s",
                     "roofline": {"bound": "hbm", "algorithmic_bytes_per_cell_stage": round(balg, 1),
                                  "with_metrics_bytes_per_cell_stage": round(bmet, 1),
                                  "achieved_GBps": round(d.ncells * balg / (stage_ms * 1e-3) / 1e9, 1),
                                  "frac": round(d.ncells * balg / (stage_ms * 1e-3) / 8.0e12, 4)},
                     "state_after": {"finite": bool(np.isfinite(mass1).all()), "max_speed": round(umax, 2),
                                     "species_mass_rel_change": [float("%.3g" % x) for x in (mass1 - mass0) / mass0]}}
    del nodes, var
    cfg.precision = FST_FP64
    out = {"workload": "3D ideal expansion over terrain %dx%dx%d (y vertical), 3 species, gravity, viscous, "
                       "no-slip ground and lid, periodic x/z, SSP-RK3, FP64 (gen3d + gen_ext)" % n,
           "dt": cfg.dt, "steps": steps + 1}
    out.update(res["fp64"])
    out["parity"] = ("bitwise equal to the oracle on the same workload at 40x20x40 over 12 steps "
                     "(tests/test_gpu_gen.py::test_terrain_expansion_against_oracle)")
    f32 = res["fp32"]
    f32["speedup_vs_fp64"] = round(res["fp64"]["ms_per_stage"] / f32["ms_per_stage"], 3)
    f32["error_vs_fp64_%d_steps" % (steps + 1)] = form_errors(cfg, states["fp32"], states["fp64"], cfg)
    out["fp32"] = f32
    return out


def expansion2d_leg(args):
    """The 2D module (cart2d: BCs once per step, then per stage EOS, WENO5,
    pressure gradient, viscous terms and the C-equation; src/cart2d.cpp:
    136-237) on the 2D ideal expansion, 1024 x 1024 cells, 2 species,
    viscous, C-equation, SSP-RK3 (3 stages per step), dt from CFL 0.5: the
    fused cart2d stage kernel and the reference-sequence kernels.  Device ms
    per step from hipEvents over 20 steps after 2 warm-up steps."""
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_VAR, Dims

    from fiesta_amd.runtime import Module

    cfg = W.expansion2d_config(1024, ceq=True, visc=True, ns=2)
    cfg.rk_order = 3
    d = Dims(cfg)
    balg = 64 * d.nvt // 3  # SSP-RK3: (2 + 3 + 3) / 3 x 8 nvt bytes per cell-stage
    out = {"workload": "cart2d expansion 1024x1024, 2 species, viscous, C-equation, SSP-RK3, FP64",
           "unit": "Gcell-stage/s", "algorithmic_bytes_per_cell_stage": balg}
    states = {}
    for path in ("fused", "reference_sequence"):
        m = Module(cfg)
        m.upload(FST_VAR, W.expansion2d_state(cfg))
        m.pre_sim()
        m.set_fused(path == "fused")
        m.step(2)
        m.sync()
        steps = 20
        m.step(steps)
        ms_step = m.last_step_ms() / steps
        states[path] = m.download(FST_VAR)
        m.close()
        stage_ms = ms_step / 3
        ach = d.ncells * balg / (stage_ms * 1e-3) / 1e9
        out[path] = {"ms_per_stage": round(stage_ms, 4), "value": round(d.ncells / (stage_ms * 1e-3) / 1e9, 4),
                     "roofline": {"bound": "hbm", "achieved_GBps": round(ach, 1),
                                  "frac": round(ach * 1e9 / 8.0e12, 4)}}
    out["fused_equals_reference_sequence"] = bool(np.array_equal(states["fused"], states["reference_sequence"]))
    out["speedup_fused"] = round(out["fused"]["value"] / out["reference_sequence"]["value"], 3)
    out["value"] = out["fused"]["value"]
    out["ms_per_stage"] = out["fused"]["ms_per_stage"]
    out["roofline"] = out["fused"]["roofline"]
    return out


def cpu_threads():
    """The host's stated core count: OMP_NUM_THREADS when the environment sets
    it (the GPU hosts state their CPU share there), else os.cpu_count()."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_leg(args, cfg_size):
    # thread placement is read when the OpenMP runtime starts (first load of the oracle)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    import oracle
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_VAR

    threads = cpu_threads()
    oracle.set_threads(threads)
    cfg = W.rt3d_config(cfg_size)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, W.rt3d_state(cfg))
    o.pre_sim()
    o.step(1)
    t0 = time.perf_counter()
    o.step(args.cpu_steps)
    wall = time.perf_counter() - t0
    cs = cfg.nci * cfg.ncj * cfg.nck * 2 * args.cpu_steps
    return dict(value=cs / wall / 1e9, unit="Gcell-stage/s", cores=threads, kind="port",
                sample="%d RK2 steps (after 1 warm-up) of the %d^3 workload with the oracle "
                       "(C++/OpenMP restatement of the reference kernels), %d threads, OMP_PROC_BIND=%s"
                       % (args.cpu_steps, cfg_size, threads, os.environ.get("OMP_PROC_BIND")))


def sod_leg():
    """The Sod shock tube in full, 1000 x 4 cells, WENO5 +
    C-equation, RK2, dt 2e-4, 1000 steps, on the GPU (the cart2d module's
    kernels) and on the CPU with the oracle; the final states compared bit
    for bit and against the exact Riemann solution at t = 0.2."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    import oracle
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_VAR, Dims
    from fiesta_amd.runtime import Module

    cfg = W.sod_config()
    d = Dims(cfg)
    var = W.sod_state(cfg)
    steps = 1000
    m = Module(cfg)
    m.upload(FST_VAR, var)
    m.pre_sim()
    m.sync()
    t0 = time.perf_counter()
    m.step(steps)
    m.sync()
    gpu_s = time.perf_counter() - t0
    got = m.download(FST_VAR)
    m.close()
    threads = cpu_threads()
    oracle.set_threads(threads)
    o = oracle.Oracle(cfg)
    o.set(FST_VAR, var)
    o.pre_sim()
    t0 = time.perf_counter()
    o.step(steps)
    cpu_s = time.perf_counter() - t0
    want = o.get(FST_VAR)
    rho = got[3, 0, d.ng, d.ng:d.ng + d.nci]  # density (one species) along x, first row
    x = (np.arange(d.nci) + 0.5) / d.nci
    ex = W.riemann_ideal(1.0, 1.0, 0.125, 0.1, 1.4)
    mid = 0.5 * (ex["rho_r"] + 0.125)
    shock = float(x[np.where((x > 0.7) & (rho < mid))[0][0]])
    cs = d.ncells * 2 * steps
    return {"workload": "Sod shock tube 1000x4, WENO5 + C-equation, RK2, dt 2e-4, 1000 steps, FP64",
            "gpu_s": round(gpu_s, 4), "gpu_gcell_stage_per_s": round(cs / gpu_s / 1e9, 5),
            "cpu_s": round(cpu_s, 4), "cpu_gcell_stage_per_s": round(cs / cpu_s / 1e9, 5), "cpu_threads": threads,
            "bitwise_equal_to_oracle": bool(np.array_equal(got, want)),
            "shock_position": round(shock, 4), "exact_shock_position": round(0.5 + 0.2 * ex["s_shock"], 4)}


def accuracy_leg(args, var, forms, weights):
    """Each RK2 form run for args.accuracy_steps steps (untimed) from the
    same state, and its final density, pressure and mass fractions compared
    with the FP64 path's (equal to the oracle bit for bit)."""
    from fiesta_amd import workloads as W
    from fiesta_amd.config import FST_FP32_BULK, FST_FP64_FAST, FST_VAR, Dims
    from fiesta_amd.runtime import Module

    n = args.accuracy_steps
    states = {}
    for f in ["fp64"] + [f for f in forms if f != "fp64" and not f.endswith("_rk3")]:
        cfg = W.rt3d_config(args.size)
        if f.startswith("fp32"):
            cfg.precision = FST_FP32_BULK
        elif f == "fp64_fast":
            cfg.precision = FST_FP64_FAST
        if f.endswith("learned"):
            cfg.learned_c = 1
        d = Dims(cfg)
        m = Module(cfg)
        m.upload(FST_VAR, var[:d.nvt])
        if cfg.learned_c:
            m.set_mlp_weights(weights)
        m.pre_sim()
        m.step(n)
        states[f] = (cfg, m.download(FST_VAR))
        m.close()
        log("accuracy leg: %s, %d steps" % (f, n))
    cfg64, ref = states["fp64"]
    out = {}
    for f, (cfg, st) in states.items():
        if f != "fp64":
            out[f] = form_errors(cfg, st, ref, cfg64)
    return out


def measured_traffic(size, form):
    """HBM bytes per RK stage of a form, from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a copy kernel
    of known size); None when no summary for this size and form exists."""
    path = os.path.join(ROOT, "profiles", "traffic_rt3d_%d.json" % size)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    e = d.get("forms", {}).get(form)
    if not e:
        return None
    return {"bytes_per_stage": e["bytes_per_stage"], "source": "profiles/%s (%s)" % (os.path.basename(path),
                                                                                    d.get("command", ""))}


def issue_floor(size, kernels):
    """VALU issue time per stage of the given stage kernels, from the committed
    rocprofv3 SQ counter summary: SQ_INSTS_VALU per dispatch / (1024 SIMDs)
    x 4 cycles per wave64 instruction at 2.4 GHz, the time the kernels would
    take if they did nothing but issue VALU instructions (None when the
    summary or a kernel is missing)."""
    path = os.path.join(ROOT, "profiles", "sq_counters_rt3d_%d.json" % size)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {})
    if not all(n in k for n in kernels):
        return None
    per = d["cycles_per_wave_instruction"] / (d["simds"] * d["clock_ghz"] * 1e9) * 1e3
    return {n: k[n]["SQ_INSTS_VALU"] * per for n in kernels}


def form_errors(cfg, got, want, cfg64):
    """Relative L2 / Linf of density, pressure and mass fractions (interior)
    of a form's final state against the FP64 path's final state (the FP64
    path equals the oracle bit for bit)."""
    from fiesta_amd.config import Dims

    RU = cfg.R

    def fields(c, var):
        d = Dims(c)
        v = var[d.interior()]
        rs = v[4:4 + c.ns]
        rho = rs.sum(axis=0)
        Y = rs / rho
        cp = sum(Y[s] * c.gamma[s] * (RU / c.M[s]) / (c.gamma[s] - 1) for s in range(c.ns))
        cv = sum(Y[s] * (RU / c.M[s]) / (c.gamma[s] - 1) for s in range(c.ns))
        p = (cp / cv - 1) * (v[3] - 0.5 / rho * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
        return {"rho": rho, "p": p, "Y": Y}

    a, b = fields(cfg, got), fields(cfg64, want)
    out = {}
    for k in ("rho", "p", "Y"):
        dd = a[k] - b[k]
        out[k] = {"rel_l2": float("%.3e" % np.sqrt((dd ** 2).sum() / (b[k] ** 2).sum())),
                  "rel_linf": float("%.3e" % (np.abs(dd).max() / np.abs(b[k]).max()))}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (multi-GPU sharding is "
              "a design note, not a measured path)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    from fiesta_amd import workloads as W
    from fiesta_amd.config import Dims

    cfg0 = W.rt3d_config(args.size)
    t0 = time.time()
    var = W.rt3d_state(cfg0)
    log("state generated in %.1fs (%d^3, nvt=%d)" % (time.time() - t0, args.size, Dims(cfg0).nvt))
    forms = [f for f in args.forms.split(",") if f]
    weights, train_info = None, None
    if any("learned" in f for f in forms):
        if args.load_weights:
            weights = np.load(args.load_weights, allow_pickle=False).astype(np.float32)
            train_info = {"loaded": os.path.basename(args.load_weights)}
        else:
            log("training the learned C-indicator")
            weights, train_info = train_learned()
            log("trained: %s" % train_info)
        if args.save_weights:
            np.save(args.save_weights, weights)
    results = {}
    for f in forms:
        log("form %s" % f)
        results[f] = gpu_leg(args, f, var, keep_state=len(forms) > 1, weights=weights)
    r = results.get("fp64") or results[forms[0]]
    d = r["dims"]
    nvt = d.nvt
    balg = 20 * nvt  # bytes per cell-stage (FP64 RK2 with tmp as a pointer swap)
    # Roofline unit: one Runge-Kutta stage (ghost fill, ghost primitives and
    # the two fused stage kernels), timed live with HIP events on the context
    # stream.  Algorithmic bytes: 20*nvt per cell-stage (FP64 RK2, stage 1
    # reads U and writes U*, stage 2 reads U*, U^n and writes U^{n+1}).
    roof = None
    if r["launches"] > 0:
        avg_ms = r["kernel_ms"] / r["launches"]
        per_launch = d.ncells * balg
        ach = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": r["kernel"],
                "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes": "%d B per cell-stage x %d cells per stage" % (balg, d.ncells)}
        ph = r.get("phases") or []
        if len(ph) == 5:
            names = ("ghost_fill_and_ghost_primitives", "learned_c_fields", "advection_kernel",
                     "viscous_rk_kernel", "ceq_kernel")
            roof["phases_ms"] = {n: round(v / r["launches"], 4) for n, v in zip(names, ph)}
            ns = r["cfg"].ns
            kern = {2: "k_adv3d_lds<%d, false, true>" % ns, 3: "k_vcr3d_ring<%d, false, false, false, false>" % ns,
                    4: "k_ceq3d<false, false>"}
            top = max(kern, key=lambda q: ph[q])
            roof["dominant_kernel"] = {"name": kern[top], "avg_ms": round(ph[top] / r["launches"], 4),
                                       "share_of_stage": round(ph[top] / r["kernel_ms"], 3)}
        tr = measured_traffic(args.size, "fp64")
        if tr:
            roof["traffic"] = tr["bytes_per_stage"]
            roof["traffic_source"] = tr["source"]
        # the FP64 stage kernels' VALU issue time: a floor under the stage
        # time that no memory optimisation removes (bitwise arithmetic)
        ns = r["cfg"].ns
        fl = issue_floor(args.size, ["k_adv3d_lds<%d, false, true>" % ns,
                                     "k_vcr3d_ring<%d, false, false, false, false>" % ns, "k_ceq3d<false, false>"])
        if fl and r["fused"]:
            tot = sum(fl.values())
            roof["valu_issue_floor"] = {
                "ms_per_stage": round(tot, 4), "kernels_ms": {n: round(v, 4) for n, v in fl.items()},
                "stage_over_floor": round(avg_ms / tot, 3),
                "hbm_frac_at_floor": round(per_launch / (tot * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "source": "profiles/sq_counters_rt3d_%d.json (SQ_INSTS_VALU x 4 cycles / 1024 SIMDs at 2.4 GHz)"
                          % args.size}
    else:
        ach = d.ncells * balg * r["value"] * 1e9 / d.ncells / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "reference kernel sequence (whole step, wall clock)"}
    out = {
        "metric": "Gcell-stage/s, 3D 256^3 2-species: FP64 vs FP32+learned-C; % HBM roofline",
        "value": round(r["value"], 4),
        "unit": "Gcell-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D Rayleigh-Taylor %d^3, 2 species, viscous, C-equation, gravity, RK2 (FP64 plain)"
                               % args.size, "cells": d.ncells, "nvt": nvt,
                   "path": "fused" if r["fused"] else "reference-sequence"},
        "roofline": roof,
    }
    if "copy_gbs" in r:
        # the attainable ceiling: a tuned streaming copy (2 GiB read + 2 GiB
        # written) timed in the same process
        out["copy_kernel_gbs"] = round(r["copy_gbs"], 1)
        if roof is not None:
            roof["ceiling_copy_gbs"] = round(r["copy_gbs"], 1)
            roof["frac_of_copy_ceiling"] = round(roof["achieved"] / r["copy_gbs"], 4)
    fm = {}
    for f, q in results.items():
        # RK2 forms are compared with fp64, SSP-RK3 forms (suffix _rk3) with fp64_rk3
        rk3 = f.endswith("_rk3")
        r0 = results.get("fp64_rk3" if rk3 else "fp64")
        stage_ms = q["kernel_ms"] / max(q["launches"], 1)
        b_form = 20 * q["dims"].nvt
        e = {"value": round(q["value"], 4), "ms_per_step": round(q["wall"] * 1e3 / args.steps, 4),
             "integrator": "ssp-rk3" if rk3 else "rk2-midpoint", "stage_ms": round(stage_ms, 4),
             "algorithmic_bytes_per_cell_stage": b_form,
             "hbm_frac": round(d.ncells * b_form / (stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if r0 is not None:
            e["speedup_vs_fp64"] = round(q["value"] / r0["value"], 3)
        if r0 is not None and r0 is not q and r0.get("state") is not None and "state" in q:
            e["error_vs_fp64"] = form_errors(q["cfg"], q["state"], r0["state"], r0["cfg"])
        if "learned" in q:
            e["learned"] = q["learned"]
        if len(q.get("phases") or []) == 5 and q["launches"]:
            e["phases_ms"] = [round(v / q["launches"], 4) for v in q["phases"]]
        tr = measured_traffic(args.size, f)
        if tr:
            e["traffic_bytes_per_stage"] = tr["bytes_per_stage"]
        fm[f] = e
    out["forms"] = fm
    if train_info:
        out["learned_training"] = train_info
    if args.accuracy_steps > 0:
        acc = accuracy_leg(args, var, forms, weights)
        for f, e in acc.items():
            if f in fm:
                fm[f]["error_vs_fp64_%d_steps" % args.accuracy_steps] = e
    if not args.no_cpu:
        log("Sod shock tube 1000 x 4 (CPU and GPU)")
        out["sod_c1"] = sod_leg()
    if args.extras:
        log("extras: noise filter and output pack")
        out["extras"] = extras_leg(args, var)
    if args.extras or args.gen3d:
        log("extras: generalized-grid module (gen3d)")
        out.setdefault("extras", {})["gen3d"] = gen3d_leg(args)
    if args.extras or args.terrain:
        log("3D ideal expansion over terrain (gen3d + gen_ext)")
        out.setdefault("extras", {})["terrain_expansion"] = terrain_leg(args)
    if args.extras or args.cart2d:
        log("2D module (expansion)")
        out.setdefault("extras", {})["cart2d_expansion"] = expansion2d_leg(args)
    if args.multiblast or args.extras:
        log("2D multi-blast 2048^2: learned C-indicator against the plain C-equation")
        w2, info2 = train_learned2d()
        log("trained 2D: %s" % info2)
        out.setdefault("extras", {})["multiblast_c4"] = multiblast_leg(args, w2, info2)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_leg(args, args.size)
        out["cpu_baseline"]["value"] = round(out["cpu_baseline"]["value"], 5)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
