"""Write the golden vectors under tests/golden/ from the oracle.

Each fixture is one .npz holding the configuration (as make_config keyword
arguments, JSON), the input arrays and the oracle's outputs:

- ``var_in`` (and ``grid`` for the generalized modules): the input state;
- ``dvar``: the right-hand side of one compute() after pre_sim;
- ``metrics`` (generalized modules): the metrics of pre_sim;
- ``var_steps``: the state after ``nsteps`` rkAdvance steps (preStep, two
  stages, postStep) from the same input.

manifest.json lists every file with its sha256, so a changed fixture is seen.
The GPU tests compare the HIP path with these arrays bit for bit, and the CPU
tests check that the oracle still reproduces them.

Run from the repository root after building the oracle:
    python3 tests/golden/make_golden.py
The files are written with fixed zip timestamps: an unchanged oracle
reproduces them byte for byte.
"""
import hashlib
import io
import json
import os
import sys
import zipfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fiesta_amd import workloads as W  # noqa: E402
from fiesta_amd.config import FST_DVAR, FST_GRID, FST_METRICS, FST_VAR, make_config  # noqa: E402

SPECIES = dict(gamma=(1.4, 1.67, 1.092), M=(0.02896, 0.004, 0.14606), mu=(1.8e-5, 2.1e-5, 1.5e-5))


def _species(ns):
    return {k: list(v[:ns]) for k, v in SPECIES.items()}


def _cart3d(**kw):
    c = dict(ndim=3, ns=2, nci=8, ncj=8, nck=8, dx=1.0e-2, dy=1.1e-2, dz=0.9e-2, visc=True, ceq=True,
             buoyancy=True, g=9.81, rhoRef=0.3, kap=10.0, eps=5.0, alpha=20.0, xper=True, yper=False, zper=True,
             bc=["outflow", "outflow", "hydrostatic", "hydrostatic", "outflow", "outflow"], dt=2e-7)
    c.update(kw)
    c.update(_species(c["ns"]))
    return c


def _cart2d(**kw):
    c = dict(ndim=2, ns=2, nci=16, ncj=8, dx=1.0e-2, dy=1.2e-2, visc=True, ceq=True, buoyancy=True, g=9.81,
             rhoRef=0.3, kap=10.0, eps=1.0, alpha=2.0, scheme="weno5", xper=False, yper=True,
             bc=["outflow", "reflective", "outflow", "outflow", "outflow", "outflow"], dt=2e-7)
    c.update(kw)
    c.update(_species(c["ns"]))
    return c


def _gen(ndim, n, ns, scheme="weno5"):
    c = dict(ndim=ndim, ns=ns, nci=n[0], ncj=n[1], nck=n[2], grid=1, R=8.314462, scheme=scheme, dt=2e-7)
    c.update(_species(ns))
    return c


# name: (make_config kwargs, input seed, steps, description)
CASES = {
    "cart3d_rt_ns2": (_cart3d(), 101, 3,
                      "3D 8^3, 2 species, viscous, C-equation, gravity; x/z periodic, y hydrostatic walls"),
    "cart3d_walls_ns3": (_cart3d(ns=3, nci=7, ncj=6, nck=5, xper=False, zper=False,
                                 bc=["outflow", "reflective", "hydrostatic", "noslip", "reflective", "outflow"]),
                         102, 3, "3D 7x6x5 ragged, 3 species, every boundary type, no periodic direction"),
    "cart3d_inviscid_ns1": (_cart3d(ns=1, visc=False, ceq=False, buoyancy=False, yper=True), 103, 4,
                            "3D 8^3 single species, Euler only, periodic in all directions"),
    "cart2d_weno_ns2": (_cart2d(), 104, 4, "2D 16x8, 2 species, WENO5, viscous, C-equation, gravity"),
    "cart2d_quick_ns1": (_cart2d(ns=1, scheme="quick", xper=True,
                                 bc=["reflective", "noslip", "hydrostatic", "outflow", "outflow", "outflow"]),
                         105, 4, "2D 16x8 single species, QUICK, walls in y"),
    "cart2d_centered4_ns3": (_cart2d(ns=3, scheme="centered4", xper=True, yper=True), 106, 4,
                             "2D 16x8, 3 species, centered4, periodic"),
    "gen3d_terrain_ns2": (_gen(3, (10, 8, 6), 2), 107, 3,
                          "gen3d on a terrain-following, vertically stretched grid"),
    "gen2d_terrain_ns2": (_gen(2, (12, 10, 1), 2), 108, 3, "gen2d on a terrain-following grid"),
    "cart3d_rt_ns2_ssp_rk3": (_cart3d(rk=3), 109, 3,
                              "as cart3d_rt_ns2 with the SSP-RK3 option (rk_order = 3) in place of RK2"),
}


def config(kw):
    kw = dict(kw)
    for k in ("gamma", "M", "mu", "bc"):
        if k in kw:
            kw[k] = tuple(kw[k])
    return make_config(**kw)


def inputs(name):
    """The input arrays of a case (var_in, grid or None), from its seed."""
    kw, seed, _, _ = CASES[name]
    cfg = config(kw)
    if kw.get("grid", 0) == 1:
        return W.random_state_all(cfg, seed=seed, umag=20.0), W.terrain_grid(cfg)
    return W.random_state(cfg, seed), None


def run_oracle(name, var_in, grid):
    """Outputs of the oracle for one case: dict of arrays."""
    import oracle
    kw, _, nsteps, _ = CASES[name]
    cfg = config(kw)
    out = {}
    for what in ("compute", "steps"):
        o = oracle.Oracle(cfg)
        if grid is not None:
            o.set(FST_GRID, grid)
        o.set(FST_VAR, var_in)
        o.pre_sim()
        if what == "compute":
            if grid is not None:
                out["metrics"] = o.get(FST_METRICS)
            o.compute()
            out["dvar"] = o.get(FST_DVAR)
        else:
            o.step(nsteps)
            out["var_steps"] = o.get(FST_VAR)
    return out


def write_npz(path, arrays):
    """An .npz (numpy's zip of .npy members, deflated) with fixed member
    timestamps, so that the same arrays give the same bytes."""
    with zipfile.ZipFile(path, "w", compression=zipfile.ZIP_DEFLATED) as z:
        for name in sorted(arrays):
            buf = io.BytesIO()
            np.lib.format.write_array(buf, np.asanyarray(arrays[name]), allow_pickle=False)
            info = zipfile.ZipInfo(name + ".npy", date_time=(1980, 1, 1, 0, 0, 0))
            info.compress_type = zipfile.ZIP_DEFLATED
            z.writestr(info, buf.getvalue())


def sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        h.update(f.read())
    return h.hexdigest()


def main():
    manifest = {"format": "npz (numpy arrays, float64, shape (v, k, j, i) as the C-ABI's array layout)",
                "generator": "tests/golden/make_golden.py (oracle/fiesta_oracle.cpp)", "cases": {}}
    for name, (kw, seed, nsteps, desc) in CASES.items():
        var_in, grid = inputs(name)
        out = run_oracle(name, var_in, grid)
        arrays = dict(var_in=var_in, **out)
        if grid is not None:
            arrays["grid"] = grid
        for k, a in arrays.items():
            assert np.isfinite(a).all(), (name, k)
        path = os.path.join(HERE, name + ".npz")
        write_npz(path, dict(config=np.array(json.dumps(kw, sort_keys=True)), **arrays))
        manifest["cases"][name] = {"file": name + ".npz", "sha256": sha256(path), "seed": seed, "steps": nsteps,
                                   "description": desc, "arrays": sorted(arrays)}
        print("%-24s %8d bytes" % (name, os.path.getsize(path)))
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
