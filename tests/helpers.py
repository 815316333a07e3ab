"""Shared test helpers: configurations and bitwise comparison."""
import numpy as np

from fiesta_amd.config import make_config


def assert_bitwise(got, want, what="", nan_ok=False):
    """Equal as IEEE values (+0 == -0); report the worst mismatch.

    nan_ok: NaN allowed where both sides hold NaN (the reference evaluates the
    EOS over zero ghosts before the first ghost fill in 3D)."""
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    if nan_ok:
        both = np.isnan(got) & np.isnan(want)
        got = np.where(both, 0.0, got)
        want = np.where(both, 0.0, want)
    assert np.isfinite(want).all(), "oracle produced non-finite values in %s" % what
    bad = got != want
    if bad.any():
        idx = np.argwhere(bad)
        den = np.maximum(np.abs(want), 1e-300)
        rel = np.abs(got - want) / den
        w = np.unravel_index(np.argmax(np.where(bad, rel, 0)), got.shape)
        raise AssertionError("%s: %d of %d values differ; first at %s; worst rel %.3e at %s (got %r want %r)"
                             % (what, bad.sum(), bad.size, tuple(idx[0]), rel[w], w, got[w], want[w]))


def cfg3d(nci=12, ncj=10, nck=8, ns=2, visc=True, ceq=True, buoyancy=True, bc=None, per=(True, False, True),
          dt=2e-7):
    if bc is None:
        bc = ("outflow", "outflow", "hydrostatic", "hydrostatic", "outflow", "outflow")
    gam = (1.4, 1.67, 1.092)[:ns]
    M = (0.02896, 0.004, 0.14606)[:ns]
    mu = (1.8e-5, 2.1e-5, 1.5e-5)[:ns]
    return make_config(ndim=3, ns=ns, nci=nci, ncj=ncj, nck=nck, dx=1.0e-2, dy=1.1e-2, dz=0.9e-2,
                       gamma=gam, M=M, mu=mu, visc=visc, ceq=ceq, buoyancy=buoyancy, g=9.81, rhoRef=0.3,
                       kap=10.0, eps=5.0, alpha=20.0, xper=per[0], yper=per[1], zper=per[2], bc=bc, dt=dt)


def cfg2d(nci=16, ncj=8, ns=2, visc=True, ceq=True, buoyancy=True, scheme="weno5", bc=None,
          per=(False, True), dt=2e-7):
    if bc is None:
        bc = ("outflow", "reflective", "outflow", "outflow", "outflow", "outflow")
    gam = (1.4, 1.67, 1.092)[:ns]
    M = (0.02896, 0.004, 0.14606)[:ns]
    mu = (1.8e-5, 2.1e-5, 1.5e-5)[:ns]
    return make_config(ndim=2, ns=ns, nci=nci, ncj=ncj, dx=1.0e-2, dy=1.2e-2, gamma=gam, M=M, mu=mu,
                       visc=visc, ceq=ceq, buoyancy=buoyancy, g=9.81, rhoRef=0.3, kap=10.0, eps=1.0,
                       alpha=2.0, scheme=scheme, xper=per[0], yper=per[1], bc=bc, dt=dt)
