"""ctypes mirror of `fst_config` (include/fiesta_hip.h) and derived sizes.

The fields are the subset of FIESTA's `struct inputConfig`
(src/input.hpp:59-148) that the RHS/RK path reads; defaults follow
executeConfiguration (src/input.cpp:130-443).
"""
import ctypes as C

MAX_SPECIES = 8

BC = {"outflow": 0, "reflective": 1, "noslip": 2, "hydrostatic": 3}
SCHEME = {"weno5": 1, "centered4": 2, "quick": 3}

FST_VAR, FST_TMP1, FST_DVAR, FST_VARX, FST_P, FST_RHO, FST_T, FST_VEL, FST_GRADRHO, FST_GRID, FST_METRICS = range(11)
FST_FP64, FST_FP32_BULK, FST_FP64_FAST = 0, 1, 2


class FstConfig(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32), ("grid", C.c_int32), ("ns", C.c_int32), ("ng", C.c_int32),
        ("nci", C.c_int32), ("ncj", C.c_int32), ("nck", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("R", C.c_double),
        ("gamma", C.c_double * MAX_SPECIES), ("M", C.c_double * MAX_SPECIES), ("mu", C.c_double * MAX_SPECIES),
        ("visc", C.c_int32), ("ceq", C.c_int32), ("buoyancy", C.c_int32), ("scheme", C.c_int32),
        ("kap", C.c_double), ("eps", C.c_double), ("alpha", C.c_double),
        ("gAccel", C.c_double), ("rhoRef", C.c_double),
        ("xPer", C.c_int32), ("yPer", C.c_int32), ("zPer", C.c_int32),
        ("bcL", C.c_int32), ("bcR", C.c_int32), ("bcB", C.c_int32), ("bcT", C.c_int32),
        ("bcH", C.c_int32), ("bcF", C.c_int32),
        ("dt", C.c_double),
        ("precision", C.c_int32), ("learned_c", C.c_int32),
        ("rk_order", C.c_int32),
        ("noise", C.c_int32), ("n_nt", C.c_int32), ("n_mode", C.c_int32),
        ("n_dh", C.c_double), ("n_eta", C.c_double), ("n_coff", C.c_double),
        ("gen_ext", C.c_int32), ("reserved", C.c_int32 * 3),
    ]


def make_config(ndim=3, ns=1, ng=3, nci=8, ncj=8, nck=8, dx=1.0, dy=1.0, dz=1.0, R=8.314462,
                gamma=(1.4,), M=(0.02896,), mu=(1.8e-5,), visc=False, ceq=False, buoyancy=False,
                scheme="weno5", kap=1.23456, eps=1.0, alpha=1.0, g=9.81, rhoRef=0.0,
                xper=False, yper=False, zper=False, bc=("outflow",) * 6, dt=1e-6,
                precision=FST_FP64, learned_c=False, grid=0, rk=2, noise=False, n_dh=0.0, n_eta=0.0,
                n_coff=0.0, n_nt=1, n_mode=1, gen_ext=False):
    """Build an FstConfig.  `bc` lists xmin, xmax, ymin, ymax, zmin, zmax.
    grid=1 selects the generalized-grid modules (gen2d/gen3d), whose cell
    sizes are 1 (src/input.cpp:344-350).  rk: 2 = the reference's RK2
    midpoint, 3 = SSP-RK3 (an added option).  noise, n_*: the noise filter
    of postStep (src/input.cpp:270-277: dh, eta, coff, nt = 1, mode = 1).
    gen_ext: on generalized grids, the cartesian ghost fill and the buoyancy
    term (an option the reference's gen modules do not have)."""
    c = FstConfig()
    c.ndim, c.grid, c.ns, c.ng = ndim, int(grid), ns, ng
    c.nci, c.ncj, c.nck = nci, ncj, (nck if ndim == 3 else 1)
    if grid == 1:
        dx = dy = dz = 1.0
    c.dx, c.dy = dx, dy
    c.dz = dz if ndim == 3 else dx  # src/input.cpp:356
    c.R = R
    assert len(gamma) == ns and len(M) == ns and len(mu) == ns
    for s in range(ns):
        c.gamma[s], c.M[s], c.mu[s] = gamma[s], M[s], mu[s]
    c.visc, c.ceq, c.buoyancy = int(visc), int(ceq), int(buoyancy)
    c.scheme = SCHEME[scheme] if isinstance(scheme, str) else int(scheme)
    c.kap, c.eps, c.alpha = kap, eps, alpha
    c.gAccel, c.rhoRef = g, rhoRef
    c.xPer, c.yPer, c.zPer = int(xper), int(yper), int(zper if ndim == 3 else 0)
    b = [BC[x] if isinstance(x, str) else int(x) for x in bc]
    c.bcL, c.bcR, c.bcB, c.bcT, c.bcH, c.bcF = b
    c.dt = dt
    c.precision = precision
    c.learned_c = int(learned_c)
    c.rk_order = int(rk)
    c.noise, c.n_nt, c.n_mode = int(noise), int(n_nt), int(n_mode)
    c.n_dh, c.n_eta, c.n_coff = n_dh, n_eta, n_coff
    c.gen_ext = int(gen_ext)
    return c


class Dims:
    """Sizes derived as src/input.cpp:354-392 derives them."""

    def __init__(self, c):
        self.ndim, self.ng, self.ns = c.ndim, c.ng, c.ns
        self.nci, self.ncj = c.nci, c.ncj
        self.nck = c.nck if c.ndim == 3 else 1
        self.ngi = c.nci + 2 * c.ng
        self.ngj = c.ncj + 2 * c.ng
        self.ngk = self.nck + 2 * c.ng if c.ndim == 3 else 1
        self.nv = (4 if c.ndim == 3 else 3) + c.ns
        # learned C-indicator: no C-equation variables are stored
        self.nvt = self.nv + (5 if (c.ceq and not c.learned_c) else 0)
        self.nvx = c.ndim + 3
        self.N = self.ngi * self.ngj * self.ngk
        self.ncells = self.nci * self.ncj * self.nck
        # grid nodes, src/input.cpp:401-406
        self.ni, self.nj = self.nci + 1, self.ncj + 1
        self.nk = self.nck + 1 if c.ndim == 3 else 1

    def grid_shape(self):
        """numpy shape (v, k, j, i) of the LayoutLeft (ni,nj,nk,3) node array."""
        return (3, self.nk, self.nj, self.ni)

    def metrics_shape(self):
        """numpy shape (n, m, k, j, i) of metrics(i,j,k,m,n)."""
        return (self.ndim, self.ndim, self.ngk, self.ngj, self.ngi)

    def shape(self, nvar):
        """numpy shape (v, k, j, i) of a LayoutLeft (i,j,k,v) array."""
        return (nvar, self.ngk, self.ngj, self.ngi)

    def interior(self):
        g = self.ng
        if self.ndim == 3:
            return (slice(None), slice(g, g + self.nck), slice(g, g + self.ncj), slice(g, g + self.nci))
        return (slice(None), slice(None), slice(g, g + self.ncj), slice(g, g + self.nci))
