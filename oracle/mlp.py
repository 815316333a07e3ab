"""TEST INFRASTRUCTURE ONLY -- numpy restatement of the learned C-indicator.

The network (5x5x5 stencil x 3 features in 3D, 5x5 x 3 in 2D -> 64 -> 64 ->
64 -> 1, GELU tanh form on the hidden layers, mean-squared-error loss; the
dimension follows from the parameter count), its gradient, the
inference-side residual check (the resting C-equation relaxation) with its
fallback to the plain method's local solution, and the ghost layer of C_hat,
all in float64.  With `bf16=True` the operands of every matrix product are
rounded to bfloat16 at the points where the GPU kernels round them (inputs,
weights W1..W3, hidden activations, the back-propagated dZ), so GPU results
can be compared with tolerances that cover only FP32 accumulation order.

Pinned in tests/test_oracle_mlp.py: the float64 gradient against central
finite differences, GELU against its closed form, and bf16 rounding against
exactly representable values.  Only tests/ use this module.
"""
import numpy as np

NH = 64


def layout(nd):
    """Offsets of the parameter blocks of the nd network (W1[NIN][64], b1,
    W2[64][64], b2, W3[64][64], b3, w4[64], b4) and NIN, NPARAM."""
    nin = 3 * 5 ** nd
    o = {"NIN": nin, "OFF_W1": 0}
    o["OFF_B1"] = nin * NH
    o["OFF_W2"] = o["OFF_B1"] + NH
    o["OFF_B2"] = o["OFF_W2"] + NH * NH
    o["OFF_W3"] = o["OFF_B2"] + NH
    o["OFF_B3"] = o["OFF_W3"] + NH * NH
    o["OFF_W4"] = o["OFF_B3"] + NH
    o["OFF_B4"] = o["OFF_W4"] + NH
    o["NPARAM"] = o["OFF_B4"] + 1
    return o


_L3, _L2 = layout(3), layout(2)
NIN = _L3["NIN"]
OFF_W1, OFF_B1, OFF_W2, OFF_B2 = _L3["OFF_W1"], _L3["OFF_B1"], _L3["OFF_W2"], _L3["OFF_B2"]
OFF_W3, OFF_B3, OFF_W4, OFF_B4, NPARAM = _L3["OFF_W3"], _L3["OFF_B3"], _L3["OFF_W4"], _L3["OFF_B4"], _L3["NPARAM"]
NPARAM2 = _L2["NPARAM"]


def _lay(w):
    n = np.asarray(w).size
    if n == _L3["NPARAM"]:
        return _L3
    if n == _L2["NPARAM"]:
        return _L2
    raise ValueError("parameter count %d is neither the 3D (%d) nor the 2D (%d) network's" %
                     (n, _L3["NPARAM"], _L2["NPARAM"]))

_K0 = 0.7978845608028654
_K1 = 0.044715


def bf16(x):
    """Round to bfloat16 (round to nearest even on the float32 bit pattern)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64).reshape(a.shape)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(_K0 * (x + _K1 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(_K0 * (x + _K1 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * _K0 * (1.0 + 3.0 * _K1 * x * x)


def unpack(w):
    L = _lay(w)
    w = np.asarray(w, dtype=np.float64)
    return (w[L["OFF_W1"]:L["OFF_B1"]].reshape(L["NIN"], NH), w[L["OFF_B1"]:L["OFF_W2"]],
            w[L["OFF_W2"]:L["OFF_B2"]].reshape(NH, NH), w[L["OFF_B2"]:L["OFF_W3"]],
            w[L["OFF_W3"]:L["OFF_B3"]].reshape(NH, NH), w[L["OFF_B3"]:L["OFF_W4"]], w[L["OFF_W4"]:L["OFF_B4"]],
            w[L["OFF_B4"]])


def forward(w, X, bf=False):
    """y [M] and the cache for backward()."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(w)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    X = r(X)
    Z1 = X @ r(W1) + b1
    H1 = r(gelu(Z1))
    Z2 = H1 @ r(W2) + b2
    H2 = r(gelu(Z2))
    Z3 = H2 @ r(W3) + b3
    H3 = r(gelu(Z3))
    y = H3 @ w4 + b4
    return y, (X, Z1, H1, Z2, H2, Z3, H3)


def loss_grad(w, X, T, bf=False):
    """Mean squared error, its gradient (parameter-vector layout) and y."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(w)
    L = _lay(w)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    y, (Xr, Z1, H1, Z2, H2, Z3, H3) = forward(w, X, bf)
    M = y.size
    e = y - np.asarray(T, np.float64)
    loss = float((e * e).mean())
    dY = 2.0 * e / M
    g = np.zeros(L["NPARAM"])
    g[L["OFF_W4"]:L["OFF_B4"]] = H3.T @ dY
    g[L["OFF_B4"]] = dY.sum()
    dZ3 = dY[:, None] * w4[None, :] * gelu_grad(Z3)
    g[L["OFF_B3"]:L["OFF_W4"]] = dZ3.sum(axis=0)
    g[L["OFF_W3"]:L["OFF_B3"]] = (H2.T @ r(dZ3)).ravel()
    dZ2 = (r(dZ3) @ r(W3).T) * gelu_grad(Z2)
    g[L["OFF_B2"]:L["OFF_W3"]] = dZ2.sum(axis=0)
    g[L["OFF_W2"]:L["OFF_B2"]] = (H1.T @ r(dZ2)).ravel()
    dZ1 = (r(dZ2) @ r(W2).T) * gelu_grad(Z1)
    g[L["OFF_B1"]:L["OFF_W2"]] = dZ1.sum(axis=0)
    g[L["OFF_W1"]:L["OFF_B1"]] = (Xr.T @ r(dZ1)).ravel()
    return loss, g, y


def stencils(F, cells, dims):
    """Network inputs of interior cells: F [3, ngk, ngj, ngi]; cells = flat
    (k,j,i) indices; input k-index = ch*125 + (dz+2)*25 + (dy+2)*5 + (dx+2)
    (2D: ch*25 + (dy+2)*5 + (dx+2))."""
    Ff = np.asarray(F, np.float64).reshape(3, -1)
    s1, s2 = dims.ngi, dims.ngi * dims.ngj
    dzs = range(-2, 3) if dims.ndim == 3 else (0,)
    offs = np.array([dx + dy * s1 + dz * s2 for dz in dzs for dy in range(-2, 3) for dx in range(-2, 3)])
    idx = np.asarray(cells)[:, None] + offs[None, :]
    return np.concatenate([Ff[c][idx] for c in range(3)], axis=1)


def active_cells(M, dims, tau):
    """Interior cells whose 3x3x3 (2D: 3x3) neighbourhood has M > tau (M =
    max |feature| in 3D, the contact source G in 2D)."""
    g = dims.ng
    Mp = np.asarray(M, np.float32)
    if dims.ndim == 2:
        M2 = Mp.reshape(dims.ngj, dims.ngi)
        acc = np.zeros((dims.ngj - 2, dims.ngi - 2), np.float32)
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                acc = np.maximum(acc, M2[1 + dj:dims.ngj - 1 + dj, 1 + di:dims.ngi - 1 + di])
        mx = np.zeros_like(M2)
        mx[1:-1, 1:-1] = acc
        inter = np.zeros_like(M2, dtype=bool)
        inter[g:g + dims.ncj, g:g + dims.nci] = True
        return np.flatnonzero((mx > tau) & inter)
    mx = np.zeros_like(Mp)
    mx[:] = 0.0
    sl = slice(1, -1)
    acc = np.zeros((dims.ngk - 2, dims.ngj - 2, dims.ngi - 2), np.float32)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                acc = np.maximum(acc, Mp[1 + dk:dims.ngk - 1 + dk, 1 + dj:dims.ngj - 1 + dj, 1 + di:dims.ngi - 1 + di])
    mx[sl, sl, sl] = acc
    inter = np.zeros_like(Mp, dtype=bool)
    inter[g:g + dims.nck, g:g + dims.ncj, g:g + dims.nci] = True
    return np.flatnonzero((mx > tau) & inter)


def _chat_ghost(ch, dims, xper, yper, zper):
    """The face ghost layer of C_hat: periodic image or mirror copy."""
    g = dims.ng
    I, J = slice(g, g + dims.nci), slice(g, g + dims.ncj)
    K = slice(g, g + dims.nck) if dims.ndim == 3 else slice(0, 1)
    ch[K, J, g - 1] = ch[K, J, g + dims.nci - 1] if xper else ch[K, J, g]
    ch[K, J, g + dims.nci] = ch[K, J, g] if xper else ch[K, J, g + dims.nci - 1]
    ch[K, g - 1, I] = ch[K, g + dims.ncj - 1, I] if yper else ch[K, g, I]
    ch[K, g + dims.ncj, I] = ch[K, g, I] if yper else ch[K, g + dims.ncj - 1, I]
    if dims.ndim == 2:
        return
    ch[g - 1, J, I] = ch[g + dims.nck - 1, J, I] if zper else ch[g, J, I]
    ch[g + dims.nck, J, I] = ch[g, J, I] if zper else ch[g + dims.nck - 1, J, I]


def residual_check(ch, G, rho, cells, dims, hmin, kap, eps, d, tol):
    """The residual check of the prediction and its fallback, as the GPU pass
    does it (float64 arithmetic on the float32 fields).  The C-equation
    relaxation (src/ceq3d.hpp:189-224) rests where
    r = G - C + a lap(C) = 0, a = kap eps (dx^2 + dy^2 + dz^2),
    lap(c) = sum_d (c[-e_d] - 2c + c[+e_d]) / d_d, G = G1 rho / h (in 2D,
    updateCeq2D src/ceq.hpp:153-195: the sums over x and y).  A cell with
    |r| h / rho > tol max G1 (or a non-finite value around it) takes the
    resting relaxation's local solution with its neighbours held:
    (G + a sum_d (c[-e_d] + c[+e_d]) / d_d) / (1 + 2 a sum_d 1 / d_d).
    Returns (new C_hat values of `cells` as float32, fallback mask, r h / rho)."""
    chf = np.asarray(ch, np.float32).ravel().astype(np.float64)
    Gf = np.asarray(G, np.float32).ravel().astype(np.float64)
    rf = np.asarray(rho, np.float64).ravel()
    g = dims.ng
    nd = dims.ndim
    Gk = np.asarray(G, np.float32).reshape(dims.ngk, dims.ngj, dims.ngi)
    K = slice(g, g + dims.nck) if nd == 3 else slice(0, 1)
    gmax = float(Gk[K, g:g + dims.ncj, g:g + dims.nci].max())
    a = kap * eps * sum(d[q] ** 2 for q in range(nd))
    st = (1, dims.ngi, dims.ngi * dims.ngj)
    c = chf[cells]
    fin = np.isfinite(c)
    lap = np.zeros(len(cells))
    nsum = np.zeros(len(cells))
    for q in range(nd):
        m = chf[cells - st[q]]
        p = chf[cells + st[q]]
        fin &= np.isfinite(m) & np.isfinite(p)
        m = np.where(np.isfinite(m), m, 0.0)
        p = np.where(np.isfinite(p), p, 0.0)
        lap += (m - 2.0 * c + p) * (1.0 / d[q])
        nsum += (m + p) * (1.0 / d[q])
    Gd = Gf[cells] * rf[cells] / hmin
    res = Gd - c + a * lap
    rn = np.abs(res) * hmin / rf[cells]
    with np.errstate(invalid="ignore"):
        ok = fin & (rn <= tol * gmax)
    rsum = 0.0
    for q in range(nd):
        rsum += 1.0 / d[q]
    rden = 1.0 / (1.0 + 2.0 * a * rsum)
    v = np.where(ok, c, (Gd + a * nsum) * rden)
    return v.astype(np.float32), ~ok, rn, tol * gmax


def chat_field(w, F, M, G, rho, dims, hmin, kap, eps, d, tau=1e-3, tol=0.05, xper=False, yper=False, zper=False):
    """C_hat over the grid as the GPU inference computes it (bf16 network):
    C_hat = max(y, 0) rho / h on the active cells, the face ghost layer, the
    residual check with its fallback, the ghost layer again.  Returns
    (chat, y, fallback mask, active cells, normalised residuals, threshold).
    The activity field is M (max |feature|) in 3D and the contact source G in
    2D."""
    cells = active_cells(M if dims.ndim == 3 else G, dims, tau)
    X = stencils(F, cells, dims)
    y, _ = forward(w, X, bf=True)
    rf = np.asarray(rho, np.float64).ravel()
    ch = np.zeros(dims.N, np.float32)
    with np.errstate(invalid="ignore"):
        ch[cells] = np.where(np.isfinite(y), np.maximum(y, 0.0) * rf[cells] / hmin, np.nan).astype(np.float32)
    ch = ch.reshape(dims.ngk, dims.ngj, dims.ngi)
    _chat_ghost(ch, dims, xper, yper, zper)
    v, fb, rn, thr = residual_check(ch, G, rho, cells, dims, hmin, kap, eps, d, tol)
    out = ch.ravel().copy()
    out[cells] = v
    out = out.reshape(dims.ngk, dims.ngj, dims.ngi)
    _chat_ghost(out, dims, xper, yper, zper)
    return out, y, fb, cells, rn, thr
