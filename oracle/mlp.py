"""TEST INFRASTRUCTURE ONLY -- numpy restatement of the learned C-indicator.

The network (5x5x5 stencil x 3 features -> 64 -> 64 -> 64 -> 1, GELU tanh
form on the hidden layers, mean-squared-error loss), its gradient, the
inference-side residual check with its fallback and the ghost layer of C_hat,
all in float64.  With `bf16=True` the operands of every matrix product are
rounded to bfloat16 at the points where the GPU kernels round them (inputs,
weights W1..W3, hidden activations, the back-propagated dZ), so GPU results
can be compared with tolerances that cover only FP32 accumulation order.

Pinned in tests/test_oracle_mlp.py: the float64 gradient against central
finite differences, GELU against its closed form, and bf16 rounding against
exactly representable values.  Only tests/ use this module.
"""
import numpy as np

NIN, NH = 375, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NIN * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NH
OFF_W4 = OFF_B3 + NH
OFF_B4 = OFF_W4 + NH
NPARAM = OFF_B4 + 1

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
    w = np.asarray(w, dtype=np.float64)
    return (w[OFF_W1:OFF_B1].reshape(NIN, NH), w[OFF_B1:OFF_W2], w[OFF_W2:OFF_B2].reshape(NH, NH),
            w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3].reshape(NH, NH), w[OFF_B3:OFF_W4], w[OFF_W4:OFF_B4], w[OFF_B4])


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
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    y, (Xr, Z1, H1, Z2, H2, Z3, H3) = forward(w, X, bf)
    M = y.size
    e = y - np.asarray(T, np.float64)
    loss = float((e * e).mean())
    dY = 2.0 * e / M
    g = np.zeros(NPARAM)
    g[OFF_W4:OFF_B4] = H3.T @ dY
    g[OFF_B4] = dY.sum()
    dZ3 = dY[:, None] * w4[None, :] * gelu_grad(Z3)
    g[OFF_B3:OFF_W4] = dZ3.sum(axis=0)
    g[OFF_W3:OFF_B3] = (H2.T @ r(dZ3)).ravel()
    dZ2 = (r(dZ3) @ r(W3).T) * gelu_grad(Z2)
    g[OFF_B2:OFF_W3] = dZ2.sum(axis=0)
    g[OFF_W2:OFF_B2] = (H1.T @ r(dZ2)).ravel()
    dZ1 = (r(dZ2) @ r(W2).T) * gelu_grad(Z1)
    g[OFF_B1:OFF_W2] = dZ1.sum(axis=0)
    g[OFF_W1:OFF_B1] = (Xr.T @ r(dZ1)).ravel()
    return loss, g, y


def stencils(F, cells, dims):
    """Network inputs of interior cells: F [3, ngk, ngj, ngi]; cells = flat
    (k,j,i) indices; input k-index = ch*125 + (dz+2)*25 + (dy+2)*5 + (dx+2)."""
    Ff = np.asarray(F, np.float64).reshape(3, -1)
    s1, s2 = dims.ngi, dims.ngi * dims.ngj
    offs = np.array([dx + dy * s1 + dz * s2 for dz in range(-2, 3) for dy in range(-2, 3) for dx in range(-2, 3)])
    idx = np.asarray(cells)[:, None] + offs[None, :]
    return np.concatenate([Ff[c][idx] for c in range(3)], axis=1)


def active_cells(M, dims, tau):
    """Interior cells whose 3x3x3 neighbourhood has max |feature| > tau."""
    g = dims.ng
    Mp = np.asarray(M, np.float32)
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


def chat_field(w, F, M, G, rho, dims, hmin, tau=1e-3, tol=0.05, xper=False, yper=False, zper=False):
    """C_hat over the grid as the GPU inference computes it (bf16 network),
    with the residual check / fallback and the face ghost layer.  Returns
    (chat, y, fallback mask, active cells, bound)."""
    cells = active_cells(M, dims, tau)
    X = stencils(F, cells, dims)
    y, _ = forward(w, X, bf=True)
    Gf = np.asarray(G, np.float64).ravel()
    g_inter = np.asarray(G, np.float32)[dims.ng:dims.ng + dims.nck, dims.ng:dims.ng + dims.ncj,
                                        dims.ng:dims.ng + dims.nci]
    gmax = float(g_inter.max()) if g_inter.size else 0.0
    bound = (1.0 + tol) * gmax + tol
    ok = np.isfinite(y) & (y >= -tol) & (y <= bound)
    s1, s2 = dims.ngi, dims.ngi * dims.ngj
    offs = np.array([dx + dy * s1 + dz * s2 for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
    gmean = Gf[cells[:, None] + offs[None, :]].sum(axis=1) / 27.0
    yy = np.where(ok, np.maximum(y, 0.0), gmean)
    ch = np.zeros(dims.N)
    ch[cells] = yy * np.asarray(rho, np.float64).ravel()[cells] / hmin
    ch = ch.reshape(dims.ngk, dims.ngj, dims.ngi)
    g = dims.ng
    I, J, K = slice(g, g + dims.nci), slice(g, g + dims.ncj), slice(g, g + dims.nck)
    ch[K, J, g - 1] = ch[K, J, g + dims.nci - 1] if xper else ch[K, J, g]
    ch[K, J, g + dims.nci] = ch[K, J, g] if xper else ch[K, J, g + dims.nci - 1]
    ch[K, g - 1, I] = ch[K, g + dims.ncj - 1, I] if yper else ch[K, g, I]
    ch[K, g + dims.ncj, I] = ch[K, g, I] if yper else ch[K, g + dims.ncj - 1, I]
    ch[g - 1, J, I] = ch[g + dims.nck - 1, J, I] if zper else ch[g, J, I]
    ch[g + dims.nck, J, I] = ch[g, J, I] if zper else ch[g + dims.nck - 1, J, I]
    return ch, y, ~ok, cells, bound
