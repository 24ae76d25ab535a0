"""CPU restatement of the learned-guess network (test infrastructure only:
tests/ use it to check the device kernels; the product never imports it).

The network has no counterpart in the reference; this file states its
definition directly: periodic 3x3 convolutions on nx*ny two-component
fields, layers (cin, cout) = (8, 32), (32, 32), (32, 32), (32, 2), SiLU
after the first three, flat FP32 parameters per layer W[tap][cout][cin] then
bias[cout], tap = 3*(dy+1) + (dx+1) reading pixel (y+dy, x+dx).

`bf16=True` rounds to BF16 (round to nearest even) where the device stores
BF16: features, weights, hidden activations and the backward's dZ; all
arithmetic is FP64.  `bf16=False` is the exact FP64 network, whose gradient
the finite-difference test pins.
"""
import numpy as np

LAYERS = [(8, 32), (32, 32), (32, 32), (32, 2)]


def bf16(x):
    """Round to the nearest BF16 (ties to even), via float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    b = a.view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64).reshape(a.shape)


def num_params():
    return sum(9 * ci * co + co for ci, co in LAYERS)


def unpack(w):
    out, o = [], 0
    for ci, co in LAYERS:
        W = np.asarray(w[o:o + 9 * co * ci], dtype=np.float64).reshape(9, co, ci)
        o += 9 * co * ci
        b = np.asarray(w[o:o + co], dtype=np.float64)
        o += co
        out.append((W, b))
    return out


def _shift(X, t):
    dy, dx = t // 3 - 1, t % 3 - 1
    return np.roll(X, shift=(-dy, -dx), axis=(0, 1))  # [y][x] <- X[y+dy][x+dx]


def _unshift(X, t):
    dy, dx = t // 3 - 1, t % 3 - 1
    return np.roll(X, shift=(dy, dx), axis=(0, 1))


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def features(un, unm1, nu, H, W, s_d, nu_mean, nu_scale, rb=True, fn=None, dt_over_sf=0.0):
    """Channels: u, v, du/s_d, dv/s_d, dt f_u/s_f, dt f_v/s_f, nu_hat, 0."""
    u = un.reshape(H, W, 2)
    d = (un - unm1).reshape(H, W, 2) / s_d
    X = np.zeros((H, W, 8))
    X[..., 0:2] = u
    X[..., 2:4] = d
    if fn is not None:
        X[..., 4:6] = fn.reshape(H, W, 2) * dt_over_sf
    X[..., 6] = (nu - nu_mean) / nu_scale
    return bf16(X) if rb else X


def forward(w, X0, bf=True):
    """Returns (output [H,W,2], cache) for the backward."""
    P = unpack(w)
    X = X0
    cache = []
    for li, (Wl, bl) in enumerate(P):
        Wr = bf16(Wl) if bf else Wl
        z = np.zeros(X.shape[:2] + (Wl.shape[1],))
        for t in range(9):
            z += _shift(X, t) @ Wr[t].T
        z += bl
        cache.append((X, Wr, z))
        if li < len(P) - 1:
            X = _silu(z)
            if bf:
                X = bf16(X)
    return z, cache


def loss_and_grad(w, un, unm1, unp1, nu, H, W, s_d, s_o, nu_mean, nu_scale, bf=True, fn=None,
                  dt_over_sf=0.0):
    X0 = features(un, unm1, nu, H, W, s_d, nu_mean, nu_scale, rb=bf, fn=fn, dt_over_sf=dt_over_sf)
    out, cache = forward(w, X0, bf)
    t = ((unp1 - 2.0 * un + unm1) / s_o).reshape(H, W, 2)
    if bf:
        t = t.astype(np.float32).astype(np.float64)
    d = out - t
    npix = H * W
    loss = float(np.sum(d * d) / (npix * 2))
    dz = 2.0 * d / (npix * 2)
    if bf:
        dz = bf16(dz)
    grads = []
    for li in range(len(LAYERS) - 1, -1, -1):
        X, Wr, z = cache[li]
        ci, co = LAYERS[li]
        dW = np.zeros((9, co, ci))
        for t_ in range(9):
            dW[t_] = dz.reshape(-1, co).T @ _shift(X, t_).reshape(-1, ci)
        db = dz.reshape(-1, co).sum(axis=0)
        grads.append(np.concatenate([dW.ravel(), db]))
        if li == 0:
            break
        dX = np.zeros(X.shape)
        for t_ in range(9):
            dX += _unshift(dz, t_) @ Wr[t_]
        zprev = cache[li - 1][2]
        dz = dX * _dsilu(zprev)
        if bf:
            dz = bf16(dz)
    return loss, out, np.concatenate(grads[::-1])
