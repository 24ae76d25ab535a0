"""CPU restatement of the learned-guess network (test infrastructure only:
tests/ use it to check the device kernels; the product never imports it).

The network has no counterpart in the reference; this file states its
definition directly: periodic 3x3 convolutions on nx*ny two-component
fields, layers (cin, cout) = (8, 32), (32, 32), (32, 32), (32, 2), SiLU
after the first three, flat FP32 parameters per layer W[tap][cout][cin] then
bias[cout], tap = 3*(dy+1) + (dx+1) reading pixel (y+dy, x+dx).  The guess
is base + output * gate with the FP64 BDF2 base
(4/3) y_n - (1/3) y_{n-1} + (2/3) dt (2 f_n - f_{n-1}) and the gate
(2/3) dt (f_n - f_{n-1}).

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


def features(un, fn, fnm1, nu, H, W, dt, s_f, s_g, nu_mean, nu_scale, rb=True):
    """Channels: u, v, dt f_u/s_f, dt f_v/s_f, dt (f_u - f_u,n-1)/s_g, the
    same for v, nu_hat, 0."""
    X = np.zeros((H, W, 8))
    X[..., 0:2] = un.reshape(H, W, 2)
    X[..., 2:4] = fn.reshape(H, W, 2) * (dt / s_f)
    X[..., 4:6] = (fn - fnm1).reshape(H, W, 2) * (dt / s_g)
    X[..., 6] = (nu - nu_mean) / nu_scale
    return bf16(X) if rb else X


def base(un, unm1, fn, fnm1, dt):
    """BDF2 with f(y_{n+1}) extrapolated linearly in time."""
    return (4.0 / 3.0) * un - (1.0 / 3.0) * unm1 + ((2.0 / 3.0) * dt) * (2.0 * fn - fnm1)


def gate(fn, fnm1, dt):
    """The base's extrapolation increment, which scales the output."""
    return ((2.0 / 3.0) * dt) * (fn - fnm1)


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


def loss_and_grad(w, un, unm1, fn, fnm1, unp1, nu, H, W, dt, s_f, s_g, s_o, nu_mean, nu_scale, bf=True):
    X0 = features(un, fn, fnm1, nu, H, W, dt, s_f, s_g, nu_mean, nu_scale, rb=bf)
    out, cache = forward(w, X0, bf)
    t = ((unp1 - base(un, unm1, fn, fnm1, dt)) / s_o).reshape(H, W, 2)
    gs = (gate(fn, fnm1, dt) / s_o).reshape(H, W, 2)
    d = out * gs - t
    if bf:
        d = d.astype(np.float32).astype(np.float64)
    npix = H * W
    loss = float(np.sum(d * d) / (npix * 2))
    dz = 2.0 * d * gs / (npix * 2)
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
