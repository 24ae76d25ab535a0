"""CPU restatement of the learned-guess network (test infrastructure only:
tests/ use it to check the device kernels; the product never imports it).

The network has no counterpart in the reference; this file states its
definition directly: periodic 3x3 convolutions on nx*ny two-component
fields, layers (cin, cout) = (8, 32), (32, 32), (32, 32), (32, 7), SiLU
after the first three, flat FP32 parameters per layer W[tap][cout][cin] then
bias[cout], tap = 3*(dy+1) + (dx+1) reading pixel (y+dy, x+dx).

The guess is the FP64 predictor P(t) of y_{n+1} of extrapolation order t:
the BDF2 relation with f(y_{n+1}) extrapolated in time with order k,
P_k = (4/3) y_n - (1/3) y_{n-1} + (2/3) dt E_k(f_n, ..., f_{n-k}), linear in
t between P_floor(t) and P_floor(t)+1.  The grid means m_j of the
network's 7 output channels estimate the step's cost from P(T_j), T_j = 1
+ j/2, relative to their mean; t = sum_j p_j T_j with p = softmax(-m /
TAU).  The loss of a sample with costs c is sum_j (m_j - (c_j - mean
c))^2 / 7.  Inputs: dt D^k f / s_k for k = 1, 2, 3
(D^k the k-th backward time difference), (nu - nu_mean) / nu_scale,
|dt D^4 f| / s_4.

`bf16=True` rounds to BF16 (round to nearest even) where the device stores
BF16: features, weights, hidden activations and the backward's dZ; all
arithmetic is FP64.  `bf16=False` is the exact FP64 network, whose gradient
the finite-difference test pins.
"""
import numpy as np

LAYERS = [(8, 32), (32, 32), (32, 32), (32, 7)]
ORDERS = tuple(1.0 + 0.5 * j for j in range(7))
TAU = 0.25


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


def diffs(f):
    """Backward time differences D^1..D^4 of f = [f_n, ..., f_{n-4}]."""
    f0, f1, f2, f3, f4 = f
    return (f0 - f1, f0 - 2.0 * f1 + f2, f0 - 3.0 * f1 + 3.0 * f2 - f3,
            f0 - 4.0 * f1 + 6.0 * f2 - 4.0 * f3 + f4)


def features(f, nu, H, W, dt, spreads, nu_mean, nu_scale, rb=True):
    """Channels: dt D^k f / s_k for k = 1, 2, 3 (u, v each), nu_hat,
    |dt D^4 f| / s_4 (the norm over the two components)."""
    d = diffs(f)
    X = np.zeros((H, W, 8))
    for k in range(3):
        X[..., 2 * k:2 * k + 2] = (d[k] * (dt / spreads[k])).reshape(H, W, 2)
    X[..., 6] = (nu - nu_mean) / nu_scale
    X[..., 7] = np.sqrt(((d[3] * (dt / spreads[3])).reshape(H, W, 2) ** 2).sum(axis=-1))
    if rb:
        # the device stores each feature as BF16 of its FP32 value
        return bf16(X.astype(np.float32))
    return X


def predictor(un, unm1, f, dt, order):
    """BDF2 with f(y_{n+1}) extrapolated with the given order (1..4) from
    f = [f_n, f_{n-1}, ...]: the binomial extrapolation sum_j (-1)^j
    C(order+1, j+1) f_{n-j}."""
    from math import comb
    E = sum(((-1) ** j) * comb(order + 1, j + 1) * f[j] for j in range(order + 1))
    return (4.0 / 3.0) * un - (1.0 / 3.0) * unm1 + ((2.0 / 3.0) * dt) * E


def extrap_weights(t):
    """Weights of f_n, ..., f_{n-4} in the order-t extrapolation (linear in t
    between the binomial weights of floor(t) and floor(t) + 1)."""
    from math import comb
    E = [[((-1) ** j) * comb(k + 1, j + 1) if j <= k else 0 for j in range(5)] for k in range(1, 5)]
    t = min(max(float(t), 1.0), 4.0)
    k = 3 if t >= 4.0 else int(t)
    w = t - k
    return [(1.0 - w) * E[k - 1][j] + w * E[k][j] for j in range(5)]


def predictor_t(un, unm1, f, dt, t):
    """P(t) as the device forms it."""
    e = extrap_weights(t)
    E = e[0] * f[0] + e[1] * f[1] + e[2] * f[2] + e[3] * f[3] + e[4] * f[4]
    return (4.0 / 3.0) * un - (1.0 / 3.0) * unm1 + ((2.0 / 3.0) * dt) * E


def estimates(w, Y3, bf=True):
    """Grid means m_j of the output channels from the last hidden layer's
    output Y3 [H,W,32]: the output layer is linear and its padding periodic,
    so m_j = bias_j + sum_i (sum_tap W[tap][j][i]) mean(Y3[..., i]), with the
    BF16 weights when bf."""
    Wl, bl = unpack(w)[-1]
    Wr = bf16(Wl) if bf else Wl
    return bl + Wr.sum(axis=0) @ np.asarray(Y3, dtype=np.float64).reshape(-1, Wl.shape[2]).mean(axis=0)


def policy(m):
    """(t, p): p = softmax(-m / TAU) of the cost estimates m, t = sum_j p_j
    T_j."""
    m = np.asarray(m, dtype=np.float64)
    p = np.exp(-(m - m.min()) / TAU)
    p /= p.sum()
    return float(min(max(p @ np.asarray(ORDERS), 1.0), 4.0)), p


def forward(w, X0, bf=True):
    """Returns (output [H,W,7], cache) for the backward."""
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


def loss_and_grad(w, f, nu, costs, H, W, dt, spreads, nu_mean, nu_scale, bf=True):
    """loss = sum_j (m_j - (c_j - mean c))^2 / 7, the network's output and
    the loss's gradient w.r.t. the parameters.  With m the output means
    (estimates()),
    d loss / d m_j = 2 (m_j - (c_j - mean c)) / 7 and d loss / d out[p][j]
    = (d loss / d m_j) / (H W) at every pixel."""
    X0 = features(f, nu, H, W, dt, spreads, nu_mean, nu_scale, rb=bf)
    out, cache = forward(w, X0, bf)
    if bf:
        out = out.astype(np.float32).astype(np.float64)
    c = np.asarray(costs, dtype=np.float64)
    m = estimates(w, cache[-1][0], bf)
    r = m - (c - c.mean())
    loss = float(r @ r / len(ORDERS))
    gm = 2.0 * r / len(ORDERS)
    npix = H * W
    dz = np.broadcast_to(gm / npix, (H, W, len(ORDERS))).copy()
    if bf:
        dz = bf16(dz.astype(np.float32))
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
