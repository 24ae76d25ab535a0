"""CPU pinning of the learned-guess network's oracle: its FP64 gradient
against central finite differences of its FP64 loss, and the BF16 rounding
helper against exact cases."""
import numpy as np

import guessnet_oracle as G


def test_bf16_rounding():
    # 1 + 2^-8 is exactly between 1 and 1 + 2^-7: ties to even -> 1
    assert G.bf16(1.0 + 2.0 ** -8) == 1.0
    assert G.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert G.bf16(-2.5) == -2.5
    x = np.random.default_rng(0).standard_normal(1000)
    r = G.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 + 1e-300)


SPREADS = (1e-3, 1e-4, 1e-5, 1e-6)


def _hist(rng, n, scale=0.1):
    """f_n, ..., f_{n-4}: a smooth random history"""
    f0 = rng.standard_normal(n)
    return [f0 + scale * k * rng.standard_normal(n) for k in range(5)]


COSTS = (9.0, 4.0, 1.0, 0.0, 2.5, 3.0, 7.0)


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    H, W = 6, 5
    w = rng.standard_normal(G.num_params()) * 0.2
    f = _hist(rng, H * W * 2)
    args = (f, 0.004, COSTS, H, W, 5e-3, (1e-2, 1e-2, 1e-2, 1e-2), 5.5e-3, 4.5e-3)
    _, _, g = G.loss_and_grad(w, *args, bf=False)
    assert np.abs(g).max() > 1e-3
    idx = rng.choice(len(w), size=60, replace=False)
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (G.loss_and_grad(wp, *args, bf=False)[0] - G.loss_and_grad(wm, *args, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_oracle_forward_closed_form():
    # with only the centre tap of the first layer's channel 0 -> 0 set and
    # identity-like later layers, the output is silu(silu(silu(x0))), x0 =
    # dt (f_n - f_{n-1}) / s_1 of the u component; the policy's logits are
    # its grid means (channel 0) and zeros, p = softmax(-m / TAU)
    H, W = 4, 4
    w = np.zeros(G.num_params())
    P = []
    o = 0
    for ci, co in G.LAYERS:
        P.append(o)
        o += 9 * co * ci + co
    w[P[0] + (4 * 32 + 0) * 8 + 0] = 1.0
    for li in (1, 2):
        w[P[li] + (4 * 32 + 0) * 32 + 0] = 1.0
    w[P[3] + (4 * 7 + 0) * 32 + 0] = 1.0
    rng = np.random.default_rng(1)
    f = _hist(rng, H * W * 2, 1.0)
    X0 = G.features(f, 0.0055, H, W, 5e-3, SPREADS, 5.5e-3, 4.5e-3, rb=False)
    out, cache = G.forward(w, X0, bf=False)
    s = lambda z: z / (1 + np.exp(-z))  # noqa: E731
    x0 = ((f[0] - f[1]) * (5e-3 / SPREADS[0])).reshape(H, W, 2)[..., 0]
    expect = s(s(s(x0)))
    assert np.allclose(out[..., 0], expect, rtol=1e-14, atol=1e-14)
    assert np.allclose(out[..., 1:], 0.0)
    # the estimates from the last hidden layer's channel means (tap sums of
    # the output weights) equal the output's grid means
    m = G.estimates(w, cache[-1][0], bf=False)
    assert np.allclose(m, out.reshape(-1, 7).mean(axis=0), rtol=1e-13, atol=1e-15)
    t, p = G.policy(m)
    e0 = np.exp(-expect.mean() / G.TAU)
    assert np.allclose(p, np.r_[e0, np.ones(6)] / (e0 + 6.0), rtol=1e-13)
    assert np.isclose(t, (e0 * 1.0 + sum(G.ORDERS[1:])) / (e0 + 6.0), rtol=1e-13)
    # channel 7: |dt D^4 f| / s_4 over the two components
    d4 = (f[0] - 4 * f[1] + 6 * f[2] - 4 * f[3] + f[4]).reshape(H, W, 2) * (5e-3 / SPREADS[3])
    assert np.allclose(X0[..., 7], np.hypot(d4[..., 0], d4[..., 1]), rtol=1e-14)


def test_predictor_order_is_exact_for_polynomial_velocities():
    # the order-k extrapolation of f from f_n, ..., f_{n-k} is exact when f
    # is a polynomial of degree k in t, so P_k equals the BDF2 relation
    # (4/3) y_n - (1/3) y_{n-1} + (2/3) dt f(t_{n+1}); one degree more is not
    rng = np.random.default_rng(4)
    yn, ynm1 = rng.standard_normal((2, 40))
    for dt, tn in ((5e-3, 0.3), (0.1, 1.7), (0.5, -2.0)):
        for k in (1, 2, 3, 4):
            c = rng.standard_normal((k + 2, 40))
            fk = lambda t: sum(c[j] * t ** j for j in range(k + 1))  # noqa: E731
            f = [fk(tn - j * dt) for j in range(5)]
            want = (4.0 / 3.0) * yn - (1.0 / 3.0) * ynm1 + (2.0 / 3.0) * dt * fk(tn + dt)
            got = G.predictor(yn, ynm1, f, dt, k)
            assert np.allclose(got, want, rtol=1e-10, atol=1e-10), (k, dt)
            fk1 = lambda t: fk(t) + c[k + 1] * t ** (k + 1)  # noqa: E731
            f1 = [fk1(tn - j * dt) for j in range(5)]
            want1 = (4.0 / 3.0) * yn - (1.0 / 3.0) * ynm1 + (2.0 / 3.0) * dt * fk1(tn + dt)
            assert np.all(np.abs(G.predictor(yn, ynm1, f1, dt, k) - want1) > 1e-12 * dt ** (k + 1)), (k, dt)
    # P(t): the order-k predictor at integer t, the mean of neighbours at
    # half orders, clamped to [1, 4]
    f = [rng.standard_normal(40) for _ in range(5)]
    P = {k: G.predictor(yn, ynm1, f, 0.1, k) for k in (1, 2, 3, 4)}
    for k in (1, 2, 3, 4):
        assert np.allclose(G.predictor_t(yn, ynm1, f, 0.1, float(k)), P[k], rtol=1e-13, atol=1e-13)
    for k in (1, 2, 3):
        assert np.allclose(G.predictor_t(yn, ynm1, f, 0.1, k + 0.5), 0.5 * (P[k] + P[k + 1]), atol=1e-13)
        assert np.allclose(G.predictor_t(yn, ynm1, f, 0.1, k + 0.25), 0.75 * P[k] + 0.25 * P[k + 1], atol=1e-13)
    assert np.allclose(G.predictor_t(yn, ynm1, f, 0.1, 7.0), P[4], atol=1e-13)
    assert np.allclose(G.predictor_t(yn, ynm1, f, 0.1, 0.0), P[1], atol=1e-13)


def test_cost_estimate_loss():
    # zero weights: estimates m = 0, uniform policy, t = 2.5, loss = mean
    # of the centred costs squared, d loss / d m_j = -2 (c_j - mean c) / 7;
    # a constant added to every cost changes nothing; estimates equal to
    # the centred costs give loss 0 and the soft arg-min of the costs
    rng = np.random.default_rng(6)
    H, W = 4, 4
    f = _hist(rng, H * W * 2)
    w = np.zeros(G.num_params())
    geo = (H, W, 5e-3, SPREADS, 5.5e-3, 4.5e-3)
    c = np.asarray(COSTS)
    cc = c - c.mean()
    loss, out, g = G.loss_and_grad(w, f, 0.004, COSTS, *geo, bf=False)
    assert abs(loss - (cc @ cc) / 7.0) < 1e-13
    t, p = G.policy(out.reshape(-1, 7).mean(axis=0))
    assert abs(t - 2.5) < 1e-14 and np.allclose(p, 1.0 / 7.0)
    # the output bias gradient is the estimate gradient (dZ is uniform / npix)
    assert np.allclose(g[-7:], -2.0 * cc / 7.0, atol=1e-14)
    loss2, _, g2 = G.loss_and_grad(w, f, 0.004, c + 5.0, *geo, bf=False)
    assert abs(loss2 - loss) < 1e-12 and np.allclose(g2, g, atol=1e-14)
    w[-7:] = cc
    loss, out, g = G.loss_and_grad(w, f, 0.004, COSTS, *geo, bf=False)
    assert loss < 1e-26 and np.abs(g).max() < 1e-13
    t, p = G.policy(out.reshape(-1, 7).mean(axis=0))
    e = np.exp(-(c - c.min()) / G.TAU)
    assert np.allclose(p, e / e.sum(), rtol=1e-13)
    assert abs(t - 2.5) < 0.01  # the cheapest order, 2.5 (costs 1 and 2.5 beside it)


def test_estimates_equal_output_means_for_random_weights():
    # the shift invariance of the grid mean under periodic padding, for a
    # random network: mean_p conv(Y3)[p] = bias + (sum_tap W) mean(Y3)
    rng = np.random.default_rng(8)
    H, W = 6, 8
    w = rng.standard_normal(G.num_params()) * 0.3
    f = _hist(rng, H * W * 2, 1.0)
    X0 = G.features(f, 0.003, H, W, 5e-3, (5e-3,) * 4, 5.5e-3, 4.5e-3, rb=False)
    out, cache = G.forward(w, X0, bf=False)
    m = G.estimates(w, cache[-1][0], bf=False)
    assert np.allclose(m, out.reshape(-1, 7).mean(axis=0), rtol=1e-12, atol=1e-13)
