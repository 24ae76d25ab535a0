"""The learned-guess network's device kernels (MFMA forward, backward-data,
backward-weight, Adam) against the CPU oracle, which mirrors the device's
BF16 rounding points and computes in FP64; tolerances are BF16-level.  The
policy (softmax of FP64 sums of the output) and the FP64 guess it selects
are checked against the oracle's on the device's own output to 1e-12."""
import numpy as np
import pytest

import guessnet_oracle as G
import pressiodemoapps_amd as pda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no HIP device")
    return _t


DT = 5e-3


def _scales(f):
    """dt, spreads of dt D^k f over the fields (k = 1..4), nu_mean, nu_scale"""
    return (DT, tuple(float(np.std(DT * d)) for d in G.diffs(f)), 5.5e-3, 4.5e-3)


def _weights(rng, n, scale):
    """random parameters, the output layer 10x smaller (as the training's
    initialisation), so that the policy starts unsaturated"""
    w = (rng.standard_normal(n) * scale).astype(np.float32)
    last = 9 * 32 * 7 + 7
    w[-last:] *= 0.1
    return w


def _fields(rng, H, W):
    y, x = np.meshgrid(np.linspace(-1, 1, H, endpoint=False), np.linspace(-1, 1, W, endpoint=False),
                       indexing="ij")
    u = np.stack([np.exp(-((x - 0.2) ** 2 + y ** 2) / 0.1), 0.5 * np.exp(-(x ** 2 + (y + 0.3) ** 2) / 0.2)],
                 axis=-1).ravel()
    un = u + 0.01 * rng.standard_normal(u.shape)
    unm1 = un - 0.004 * (1 + rng.standard_normal(u.shape))
    f0 = rng.standard_normal(u.shape)
    f = [f0 + 0.08 * k * rng.standard_normal(u.shape) for k in range(5)]
    return un, unm1, f


COSTS = np.array([9.0, 4.0, 1.0, 0.0, 2.5, 3.0, 7.0])  # cheapest at order 2.5


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def test_forward_and_gradients_match_oracle(torch):
    print("[gpu-nn] forward/backward vs oracle, 32x32", flush=True)
    H, W = 32, 32
    rng = np.random.default_rng(5)
    net = pda.GuessNet(W, H)
    w = _weights(rng, net.nparams, 0.15)
    net.set_params(w)
    un, unm1, f = _fields(rng, H, W)
    net.set_scales(*_scales(f))
    t = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    loss, g, out = net.gradients([t(x) for x in f], 0.004, COSTS)
    lo, oo, go = G.loss_and_grad(w.astype(np.float64), f, 0.004, COSTS, H, W, *_scales(f), bf=True)
    mo = oo.reshape(-1, 7).mean(axis=0)
    to, po = G.policy(mo)
    assert np.abs(go).max() > 0
    print(f"[gpu-nn] out {rel(out, oo):.2e} loss {abs(loss - lo) / lo:.2e}", flush=True)
    assert rel(out, oo) <= 1e-2
    assert abs(loss - lo) <= 1e-2 * lo
    td, pd, md = net.policy()
    print(f"[gpu-nn] estimates {rel(md, mo):.2e} order {td:.4f} vs {to:.4f}", flush=True)
    assert rel(md, mo) <= 1e-2
    assert abs(td - to) <= 1e-2 and np.abs(pd - po).max() <= 1e-2
    o = 0
    for li, (ci, co) in enumerate(G.LAYERS):
        nW = 9 * ci * co
        eW, eb = rel(g[o:o + nW], go[o:o + nW]), rel(g[o + nW:o + nW + co], go[o + nW:o + nW + co])
        print(f"[gpu-nn] layer {li}: dW {eW:.2e} db {eb:.2e}", flush=True)
        assert eW <= 3e-2 and eb <= 3e-2
        o += nW + co


def test_predict_is_the_policy_predictor(torch):
    H, W = 32, 32
    rng = np.random.default_rng(2)
    net = pda.GuessNet(W, H)
    w = _weights(rng, net.nparams, 0.1)
    net.set_params(w)
    un, unm1, f = _fields(rng, H, W)
    net.set_scales(*_scales(f))
    t = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    ft = [t(x) for x in f]
    gs = net.predict(t(un), t(unm1), ft, 0.004)
    order, p, m = net.policy()
    _, _, out = net.gradients(ft, 0.004, COSTS)
    torch.cuda.synchronize()
    # the estimates (from the last hidden layer's channel sums) are the
    # grid means of the full output layer, up to FP32 rounding; the policy
    # is their soft arg-min, the order its expectation over 1, 1.5, ..., 4
    mo = out.reshape(-1, 7).astype(np.float64).mean(axis=0)
    assert np.allclose(m, mo, rtol=1e-4, atol=1e-5)
    to, po = G.policy(m)
    assert np.allclose(p, po, rtol=1e-12, atol=1e-15) and abs(order - to) <= 1e-12
    assert 1.0 <= order <= 4.0
    expect = G.predictor_t(un, unm1, f, DT, order)
    assert np.allclose(gs.cpu().numpy(), expect, rtol=0, atol=1e-12)


def test_training_reduces_expected_cost(torch):
    print("[gpu-nn] training", flush=True)
    H, W = 32, 32
    rng = np.random.default_rng(9)
    net = pda.GuessNet(W, H)
    w = _weights(rng, net.nparams, 0.05)
    net.set_params(w)
    un, unm1, f = _fields(rng, H, W)
    net.set_scales(*_scales(f))
    t = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    ft = [t(x) for x in f]
    first = net.train_step(ft, 0.004, COSTS, 3e-3, want_loss=True)
    for _ in range(300):
        net.train_step(ft, 0.004, COSTS, 3e-3)
    last = net.train_step(ft, 0.004, COSTS, 3e-3, want_loss=True)
    order, p, _ = net.policy()
    print(f"[gpu-nn] cost-estimate loss {first:.4f} -> {last:.4f}, order {order:.3f}", flush=True)
    assert last < 0.01 * first
    assert abs(order - 2.5) < 0.05  # the cheapest order


def test_inputs_are_validated_before_any_launch(torch):
    """predict / gradients / train_step refuse a field of the wrong size,
    dtype or device (the kernels read and write 2*nx*ny entries of each),
    and costs that are not 7 finite values."""
    net = pda.GuessNet(32, 16)
    n = 2 * 32 * 16
    good = torch.zeros(n, dtype=torch.float64, device="cuda")
    bad = [torch.zeros(2 * n, dtype=torch.float64, device="cuda"),     # another grid's state
           torch.zeros(n, dtype=torch.float32, device="cuda"),         # wrong dtype
           torch.zeros(n, dtype=torch.float64)]                        # host memory
    f = [good] * 5
    for b in bad:
        with pytest.raises(pda.PdaError):
            net.predict(good, good, [good, b, good, good, good], 0.005)
        with pytest.raises(pda.PdaError):
            net.predict(good, good, f, 0.005, out=b)
        with pytest.raises(pda.PdaError):
            net.predict(b, good, f, 0.005)
        with pytest.raises(pda.PdaError):
            net.gradients([b] + f[1:], 0.005, COSTS)
        with pytest.raises(pda.PdaError):
            net.train_step(f[:4] + [b], 0.005, COSTS, 1e-3)
    with pytest.raises(pda.PdaError):  # four velocities instead of five
        net.predict(good, good, [good] * 4, 0.005)
    for c in (COSTS[:6], np.r_[COSTS, 0.0], np.r_[np.nan, COSTS[1:]], np.r_[np.inf, COSTS[1:]]):
        with pytest.raises(pda.PdaError):
            net.train_step(f, 0.005, c, 1e-3)
    out = net.predict(good, good, f, 0.005)
    torch.cuda.synchronize()
    assert out.shape == good.shape and torch.isfinite(out).all()


def test_batch_autotune_validates_state(torch):
    mesh = pda.create_structured_mesh((64, 64), (-5.0, 5.0, -5.0, 5.0), 7)
    probs = [pda.create_problem(mesh, pda.Swe2d.SlipWall, pda.InviscidFluxReconstruction.Weno5, 1)
             for _ in range(2)]
    b = pda.Batch(probs)
    with pytest.raises(pda.PdaError):
        b.autotune(torch.zeros(b.total // 2, dtype=torch.float64, device="cuda"))
    with pytest.raises(pda.PdaError):
        b.autotune(torch.zeros(b.total, dtype=torch.float32, device="cuda"))
