"""The learned-guess network's device kernels (MFMA forward, backward-data,
backward-weight, Adam) against the CPU oracle, which mirrors the device's
BF16 rounding points and computes in FP64; tolerances are BF16-level."""
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
SCALES = (DT, 5e-3, 4e-4, 3e-5, 5.5e-3, 4.5e-3)  # dt, s_f, s_g, s_o, nu_mean, nu_scale


def _fields(rng, H, W):
    y, x = np.meshgrid(np.linspace(-1, 1, H, endpoint=False), np.linspace(-1, 1, W, endpoint=False),
                       indexing="ij")
    u = np.stack([np.exp(-((x - 0.2) ** 2 + y ** 2) / 0.1), 0.5 * np.exp(-(x ** 2 + (y + 0.3) ** 2) / 0.2)],
                 axis=-1).ravel()
    un = u + 0.01 * rng.standard_normal(u.shape)
    unm1 = un - 0.004 * (1 + rng.standard_normal(u.shape))
    fn = rng.standard_normal(u.shape)
    fnm1 = fn + 0.08 * rng.standard_normal(u.shape)
    unp1 = G.base(un, unm1, fn, fnm1, DT) + G.gate(fn, fnm1, DT) * rng.standard_normal(u.shape)
    return un, unm1, fn, fnm1, unp1


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def test_forward_and_gradients_match_oracle(torch):
    print("[gpu-nn] forward/backward vs oracle, 32x32", flush=True)
    H, W = 32, 32
    rng = np.random.default_rng(5)
    net = pda.GuessNet(W, H)
    w = (rng.standard_normal(net.nparams) * 0.15).astype(np.float32)
    net.set_params(w)
    net.set_scales(*SCALES)
    un, unm1, fn, fnm1, unp1 = _fields(rng, H, W)
    t = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    loss, g, out = net.gradients(t(un), t(unm1), t(fn), t(fnm1), t(unp1), 0.004)
    lo, oo, go = G.loss_and_grad(w.astype(np.float64), un, unm1, fn, fnm1, unp1, 0.004, H, W, *SCALES,
                                 bf=True)
    print(f"[gpu-nn] out {rel(out, oo):.2e} loss {abs(loss - lo) / lo:.2e}", flush=True)
    assert rel(out, oo) <= 1e-2
    assert abs(loss - lo) <= 1e-2 * lo
    o = 0
    for li, (ci, co) in enumerate(G.LAYERS):
        nW = 9 * ci * co
        eW, eb = rel(g[o:o + nW], go[o:o + nW]), rel(g[o + nW:o + nW + co], go[o + nW:o + nW + co])
        print(f"[gpu-nn] layer {li}: dW {eW:.2e} db {eb:.2e}", flush=True)
        assert eW <= 3e-2 and eb <= 3e-2
        o += nW + co


def test_predict_is_base_plus_scaled_output(torch):
    H, W = 32, 32
    rng = np.random.default_rng(2)
    net = pda.GuessNet(W, H)
    w = (rng.standard_normal(net.nparams) * 0.1).astype(np.float32)
    net.set_params(w)
    net.set_scales(*SCALES)
    un, unm1, fn, fnm1, unp1 = _fields(rng, H, W)
    t = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    gs = net.predict(t(un), t(unm1), t(fn), t(fnm1), 0.004)
    _, _, out = net.gradients(t(un), t(unm1), t(fn), t(fnm1), t(unp1), 0.004)
    torch.cuda.synchronize()
    expect = G.base(un, unm1, fn, fnm1, DT) + G.gate(fn, fnm1, DT) * out.astype(np.float64)
    assert np.allclose(gs.cpu().numpy(), expect, rtol=0, atol=1e-12)


def test_training_reduces_loss(torch):
    print("[gpu-nn] training", flush=True)
    H, W = 32, 32
    rng = np.random.default_rng(9)
    net = pda.GuessNet(W, H)
    w = (rng.standard_normal(net.nparams) * 0.05).astype(np.float32)
    net.set_params(w)
    net.set_scales(*SCALES)
    un, unm1, fn, fnm1, _ = _fields(rng, H, W)
    # a target the network can represent: a smooth local function of the fields
    unp1 = G.base(un, unm1, fn, fnm1, DT) + G.gate(fn, fnm1, DT) * 0.3 * np.tanh(un)
    t = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    a, b, f0, f1, c = t(un), t(unm1), t(fn), t(fnm1), t(unp1)
    first = net.train_step(a, b, f0, f1, c, 0.004, 3e-3, want_loss=True)
    for _ in range(300):
        net.train_step(a, b, f0, f1, c, 0.004, 3e-3)
    last = net.train_step(a, b, f0, f1, c, 0.004, 3e-3, want_loss=True)
    print(f"[gpu-nn] loss {first:.3e} -> {last:.3e}", flush=True)
    assert last < 0.1 * first


def test_inputs_are_validated_before_any_launch(torch):
    """predict / gradients / train_step refuse a field of the wrong size,
    dtype or device (the kernels read and write 2*nx*ny entries of each)."""
    net = pda.GuessNet(32, 16)
    n = 2 * 32 * 16
    good = torch.zeros(n, dtype=torch.float64, device="cuda")
    bad = [torch.zeros(2 * n, dtype=torch.float64, device="cuda"),     # another grid's state
           torch.zeros(n, dtype=torch.float32, device="cuda"),         # wrong dtype
           torch.zeros(n, dtype=torch.float64)]                        # host memory
    for b in bad:
        with pytest.raises(pda.PdaError):
            net.predict(good, good, b, good, 0.005)
        with pytest.raises(pda.PdaError):
            net.predict(good, good, good, good, 0.005, out=b)
        with pytest.raises(pda.PdaError):
            net.gradients(good, good, good, good, b, 0.005)
        with pytest.raises(pda.PdaError):
            net.train_step(b, good, good, good, good, 0.005, 1e-3)
    out = net.predict(good, good, good, good, 0.005)
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
