"""Learned proposals: the numpy restatement of the network and its training
kernels (oracle/lp_ref.py) is pinned by finite differences; the GPU kernels
(include/dme_learned.h) are checked against it; the trained network is
checked against the method without it."""
import ctypes as C
import os

import numpy as np
import pytest

import lp_ref as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _small_problem(seed, n=1, h=4, w=5):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, h, w, 32))
    x[..., :8] = rng.standard_normal((n, h, w, 8))
    params = np.concatenate([np.concatenate([rng.standard_normal(9 * 32 * 32) * 0.15, rng.standard_normal(32) * 0.1])
                             for _ in range(4)])
    lab = np.zeros((n, h, w, 8))
    for r in range(2):
        lab[..., 4 * r:4 * r + 2] = rng.uniform(-6, 6, (n, h, w, 2))
        lab[..., 4 * r + 2] = rng.uniform(2, 8, (n, h, w))
        lab[..., 4 * r + 3] = rng.random((n, h, w)) < 0.8
    return x, params, lab


def _loss64(x, params, lab, alpha=0.1, beta=0.5):
    out, _ = R.forward(x, params, rounding=False)
    L, _ = R.wta_loss(out, lab, alpha, beta, 1.0, dtype=np.float64)
    return L


def test_bf16_rounding_ties_to_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    r = R.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -6) and r[3] == -2.5
    assert abs(r[4] - 3.0e-3) < 3.0e-3 * 2 ** -8


def test_oracle_gradient_matches_finite_differences():
    x, params, lab = _small_problem(1)
    out, acts = R.forward(x, params, rounding=False)
    _, dout = R.wta_loss(out, lab, 0.1, 0.5, 1.0, dtype=np.float64)
    g = R.backward(acts, dout, params, rounding=False)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(R.LAYER * l + np.arange(R.LAYER), 8, replace=False) for l in range(4)])
    eps = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (_loss64(x, p1, lab) - _loss64(x, p2, lab)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_wta_loss_picks_the_best_hypothesis():
    out = np.zeros((1, 1, 1, 32), np.float32)
    out[0, 0, 0, 0:8] = [0.0, 0.0, 1.0, 0.5, -1.0, 0.0, 0.25, 0.25]  # ref 0: 4 hypotheses (x, y) / 8 pel
    lab = np.zeros((1, 1, 1, 8), np.float32)
    lab[0, 0, 0, 0:4] = [8.0, 4.0, 0.0, 1.0]  # target (8, 4) pel, valid
    L, d = R.wta_loss(out, lab, 0.0, 0.0, 1.0)
    assert L == 0.0  # hypothesis 1 is exact
    lab[0, 0, 0, 0:2] = [9.0, 4.0]
    L, d = R.wta_loss(out, lab, 0.0, 0.0, 1.0)
    assert L == 1.0 and d[0, 0, 0, 2] == -8.0 and np.count_nonzero(d) == 1


def test_adam_oracle_first_step_is_sign_times_lr():
    p, m, v = R.adam(np.zeros(3), np.array([2.0, -0.5, 0.0]), np.zeros(3), np.zeros(3), 0.01, 0.9, 0.999, 1e-12, 1)
    assert np.allclose(p, [-0.01, 0.01, 0.0])


# ---- GPU ----------------------------------------------------------------


@pytest.fixture(scope="module")
def gctx():
    import dirac_me

    c = dirac_me.Context(0)
    yield c
    c.close()


def _bits(a):
    """float32 values that are bfloat16 -> their 16-bit patterns."""
    return (np.ascontiguousarray(R.bf16(a), np.float32).view(np.uint32) >> 16).astype(np.uint16)


def _dev(ctx, a, dtype):
    from dirac_me.learned import DeviceArray

    a = np.ascontiguousarray(a, dtype)
    d = DeviceArray(ctx, a.shape, dtype)
    d.upload(a)
    return d


@pytest.mark.gpu
def test_gpu_conv_forward_and_data_gradient_match_oracle(gctx):
    from dirac_me import learned as L
    from dirac_me.abi import check, lib

    rng = np.random.default_rng(3)
    n, h, w = 2, 7, 9
    x = R.bf16(rng.standard_normal((n, h, w, 32)).astype(np.float32))
    wt = rng.standard_normal((9, 32, 32)).astype(np.float32) * 0.2
    b = rng.standard_normal(32).astype(np.float32)
    dw = _dev(gctx, wt.ravel(), np.float32)
    ff = L.DeviceArray(gctx, L.FRAG, np.uint16)
    fb = L.DeviceArray(gctx, L.FRAG, np.uint16)
    dx = _dev(gctx, _bits(x).ravel(), np.uint16)
    db = _dev(gctx, b, np.float32)
    out = L.DeviceArray(gctx, (n, h, w, 32), np.float32)
    outb = L.DeviceArray(gctx, (n, h, w, 32), np.uint16)
    check(lib().dme_lp_prepare(gctx.h, dw.at(), ff.at(), fb.at()))
    check(lib().dme_lp_conv(gctx.h, dx.at(), n, h, w, ff.at(), db.at(), 1, outb.at(), out.at()))
    ref = R.conv(x, R.bf16(wt), b, relu=True)
    got = out.download()
    assert np.max(np.abs(got - ref)) <= 1e-4 * (1 + np.max(np.abs(ref)))
    gotb = (outb.download().astype(np.uint32) << 16).view(np.float32)
    assert np.array_equal(gotb, R.bf16(got))
    # data gradient: the same kernel with the backward fragments, no bias
    check(lib().dme_lp_conv(gctx.h, dx.at(), n, h, w, fb.at(), None, 0, None, out.at()))
    ref = R.conv_data_grad(x, R.bf16(wt))
    got = out.download()
    assert np.max(np.abs(got - ref)) <= 1e-4 * (1 + np.max(np.abs(ref)))


@pytest.mark.gpu
def test_gpu_weight_and_bias_gradient_match_oracle(gctx):
    # dme_lp_wgrad alone on BF16 inputs: every product is exact in f32, so
    # each element of dw and db agrees with the float64 sum to the f32
    # accumulation error of its own terms (a mis-indexed tap or channel is
    # off by whole terms)
    from dirac_me import learned as L
    from dirac_me.abi import check, lib

    rng = np.random.default_rng(11)
    n, h, w = 2, 6, 7
    x = R.bf16(rng.standard_normal((n, h, w, 32)).astype(np.float32))
    g = R.bf16(rng.standard_normal((n, h, w, 32)).astype(np.float32))
    dx, dg = _dev(gctx, _bits(x).ravel(), np.uint16), _dev(gctx, _bits(g).ravel(), np.uint16)
    dw, db = L.DeviceArray(gctx, (9, 32, 32), np.float32), L.DeviceArray(gctx, (32,), np.float32)
    dw.zero()
    db.zero()
    check(lib().dme_lp_wgrad(gctx.h, dx.at(), dg.at(), n, h, w, dw.at(), db.at()))
    want_w, want_b = R.conv_weight_grad(x, g)
    bound_w, bound_b = R.conv_weight_grad(np.abs(x), np.abs(g))
    assert np.all(np.abs(dw.download() - want_w) <= 1e-5 * bound_w + 1e-7)
    assert np.all(np.abs(db.download() - want_b) <= 1e-5 * bound_b + 1e-7)


@pytest.mark.gpu
def test_gpu_relu_backward_and_wta_loss_match_oracle(gctx):
    from dirac_me import learned as L
    from dirac_me.abi import check, lib

    rng = np.random.default_rng(12)
    # ReLU backward: bf16(g) where the activation is positive, else 0, exactly
    count = 4 * 5 * 6 * 32
    act = R.bf16(rng.standard_normal(count).astype(np.float32))
    g = rng.standard_normal(count).astype(np.float32)
    da, dg = _dev(gctx, _bits(act), np.uint16), _dev(gctx, g, np.float32)
    out = L.DeviceArray(gctx, (count,), np.uint16)
    check(lib().dme_lp_relu_backward(gctx.h, da.at(), dg.at(), count, out.at()))
    assert np.array_equal(out.download(), _bits(np.where(act > 0, g, 0.0).astype(np.float32)))
    # winner-takes-all loss and its gradient on random outputs (no ties)
    _, _, lab = _small_problem(13, n=2, h=5, w=6)
    npos = lab.size // 8
    o = rng.standard_normal((npos, 32)).astype(np.float32)
    do, dl = _dev(gctx, o.ravel(), np.float32), _dev(gctx, lab.astype(np.float32).ravel(), np.float32)
    dd, loss = L.DeviceArray(gctx, (npos, 32), np.float32), L.DeviceArray(gctx, (1,), np.float32)
    loss.zero()
    scale = 1.0 / npos
    check(lib().dme_lp_wta_loss(gctx.h, do.at(), dl.at(), npos, C.c_float(0.1), C.c_float(0.5), C.c_float(scale),
                                dd.at(), loss.at()))
    want_l, want_d = R.wta_loss(o, lab.reshape(npos, 8), 0.1, 0.5, scale, dtype=np.float64)
    assert np.allclose(dd.download(), want_d, rtol=1e-6, atol=1e-9)
    assert abs(float(loss.download()[0]) - want_l) <= 1e-5 * abs(want_l)


@pytest.mark.gpu
def test_gpu_training_step_matches_oracle(gctx):
    from dirac_me import learned as L

    x, params, lab = _small_problem(4, n=2, h=6, w=7)
    params = params.astype(np.float32)
    net = L.Network(gctx, params)
    n, h, w = x.shape[:3]
    fx = _dev(gctx, _bits(x.astype(np.float32)).ravel(), np.uint16)
    fl = _dev(gctx, lab.astype(np.float32).ravel(), np.float32)
    net.loss_and_grad(fx.at(), fl.at(), n, h, w, alpha=0.1, beta=0.5)
    g_gpu = net.g.download().astype(np.float64)
    loss_gpu = net.loss_value()
    out, acts = R.forward(x.astype(np.float32), params)
    loss, dout = R.wta_loss(out, lab, 0.1, 0.5, 1.0 / (n * h * w))
    g_ref = R.backward(acts, dout, params)
    assert abs(loss_gpu - loss) <= 1e-3 * abs(loss)
    # BF16 activations may differ by one rounding where the f32 sums differ
    # in order; the gradient agrees to well within that
    assert np.linalg.norm(g_gpu - g_ref) <= 2e-2 * np.linalg.norm(g_ref)
    # Adam step on the GPU equals the oracle's on the GPU gradient
    p_ref, _, _ = R.adam(params.astype(np.float64), g_gpu, 0, 0, 1e-3, 0.9, 0.999, 1e-8, 1)
    net.step(1e-3)
    assert np.allclose(net.params(), p_ref, rtol=1e-5, atol=1e-7)


@pytest.mark.gpu
def test_trained_network_beats_the_method_without_it(gctx):
    """On held-out 1080p pictures the committed network keeps the mean
    Lagrangian (SB) cost within 1% of the plain search and needs the wide
    fallback search for far fewer blocks than an untrained network (whose
    proposals sit on the level-2 guide), which is what it saves."""
    from dirac_me import learned as L

    trained = L.evaluate(gctx, L.load_weights(), count=2)
    untrained = L.evaluate(gctx, L.init_params(0), count=2)
    assert trained["gap"] <= 0.01, trained
    assert trained["fallback_rate"] < 0.6 * untrained["fallback_rate"], (trained, untrained)


@pytest.mark.gpu
def test_learned_batch_matches_single_picture_calls(gctx):
    """A batch of pictures through dme_run_batch_learned gives the same
    MEData as one dme_motion_estimate_learned call per picture."""
    import dirac_me
    from dirac_me import learned as L, synth

    gctx.set_lp_weights(L.load_weights())
    sets = [synth.config1(L._picture_seed(i, True)) for i in range(2)]
    prm = dirac_me.make_params(*sets[0]["block"], precision=2, depth=4, lam=10.0)
    gctx.set_params(prm)
    g = dirac_me.geometry(prm, 1920, 1080)
    for i, d in enumerate(sets):
        gctx.upload(10 * i, d["refs"][0])
        gctx.upload(10 * i + 2, d["refs"][1])
        gctx.upload(10 * i + 1, d["cur"], d["u"], d["v"])
    batch = [dirac_me.MEData(g) for _ in sets]
    gctx.run_batch_learned([(10 * i + 1, [10 * i, 10 * i + 2], prm) for i in range(len(sets))], batch)
    for i in range(len(sets)):
        one = dirac_me.MEData(g)
        gctx.motion_estimate_learned(10 * i + 1, [10 * i, 10 * i + 2], one)
        for k, a in one.arrays().items():
            assert np.array_equal(a, batch[i].arrays()[k]), k
