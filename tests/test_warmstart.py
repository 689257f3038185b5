"""Learned warm start: the network layers against the CPU layer oracle
(oracle/nn_oracle.py), the whole U-Net gradient against a float64 CPU
composition of the same layers, the Chapman-Enskog lift against its closed
form moments, and a short training run.

The layer oracle is pinned here too: its convolution against
scipy.signal.correlate, and its weight and data gradients against central
finite differences of a scalar loss."""
import math

import numpy as np
import pytest
from scipy import signal

from mplb_amd import warmstart as ws
from oracle import nn_oracle as nno


# ---- pins of the layer oracle -------------------------------------------------
def test_oracle_conv_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 7, 6, 3))
    w = rng.standard_normal((9, 3, 4))
    out = nno.conv3x3(x, w, round_bf16=False)
    for n in range(2):
        for co in range(4):
            ref = sum(signal.correlate(x[n, :, :, ci], w[:, ci, co].reshape(3, 3), mode="same")
                      for ci in range(3))
            assert np.allclose(out[n, :, :, co], ref, atol=1e-12)


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 5, 4, 2))
    w = rng.standard_normal((9, 2, 3))
    r = rng.standard_normal((1, 5, 4, 3))  # loss = sum(r * conv(x, w))
    loss = lambda xx, ww: float((r * nno.conv3x3(xx, ww, round_bf16=False)).sum())
    dw, db = nno.conv3x3_wgrad(x, r, round_bf16=False)
    dx = nno.conv3x3_dgrad(r, w, round_bf16=False)
    e = 1e-6
    for idx in [(0, 0, 0), (4, 1, 2), (8, 0, 1), (2, 1, 0)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += e
        wm[idx] -= e
        assert abs((loss(x, wp) - loss(x, wm)) / (2 * e) - dw[idx]) < 1e-6
    for idx in [(0, 0, 0, 0), (0, 4, 3, 1), (0, 2, 1, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += e
        xm[idx] -= e
        assert abs((loss(xp, w) - loss(xm, w)) / (2 * e) - dx[idx]) < 1e-6
    assert np.allclose(db, r.sum(axis=(0, 1, 2)))


def test_oracle_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], np.float32)
    r = nno.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7        # rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) / 3.0e-3 < 2 ** -8


def test_distance_inputs_match_a_distance_transform():
    """The closed-form distance channels equal scipy's exact Euclidean
    distance transform of the obstacle mask framed by the walls."""
    from scipy import ndimage
    from mplb_amd import cases
    n = 64
    res, rects = ws.case_params(24, seed=21, n=n)
    x = ws.inputs(n, res, rects)
    for k in range(len(res)):
        solid = cases.rect_mask(n, rects[k]).astype(bool)
        framed = np.pad(solid, 1, constant_values=True)
        d = ndimage.distance_transform_edt(~framed)[1:-1, 1:-1]
        assert np.max(np.abs(x[k, :, :, 6] - d / (n / 2))) < 1e-6
        do = (np.minimum(ndimage.distance_transform_edt(~solid), n) / n if solid.any()
              else np.ones((n, n)))
        assert np.max(np.abs(x[k, :, :, 7] - do)) < 1e-6


def test_network_size():
    # 24-48-64-64 encoder, 64-channel bottleneck: 9 ci co + co per conv
    assert ws.param_count() == 563_688
    # every convolution input has a multiple of 8 channels (the MFMA K step)
    for name, ci, co, relu, lev in ws.LAYERS:
        assert ci % 8 == 0, name


def test_library_exports_header():
    import ctypes
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(ws.__file__), "..", "include", "mplbws.h")).read()
    names = sorted(set(re.findall(r"\b(mplbws_[a-z0-9_]+)\s*\(", hdr)))
    lib = ctypes.CDLL(ws.WS_PATH)
    for n in names:
        assert hasattr(lib, n), n


# ---- the HIP kernels ----------------------------------------------------------
def _t(a):
    return ws.DeviceArray.from_numpy(a)


def _np(t):
    return t.numpy().astype(np.float64)


def _empty(shape, zero=False):
    return ws.DeviceArray(shape, zero=zero)


@pytest.mark.gpu
@pytest.mark.parametrize("n,h,w,ci,co,relu", [(2, 5, 6, 8, 16, 1), (1, 8, 8, 16, 8, 0),
                                                (3, 4, 9, 64, 64, 1), (1, 6, 6, 128, 32, 0),
                                                (2, 16, 16, 32, 16, 1),
                                                # several row chunks per image, a ragged row
                                                (2, 40, 130, 16, 16, 1),
                                                # the U-Net's own shapes: 128-wide rows
                                                # with a 72-channel concatenation, and a
                                                # 16^2 level with 4 input-channel chunks
                                                (2, 32, 128, 72, 24, 1),
                                                (3, 16, 16, 128, 64, 0)])
def test_gpu_conv_fwd_dgrad_wgrad(n, h, w, ci, co, relu):
    print("conv %s" % ((n, h, w, ci, co, relu),), flush=True)
    lib = ws.load()
    rng = np.random.default_rng(7)
    x = rng.standard_normal((n, h, w, ci)).astype(np.float32)
    wt = (rng.standard_normal((9, ci, co)) / math.sqrt(9 * ci)).astype(np.float32)
    b = rng.standard_normal(co).astype(np.float32)
    dout = rng.standard_normal((n, h, w, co)).astype(np.float32)
    X, W, Bb, D = _t(x), _t(wt), _t(b), _t(dout)
    out = _empty((n, h, w, co))
    assert lib.mplbws_conv3x3_fwd(ws._p(X), ws._p(W), ws._p(Bb), ws._p(out), n, h, w, ci, co,
                                  relu) == 0
    ref = nno.conv3x3(x, wt, b, relu=bool(relu))
    scale = np.abs(ref).max()
    assert np.max(np.abs(_np(out) - ref)) <= 1e-5 * scale
    # data gradient through the flipped weights
    wf = _empty((9 * ci * co,))
    din = _empty((n, h, w, ci))
    assert lib.mplbws_flip_weights(ws._p(W), ws._p(wf), ci, co) == 0
    assert lib.mplbws_conv3x3_fwd(ws._p(D), ws._p(wf), None, ws._p(din), n, h, w, co, ci, 0) == 0
    refd = nno.conv3x3_dgrad(dout, wt)
    assert np.max(np.abs(_np(din) - refd)) <= 1e-5 * np.abs(refd).max()
    # weight and bias gradients
    wsz = lib.mplbws_wgrad_workspace(n, h, w, ci, co)
    work = _empty((int(wsz),))
    dw = _empty((9, ci, co))
    db = _empty((co,))
    assert lib.mplbws_conv3x3_wgrad(ws._p(X), ws._p(D), ws._p(dw), ws._p(db), ws._p(work), n, h,
                                    w, ci, co) == 0
    rw, rb = nno.conv3x3_wgrad(x, dout)
    assert np.max(np.abs(_np(dw) - rw)) <= 1e-5 * np.abs(rw).max()
    assert np.max(np.abs(_np(db) - rb)) <= 1e-5 * np.abs(rb).max()


@pytest.mark.gpu
def test_gpu_pool_up_channels():
    lib = ws.load()
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 8, 6, 4)).astype(np.float32)
    X = _t(x)
    p = _empty((2, 4, 3, 4))
    assert lib.mplbws_pool2_fwd(ws._p(X), ws._p(p), 2, 8, 6, 4) == 0
    assert np.max(np.abs(_np(p) - nno.pool2(x))) < 1e-6
    # adjoint identities: <pool x, y> = <x, pool^T y>, <up x, y> = <x, up^T y>
    y = rng.standard_normal((2, 4, 3, 4)).astype(np.float32)
    gx = _empty(X.shape)
    assert lib.mplbws_pool2_bwd(ws._p(_t(y)), ws._p(gx), 2, 8, 6, 4) == 0
    assert abs(float((nno.pool2(x) * y).sum()) - float((x * _np(gx)).sum())) < 1e-4
    c = _empty((2, 8, 6, 7), zero=True)
    assert lib.mplbws_up2_fwd(ws._p(_t(y)), ws._p(c), 2, 4, 3, 4, 7, 2) == 0
    cu = _np(c)
    assert np.array_equal(cu[..., 2:6], nno.up2(y))
    assert np.all(cu[..., :2] == 0) and np.all(cu[..., 6:] == 0)
    z = rng.standard_normal((2, 8, 6, 7)).astype(np.float32)
    gy = _empty((2, 4, 3, 4))
    assert lib.mplbws_up2_bwd(ws._p(_t(z)), ws._p(gy), 2, 4, 3, 4, 7, 2) == 0
    assert abs(float((nno.up2(y) * z[..., 2:6]).sum()) - float((y * _np(gy)).sum())) < 1e-4


@pytest.mark.gpu
def test_gpu_gather_rows():
    lib = ws.load()
    rng = np.random.default_rng(5)
    src = rng.standard_normal((9, 3, 5, 4)).astype(np.float32)
    idx = np.array([7, 0, 7, 3, 8], np.int32)
    S, I = _t(src), ws.DeviceArray((5,))
    I.upload(idx)
    dst = _empty((5, 3, 5, 4))
    assert lib.mplbws_gather(ws._p(S), ws._p(dst), ws._p(I), 5, 60) == 0
    assert np.array_equal(_np(dst), src[idx])


def _cpu_unet(params, x, target, mask, enc, bott, rb=True):
    """float64 forward/backward of the U-Net from the oracle layers (BF16
    rounding of conv operands as on the GPU when rb); returns (loss, grads)."""
    P = params
    spec = {nm: (ci, co, r) for nm, ci, co, r, _ in ws.layer_specs(enc, bott)}
    act, inp = {}, {}

    def conv(name, src):
        inp[name] = src
        act[name] = nno.conv3x3(src, P["w_" + name], P["b_" + name], relu=bool(spec[name][2]),
                                round_bf16=rb)
        return act[name]

    nl = len(enc)
    skips = []
    cur = x
    for k in range(nl):
        e = conv("e%db" % k, conv("e%da" % k, cur))
        skips.append(e)
        cur = nno.pool2(e)
    cur = conv("bb", conv("ba", cur))
    for k in reversed(range(nl)):
        cat = np.concatenate([nno.up2(cur), skips[k]], axis=3)
        cur = conv("d%db" % k, conv("d%da" % k, cat))
    out = conv("out", cur)
    fluid = (mask == 0)[..., None]
    cnt = fluid.sum() * 3
    diff = (out[..., :3] - target) * fluid
    loss = float((diff ** 2).sum() / cnt)
    gout = np.zeros_like(out)
    gout[..., :3] = 2 * diff / cnt
    grads = {}

    def back(name, g):
        if spec[name][2]:
            g = np.where(act[name] > 0, g, 0.0)
        grads["w_" + name], grads["b_" + name] = nno.conv3x3_wgrad(inp[name], g, round_bf16=rb)
        return nno.conv3x3_dgrad(g, P["w_" + name], round_bf16=rb)

    def pool_t(g):
        return nno.up2(g) / 4.0

    def up_t(g):
        n, h, w, c = g.shape
        return g.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))

    g = back("out", gout)
    gskip = [None] * nl
    for k in range(nl):
        c = enc[k]
        gcat = back("d%da" % k, back("d%db" % k, g))
        cc = gcat.shape[3] - c
        gskip[k] = gcat[..., cc:]
        g = up_t(gcat[..., :cc])
    g = back("ba", back("bb", g))
    for k in reversed(range(nl)):
        g = pool_t(g) + gskip[k]
        g = back("e%da" % k, back("e%db" % k, g))
    return loss, grads


def test_cpu_unet_gradient_matches_finite_differences():
    """The CPU composition used to check the GPU U-Net, in plain float64,
    against central differences of its own loss."""
    enc, bott, n = (8, 8), 8, 8
    rng = np.random.default_rng(4)
    P = {}
    for name, ci, co, relu, _ in ws.layer_specs(enc, bott):
        P["w_" + name] = rng.standard_normal((9, ci, co)) / math.sqrt(9 * ci)
        P["b_" + name] = 0.1 * rng.standard_normal(co)
    x = rng.standard_normal((1, n, n, ws.IN_CH))
    target = rng.standard_normal((1, n, n, 3))
    mask = np.zeros((1, n, n))
    mask[0, 2:4, 3:5] = 1
    loss, grads = _cpu_unet(P, x, target, mask, enc, bott, rb=False)
    e = 1e-6
    for key, idx in [("w_e0a", (4, 2, 3)), ("w_ba", (0, 1, 5)), ("w_d1a", (8, 9, 2)),
                     ("w_out", (4, 3, 1)), ("b_e1b", (2,)), ("w_d0b", (1, 7, 0))]:
        Pp = {k: v.copy() for k, v in P.items()}
        Pm = {k: v.copy() for k, v in P.items()}
        Pp[key][idx] += e
        Pm[key][idx] -= e
        fd = (_cpu_unet(Pp, x, target, mask, enc, bott, rb=False)[0]
              - _cpu_unet(Pm, x, target, mask, enc, bott, rb=False)[0]) / (2 * e)
        assert abs(fd - grads[key][idx]) <= 1e-6 * max(1.0, abs(fd)), (key, fd, grads[key][idx])


@pytest.mark.gpu
def test_gpu_unet_gradients_match_cpu():
    print("U-Net 16x16 batch 2 (3 levels) gradients vs CPU", flush=True)
    n, B = 16, 2
    enc, bott = (8, 16, 16), 16
    net = ws.UNet(n=n, batch=B, seed=5, enc=enc, bott=bott)
    res, rects = [150.0, 700.0], [[(3, 4, 5, 6)], []]
    x = ws.inputs(n, res, rects)
    rng = np.random.default_rng(2)
    target = rng.standard_normal((B, n, n, 3)).astype(np.float32)
    mask = x[..., 0].copy()
    net.forward(x)
    loss_gpu, gout = net.loss_grad(target, mask)
    net.backward(gout)
    params = net.state_dict()
    loss_cpu, grads = _cpu_unet(params, x, target.astype(np.float64), mask, enc, bott)
    assert abs(loss_gpu - loss_cpu) <= 1e-4 * abs(loss_cpu)
    for name, _, _, _, _ in ws.layer_specs(enc, bott):
        gw = net.gw[name].numpy()
        gb = net.gb[name].numpy()
        rw, rb = grads["w_" + name], grads["b_" + name]
        # BF16 rounding of the chain's intermediate operands compounds over
        # the layers: a relative tolerance of a few 1e-3 of the largest entry
        assert np.max(np.abs(gw - rw)) <= 5e-3 * np.abs(rw).max() + 1e-12, name
        assert np.max(np.abs(gb - rb)) <= 5e-3 * np.abs(rb).max() + 1e-12, name


@pytest.mark.gpu
def test_gpu_multiscale_loss_gradient():
    """The output gradient with large-scale terms, against the same loss
    written with numpy reshapes: fine masked MSE plus, per level k, w times
    the mean of (P_k(m p) - P_k(t))^2 over the 2^k-pooled pixels."""
    n, B = 32, 2
    coarse = ((1, 0.5), (3, 2.0))
    net = ws.UNet(n=n, batch=B, seed=3, enc=(16, 32, 32), bott=32, coarse=coarse)
    res, rects = ws.case_params(B, seed=4, n=n)
    x = ws.inputs(n, res, rects)
    m = x[..., 0] == 0
    rng = np.random.default_rng(11)
    t = (rng.standard_normal((B, n, n, 3)) * 0.1 * m[..., None]).astype(np.float32)
    out = net.forward(x).numpy().astype(np.float64)
    loss, g = net.loss_grad(t, x[..., 0])
    g = g.numpy()
    p = out[..., :3]
    ref = np.zeros((B, n, n, ws.OUT_CH))
    ref[..., :3] = np.where(m[..., None], 2.0 * (p - t) / (m.sum() * 3.0), 0.0)
    assert abs(loss - ((p - t) ** 2 * m[..., None]).sum() / (m.sum() * 3.0)) <= 1e-5 * loss
    pm = p * m[..., None]
    for k, w in coarse:
        f = 1 << k
        sz = n // f
        pool = lambda a: a.reshape(B, sz, f, sz, f, 3).mean(axis=(2, 4))
        gc = 2.0 * w * (pool(pm) - pool(t.astype(np.float64))) / (B * sz * sz * 3.0)
        up = np.repeat(np.repeat(gc, f, axis=1), f, axis=2) / (f * f)
        ref[..., :3] += up * m[..., None]
    assert np.max(np.abs(g - ref)) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.gpu
def test_gpu_training_reduces_loss():
    print("U-Net 32x32 training steps", flush=True)
    n, B = 32, 8
    net = ws.UNet(n=n, batch=B, seed=1, enc=(16, 32, 32), bott=32)
    res, rects = ws.case_params(B, seed=9, n=n)
    x = ws.inputs(n, res, rects)
    yy, xx = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n), indexing="ij")
    target = np.stack([np.sin(3 * yy) * xx, yy * (1 - xx), 0.1 * xx], -1)
    target = np.broadcast_to(target, (B, n, n, 3)).astype(np.float32)
    losses = []
    for it in range(60):
        net.forward(x)
        loss, g = net.loss_grad(target, x[..., 0])
        net.backward(g)
        net.adam(lr=3e-3)
        losses.append(loss)
    print("loss %.4g -> %.4g" % (losses[0], losses[-1]), flush=True)
    assert losses[-1] < 0.2 * losses[0]


# ---- the lift ------------------------------------------------------------------
@pytest.mark.gpu
def test_gpu_lift_moments_of_a_shear_flow():
    """Linear shear u = (a y, b x) in a periodic-free 2D block: the lifted
    f has density rho, momentum rho u and non-equilibrium momentum flux
    sum_q c_x c_y f1 = -(tau_L rho / 3)(d_x u_y + d_y u_x) (lattice units),
    the closed form of the first-order BGK stress."""
    from mplb_amd import cases, hip
    from mplb_amd.lattice import lattice
    n = 12
    g = hip.LatticeCase(2, storage=hip.STORE_F64, device=0)
    p = cases.cavity2d(g, n, re=100.0)
    a, b = 2e-3, -1e-3
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    rho = 1.0 + 1e-3 * np.cos(xx)
    u = cases.CS * a * yy  # code units
    v = cases.CS * b * xx
    g.set_macros(0, rho, u, v)
    g.init_lift()
    f = g.get_f(0)
    c, w = lattice(2)
    assert np.allclose(f.sum(0), rho, atol=1e-14)
    jx = sum(cases.CS * c[q, 0] * f[q] for q in range(9))
    assert np.allclose(jx, rho * u, atol=1e-14)
    tau_l = p["tau_l"]
    pxy = sum(c[q, 0] * c[q, 1] * f[q] for q in range(9))
    feq_xy = rho * (u / cases.CS) * (v / cases.CS)
    expect = feq_xy - tau_l * rho / 3.0 * (a + b)
    inner = (slice(1, n - 1), slice(1, n - 1))
    assert np.allclose(pxy[inner], expect[inner], rtol=0, atol=1e-13)


@pytest.mark.gpu
def test_gpu_warmstart_pipeline_small():
    """The whole pipeline at a small size: data from the batched solver,
    training, and learned versus rest starts that both converge."""
    print("warm-start pipeline 32^2, 96 train / 16 test", flush=True)
    res, _ = ws.run(n=32, n_train=96, n_test=16, epochs=4, criteria=1e-8, check_period=500,
                    chunk=64, log=lambda m: print(m, flush=True))
    print(res, flush=True)
    assert res["loss_last"] < res["loss_first"]
    assert res["learned_steps_max"] > 0 and res["rest_steps_max"] > 0
    # both starts converge to the same steady flow, to what the criterion
    # leaves: FP32 runs stopped at a relative change of 1e-8 per unit time
    # still carry the slowest transient, at the 1e-4 .. 1e-3 level here
    assert res["velocity_rel_l2_learned_vs_rest_median"] < 1e-3
    assert res["velocity_rel_l2_learned_vs_rest_max"] < 5e-3
