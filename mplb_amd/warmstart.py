"""Learned warm start for batched 2D cavities with obstacles.

A U-Net (563,688 parameters in the default widths) maps (obstacle mask, log Re, lid speed, lid
row, node coordinates) on an n x n cavity to the steady flow (u, v and the
density deviation); the lattice then starts from that flow lifted to
populations (mplbhip_init_lift: feq plus the first-order Chapman-Enskog
part) instead of from rest.  There is no reference counterpart: the
reference starts every run from the App's initial macros.

Compute runs in libmplbws.so (include/mplbws.h): 3x3 convolutions on the
matrix cores (BF16 operands, FP32 accumulation), their gradients, pooling,
upsampling, loss and Adam; device memory comes from the same library
(DeviceArray), so the process holds one HIP runtime.
"""
import ctypes
import math
import os
import time

import numpy as np

from . import cases, hip

_HERE = os.path.dirname(os.path.abspath(__file__))
WS_PATH = os.path.join(_HERE, "libmplbws.so")
_ws = None

LID = 0.05                     # lattice lid speed of the cavities
# scale of the density-deviation target: a tenth of 1/(3 U^2), so that the
# density channel (dominated by the lid-corner singularities, and only a
# small part of the lift) weighs less than the velocity in the loss
RHO_SCALE = 1.0 / (30 * LID * LID)
IN_CH = 8                      # input channels
LOSS_PARTIALS = 1024           # partial sums of one loss evaluation
OUT_CH = 8                     # output channels (3 used)


class WarmStartError(RuntimeError):
    pass


def load():
    global _ws
    if _ws is not None:
        return _ws
    if not os.path.exists(WS_PATH):
        raise WarmStartError("libmplbws.so is not built (python -m mplb_amd.build)")
    lib = ctypes.CDLL(WS_PATH)
    P, I, F, L = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int64
    sig = {
        "mplbws_conv3x3_fwd": [P, P, P, P, I, I, I, I, I, I],
        "mplbws_flip_weights": [P, P, I, I],
        "mplbws_wgrad_workspace": [I, I, I, I, I],
        "mplbws_conv3x3_wgrad": [P, P, P, P, P, I, I, I, I, I],
        "mplbws_relu_bwd": [P, P, P, L],
        "mplbws_pool2_fwd": [P, P, I, I, I, I],
        "mplbws_pool2_bwd": [P, P, I, I, I, I],
        "mplbws_up2_fwd": [P, P, I, I, I, I, I, I],
        "mplbws_up2_bwd": [P, P, I, I, I, I, I, I],
        "mplbws_channels": [P, P, L, I, I, I, I],
        "mplbws_gather": [P, P, P, I, L],
        "mplbws_mask_channels": [P, P, P, L, I, I],
        "mplbws_mse": [P, P, P, P, P, I, L, I, I, F],
        "mplbws_adam": [P, P, P, P, L, F, F, F, F, I],
        "mplbws_sync": [],
        "mplbws_alloc": [ctypes.POINTER(ctypes.c_void_p), L],
        "mplbws_free": [P],
        "mplbws_upload": [P, P, L],
        "mplbws_download": [P, P, L],
        "mplbws_zero": [P, L],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = L if name == "mplbws_wgrad_workspace" else I
    _ws = lib
    return lib


class DeviceArray:
    """An FP32 device array owned by libmplbws.so (hipMalloc)."""

    def __init__(self, shape, zero=False):
        self.lib = load()
        self.shape = tuple(int(s) for s in shape)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        p = ctypes.c_void_p()
        _ck(self.lib.mplbws_alloc(ctypes.byref(p), self.size * 4), "alloc")
        self.p = p
        if zero:
            self.zero()

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        d = cls(a.shape)
        d.upload(a)
        return d

    def upload(self, a):
        """Copy a host array in: FP32 values, or int32 ones stored as they
        are (index arrays for mplbws_gather)."""
        a = np.ascontiguousarray(a)
        if a.dtype != np.int32:
            a = a.astype(np.float32, copy=False)
        if a.size != self.size:
            raise WarmStartError("upload of %d values into %d" % (a.size, self.size))
        _ck(self.lib.mplbws_upload(self.p, a.ctypes.data, a.size * 4), "upload")

    def numpy(self):
        out = np.empty(self.shape, np.float32)
        _ck(self.lib.mplbws_download(out.ctypes.data, self.p, self.size * 4), "download")
        return out

    def zero(self):
        _ck(self.lib.mplbws_zero(self.p, self.size * 4), "zero")

    def numel(self):
        return self.size

    def __del__(self):
        try:
            if self.p:
                self.lib.mplbws_free(self.p)
                self.p = None
        except Exception:
            pass


def _ck(rc, what):
    if rc != 0:
        raise WarmStartError("%s failed (%d)" % (what, rc))


def _p(t):
    return t.p


def _at(t, offset):
    """Device pointer offset by `offset` 4-byte elements into t."""
    return ctypes.c_void_p(t.p.value + 4 * int(offset))


# ---- the network -----------------------------------------------------------
# Encoder channels per level (each level: two 3x3 conv + ReLU, then 2x2
# average pooling), a bottleneck of two conv + ReLU, and a mirrored decoder
# (nearest 2x upsampling, concatenation with the encoder level, two conv +
# ReLU), then a linear 3x3 output conv.  Four levels give a receptive field
# wider than a 128^2 cavity, whose flow is global.
ENC = (24, 48, 64, 64)
BOTTLENECK = 64
# large-scale loss terms of the pipeline's training: (pooling levels,
# weight), i.e. errors averaged over 4x4 and 16x16 pixel squares
COARSE = ((2, 4.0), (4, 4.0))


def layer_specs(enc=ENC, bott=BOTTLENECK):
    """[(name, cin, cout, relu, level)]; level = image size divisor 2^level."""
    L = []
    cin = IN_CH
    for k, c in enumerate(enc):
        L += [("e%da" % k, cin, c, 1, k), ("e%db" % k, c, c, 1, k)]
        cin = c
    nl = len(enc)
    L += [("ba", cin, bott, 1, nl), ("bb", bott, bott, 1, nl)]
    cin = bott
    for k in reversed(range(nl)):
        c = enc[k]
        L += [("d%da" % k, cin + c, c, 1, k), ("d%db" % k, c, c, 1, k)]
        cin = c
    L.append(("out", cin, OUT_CH, 0, 0))
    return L


LAYERS = layer_specs()


def param_count(layers=None):
    return sum(9 * ci * co + co for _, ci, co, _, _ in (layers or LAYERS))


class UNet:
    """U-Net on n x n images (n divisible by 2^levels), batch `batch`;
    parameters in FP32 with Adam state; activations kept for backward."""

    def __init__(self, n=128, batch=32, seed=0, enc=ENC, bott=BOTTLENECK, coarse=()):
        """coarse: ((pools, weight), ...) adds, per entry, the squared error
        of the fluid-masked prediction and the target both averaged over
        2^pools x 2^pools pixel squares, times weight (see loss_grad_device)."""
        self.lib = load()
        self.n, self.batch = n, batch
        self.coarse = tuple((int(k), float(w)) for k, w in coarse)
        if any(k < 1 or n % (1 << k) for k, _ in self.coarse):
            raise WarmStartError("coarse levels must divide the image size")
        self.enc, self.bott = tuple(enc), bott
        self.layers = layer_specs(self.enc, bott)
        nl = len(self.enc)
        if n % (1 << nl):
            raise WarmStartError("image size %d is not divisible by %d" % (n, 1 << nl))
        rng = np.random.default_rng(seed)
        self.w, self.b = {}, {}
        self.gw, self.gb, self.state = {}, {}, {}
        self.spec = {}
        for name, ci, co, relu, lev in self.layers:
            std = math.sqrt(2.0 / (9 * ci)) if relu else math.sqrt(1.0 / (9 * ci))
            w = (rng.standard_normal((9, ci, co)) * std).astype(np.float32)
            self.w[name] = DeviceArray.from_numpy(w)
            self.b[name] = DeviceArray((co,), zero=True)
            self.gw[name] = DeviceArray(w.shape, zero=True)
            self.gb[name] = DeviceArray((co,), zero=True)
            self.state[name] = [DeviceArray(w.shape, zero=True), DeviceArray(w.shape, zero=True),
                                DeviceArray((co,), zero=True), DeviceArray((co,), zero=True)]
            self.spec[name] = (ci, co, relu, n >> lev)
        self.step_count = 0
        self._alloc(batch)

    def _alloc(self, batch):
        n, nl = self.n, len(self.enc)
        self.batch = batch
        e = lambda s, c: DeviceArray((batch, s, s, c))
        a = {"x": e(n, IN_CH)}
        for name, ci, co, relu, lev in self.layers:
            a[name] = e(n >> lev, co)
        for k, c in enumerate(self.enc):
            a["p%d" % k] = e(n >> (k + 1), c)            # pooled level k
            below = self.bott if k == nl - 1 else self.enc[k + 1]
            a["c%d" % k] = e(n >> k, below + c)           # decoder input of level k
        self.act = a
        self.grad = {k: DeviceArray(v.shape) for k, v in a.items()}
        self.gout = DeviceArray(a["out"].shape)
        ws, wf = 0, 0
        for name, (ci, co, _, s) in self.spec.items():
            ws = max(ws, self.lib.mplbws_wgrad_workspace(batch, s, s, ci, co))
            wf = max(wf, 9 * ci * co)
        self.workspace = DeviceArray((int(ws),))
        self.wflip = DeviceArray((wf,))
        self.lossbuf = DeviceArray((LOSS_PARTIALS,), zero=True)
        self._tgt = DeviceArray((batch, n, n, 3))
        self._msk = DeviceArray((batch, n, n))
        # multi-scale loss: masked prediction, then per pooling level the
        # pooled prediction, target and gradient (index 0 = full size)
        kmax = max([k for k, _ in self.coarse], default=0)
        self._pm = DeviceArray((batch, n, n, OUT_CH)) if kmax else None
        self._pp = [None] + [DeviceArray((batch, n >> k, n >> k, OUT_CH)) for k in range(1, kmax + 1)]
        self._pt = [None] + [DeviceArray((batch, n >> k, n >> k, 3)) for k in range(1, kmax + 1)]
        self._pg = [DeviceArray((batch, n >> k, n >> k, OUT_CH)) for k in range(kmax + 1)]
        self._pgo = [None] + [DeviceArray((batch, n >> k, n >> k, OUT_CH)) if k < kmax else None
                              for k in range(1, kmax + 1)]
        self._pzero = [None] + [DeviceArray((batch, n >> k, n >> k), zero=True)
                                for k in range(1, kmax + 1)]
        self._pparts = DeviceArray((LOSS_PARTIALS,), zero=True) if kmax else None

    # -- building blocks
    def _conv(self, name, src, dst):
        ci, co, relu, s = self.spec[name]
        _ck(self.lib.mplbws_conv3x3_fwd(_p(src), _p(self.w[name]), _p(self.b[name]), _p(dst),
                                        self.batch, s, s, ci, co, relu), name)

    def _conv_bwd(self, name, src, out, dout, din):
        """dout: gradient at the layer output (overwritten with the
        pre-activation gradient for ReLU layers); din: input gradient (or
        None for the first layer)."""
        ci, co, relu, s = self.spec[name]
        if relu:
            _ck(self.lib.mplbws_relu_bwd(_p(dout), _p(out), _p(dout), dout.numel()), "relu")
        _ck(self.lib.mplbws_conv3x3_wgrad(_p(src), _p(dout), _p(self.gw[name]), _p(self.gb[name]),
                                          _p(self.workspace), self.batch, s, s, ci, co), "wgrad")
        if din is not None:
            _ck(self.lib.mplbws_flip_weights(_p(self.w[name]), _p(self.wflip), ci, co), "flip")
            _ck(self.lib.mplbws_conv3x3_fwd(_p(dout), _p(self.wflip), None, _p(din), self.batch,
                                            s, s, co, ci, 0), "dgrad")

    def forward(self, x):
        """x: (batch, n, n, IN_CH) host array, or None when the input
        activation is already on the device; returns the output activation
        (device, (batch, n, n, OUT_CH), channels 0..2 used)."""
        a, L, n, B, nl = self.act, self.lib, self.n, self.batch, len(self.enc)
        if x is not None:
            a["x"].upload(x)
        src = a["x"]
        for k, c in enumerate(self.enc):
            self._conv("e%da" % k, src, a["e%da" % k])
            self._conv("e%db" % k, a["e%da" % k], a["e%db" % k])
            s = n >> k
            _ck(L.mplbws_pool2_fwd(_p(a["e%db" % k]), _p(a["p%d" % k]), B, s, s, c), "pool")
            src = a["p%d" % k]
        self._conv("ba", src, a["ba"])
        self._conv("bb", a["ba"], a["bb"])
        cur, cc = a["bb"], self.bott
        for k in reversed(range(nl)):
            c, s = self.enc[k], n >> (k + 1)
            cat = a["c%d" % k]
            _ck(L.mplbws_up2_fwd(_p(cur), _p(cat), B, s, s, cc, cc + c, 0), "up")
            _ck(L.mplbws_channels(_p(a["e%db" % k]), _p(cat), B * 4 * s * s, c, cc + c, cc, 0),
                "cat")
            self._conv("d%da" % k, cat, a["d%da" % k])
            self._conv("d%db" % k, a["d%da" % k], a["d%db" % k])
            cur, cc = a["d%db" % k], c
        self._conv("out", cur, a["out"])
        return a["out"]

    def backward(self, gout):
        """gout: gradient at the output activation; fills self.gw/self.gb."""
        a, g, L, n, B, nl = self.act, self.grad, self.lib, self.n, self.batch, len(self.enc)
        self._conv_bwd("out", a["d0b"], a["out"], gout, g["d0b"])
        for k in range(nl):
            c, s = self.enc[k], n >> (k + 1)
            cc = self.bott if k == nl - 1 else self.enc[k + 1]
            self._conv_bwd("d%db" % k, a["d%da" % k], a["d%db" % k], g["d%db" % k], g["d%da" % k])
            self._conv_bwd("d%da" % k, a["c%d" % k], a["d%da" % k], g["d%da" % k], g["c%d" % k])
            below = "bb" if k == nl - 1 else "d%db" % (k + 1)
            _ck(L.mplbws_up2_bwd(_p(g["c%d" % k]), _p(g[below]), B, s, s, cc, cc + c, 0), "up")
        self._conv_bwd("bb", a["ba"], a["bb"], g["bb"], g["ba"])
        self._conv_bwd("ba", a["p%d" % (nl - 1)], a["ba"], g["ba"], g["p%d" % (nl - 1)])
        for k in reversed(range(nl)):
            c, s = self.enc[k], n >> k
            cc = self.bott if k == nl - 1 else self.enc[k + 1]
            # e_kb feeds the pool and the skip into c_k
            _ck(L.mplbws_pool2_bwd(_p(g["p%d" % k]), _p(g["e%db" % k]), B, s, s, c), "pool")
            _ck(L.mplbws_channels(_p(g["c%d" % k]), _p(g["e%db" % k]), B * s * s, c, cc + c, cc,
                                  2), "cat")
            self._conv_bwd("e%db" % k, a["e%da" % k], a["e%db" % k], g["e%db" % k], g["e%da" % k])
            src = a["x"] if k == 0 else a["p%d" % (k - 1)]
            gin = None if k == 0 else g["p%d" % (k - 1)]
            self._conv_bwd("e%da" % k, src, a["e%da" % k], g["e%da" % k], gin)

    def loss_grad(self, target, mask):
        """Masked MSE over the 3 used output channels (the density channel
        weighted by RHO_WEIGHT through its target scale).  target (batch, n,
        n, 3) and mask (batch, n, n) are host arrays.  Returns (loss,
        gradient at the output, device)."""
        fluid = float((np.asarray(mask) == 0).sum())
        inv = 1.0 / max(fluid * 3.0, 1.0)
        self._tgt.upload(target)
        self._msk.upload(mask)
        self.loss_grad_device(inv, _p(self.lossbuf))
        loss = float(self.lossbuf.numpy().astype(np.float64).sum()) * inv
        return loss, self.gout

    def loss_grad_device(self, inv, partial):
        """The loss of the targets and masks already in self._tgt and
        self._msk: gradient into self.gout, LOSS_PARTIALS squared-error
        partial sums at the device pointer `partial` (their sum times inv is
        the loss); nothing is read back."""
        pixels = self.batch * self.n * self.n
        _ck(self.lib.mplbws_mse(_p(self.act["out"]), _p(self._tgt), _p(self._msk), _p(self.gout),
                                partial, LOSS_PARTIALS, pixels, 3, OUT_CH, inv), "mse")
        if self.coarse:
            self._coarse_loss()

    def _coarse_loss(self):
        """Large-scale terms: for each (k, w), w times the mean over the
        2^k-pooled pixels and the 3 channels of (P_k(m p) - P_k(t))^2, with
        m the fluid mask and P_k the 2^k x 2^k average (the targets are 0
        on obstacles).  Their gradient, m P_k^T(2 w (P_k(m p) - P_k(t)) /
        count), is added to self.gout.  Slow, large-scale errors of the
        initial flow cost the most solver steps."""
        L, B, n = self.lib, self.batch, self.n
        kmax = max(k for k, _ in self.coarse)
        _ck(L.mplbws_mask_channels(_p(self.act["out"]), _p(self._msk), _p(self._pm),
                                   B * n * n, OUT_CH, 3), "mask")
        src, tsrc = self._pm, self._tgt
        for k in range(1, kmax + 1):
            s = n >> (k - 1)
            _ck(L.mplbws_pool2_fwd(_p(src), _p(self._pp[k]), B, s, s, OUT_CH), "pool")
            _ck(L.mplbws_pool2_fwd(_p(tsrc), _p(self._pt[k]), B, s, s, 3), "pool")
            src, tsrc = self._pp[k], self._pt[k]
        weight = dict(self.coarse)
        # gradient with respect to level k's pooled prediction: its own term
        # plus, carried down by the pooling adjoint, level k + 1's gradient
        # (already in self._pg[k] for k < kmax when the loop gets there)
        for k in range(kmax, 0, -1):
            s = n >> k
            pix = B * s * s
            if k in weight:
                own = self._pg[k] if k == kmax else self._pgo[k]
                _ck(L.mplbws_mse(_p(self._pp[k]), _p(self._pt[k]), _p(self._pzero[k]), _p(own),
                                 _p(self._pparts), LOSS_PARTIALS, pix, 3, OUT_CH,
                                 weight[k] / (pix * 3.0)), "mse")
                if k < kmax:
                    _ck(L.mplbws_channels(_p(own), _p(self._pg[k]), pix, OUT_CH, OUT_CH, 0, 2),
                        "add")
            _ck(L.mplbws_pool2_bwd(_p(self._pg[k]), _p(self._pg[k - 1]), B, 2 * s, 2 * s,
                                   OUT_CH), "pool")
        _ck(L.mplbws_mask_channels(_p(self._pg[0]), _p(self._msk), _p(self._pg[0]),
                                   B * n * n, OUT_CH, 3), "mask")
        _ck(L.mplbws_channels(_p(self._pg[0]), _p(self.gout), B * n * n, OUT_CH, OUT_CH, 0, 2),
            "add")

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for name in self.w:
            m, v, mb, vb = self.state[name]
            _ck(self.lib.mplbws_adam(_p(self.w[name]), _p(self.gw[name]), _p(m), _p(v),
                                     self.w[name].numel(), lr, b1, b2, eps, self.step_count),
                "adam")
            _ck(self.lib.mplbws_adam(_p(self.b[name]), _p(self.gb[name]), _p(mb), _p(vb),
                                     self.b[name].numel(), lr, b1, b2, eps, self.step_count),
                "adam")

    def state_dict(self):
        return {**{"w_" + k: v.numpy() for k, v in self.w.items()},
                **{"b_" + k: v.numpy() for k, v in self.b.items()}}

    def load_state(self, d):
        for k in self.w:
            self.w[k].upload(d["w_" + k])
            self.b[k].upload(d["b_" + k])


def save_weights(net, path, n_train):
    """The trained network as an .npz file: every layer's weights and
    biases plus its shape (grid size, encoder widths, bottleneck) and the
    number of training cases, which fixes the held-out range of
    case_params."""
    meta = np.array([net.n, net.bott, n_train] + list(net.enc), dtype=np.int64)
    np.savez(path, __meta__=meta, **net.state_dict())


def load_weights(path, batch=32):
    """A UNet with the weights save_weights wrote, and its number of
    training cases."""
    d = np.load(path, allow_pickle=False)
    meta = [int(x) for x in d["__meta__"]]
    n, bott, n_train, enc = meta[0], meta[1], meta[2], tuple(meta[3:])
    net = UNet(n=n, batch=batch, enc=enc, bott=bott)
    net.load_state({k: d[k] for k in d.files if k != "__meta__"})
    return net, n_train


# ---- data -------------------------------------------------------------------
def case_params(count, seed, re_range=(100.0, 1000.0), n=128):
    """Re per case (splitmix64 stream, log-uniform in re_range) and the
    obstacle rectangles (cases.obstacle_rects; sides 8-32 nodes at n = 128,
    scaled with n)."""
    i = np.arange(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = cases.splitmix64((2 * i + np.uint64(1)) ^ np.uint64(seed))
    u = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    lo, hi = math.log(re_range[0]), math.log(re_range[1])
    res = np.exp(lo + (hi - lo) * u)
    side = (max(2, 8 * n // 128), max(3, 32 * n // 128))
    return [float(r) for r in res], cases.obstacle_rects(count, n=n, seed=seed, side=side)


def inputs(n, res, rects, re_range=(100.0, 1000.0), sdf=True):
    """Network inputs (count, n, n, IN_CH): mask, normalised log Re, lid
    speed / 0.05, lid row, x and y coordinates in [0, 1], and (sdf) the
    Euclidean distance to the nearest wall or obstacle node / (n / 2) and
    to the nearest obstacle node / n (1 without obstacles).  A stack of
    3x3 convolutions cannot measure long distances itself; the distance
    fields hand it the geometry at every pixel."""
    cnt = len(res)
    x = np.zeros((cnt, n, n, IN_CH), np.float32)
    lo, hi = math.log(re_range[0]), math.log(re_range[1])
    yy, xx = np.meshgrid(np.arange(n) / (n - 1), np.arange(n) / (n - 1), indexing="ij")
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    # the walls are the node frame just outside the n x n nodes
    d_wall = np.minimum(np.minimum(ix + 1, n - ix), np.minimum(iy + 1, n - iy)).astype(float)
    for k in range(cnt):
        solid = cases.rect_mask(n, rects[k])
        x[k, :, :, 0] = solid
        x[k, :, :, 1] = (math.log(res[k]) - lo) / (hi - lo)
        x[k, :, :, 2] = 1.0
        x[k, n - 1, :, 3] = 1.0
        x[k, :, :, 4] = xx
        x[k, :, :, 5] = yy
        if sdf:
            d_obs = np.full((n, n), float(n))
            for x0, y0, w, h in rects[k]:
                # node distance to an axis-aligned box of nodes, in closed form
                dx = np.maximum(np.maximum(x0 - ix, ix - (x0 + w - 1)), 0)
                dy = np.maximum(np.maximum(y0 - iy, iy - (y0 + h - 1)), 0)
                d_obs = np.minimum(d_obs, np.sqrt(dx * dx + dy * dy))
            x[k, :, :, 6] = np.minimum(d_wall, d_obs) / (n / 2)
            x[k, :, :, 7] = d_obs / n
    return x


def solve(n, res, rects, criteria, check_period=1000, max_steps=2000000,
          storage=hip.STORE_F32, init=None, device=0, restart=0.0, timings=None):
    """Run the cavities (one block each) to the velocity residual criteria.
    init: None (rest) or (rho, u, v) arrays (count, n, n) in code units,
    lifted to populations.  restart > 0: an instance whose residual climbs
    above restart times its lowest restarts once at rest.  Returns (case,
    steps per case, residual per case, wall seconds from case set-up to
    convergence).  timings, a dict, receives the seconds of the case
    set-up ("setup"), of the initial state ("init") and of the iteration
    ("iterate")."""
    t0 = time.perf_counter()
    g = hip.LatticeCase(2, storage=storage, device=device)
    masks = [cases.rect_mask(n, r) for r in rects]
    cases.batched_cavities(g, [n] * len(res), res, ulid=LID, masks=masks)
    t1 = time.perf_counter()
    if init is not None:
        rho, u, v = init
        for b in range(len(res)):
            g.set_macros(b, rho[b], u[b], v[b])
        g.init_lift()
    if restart > 0:
        g.set_batch_restart(restart)
    t2 = time.perf_counter()
    steps, resid, _ = g.iterate_batch(criteria, check_period, max_steps)
    g.sync()
    t3 = time.perf_counter()
    if timings is not None:
        timings.update(setup=t1 - t0, init=t2 - t1, iterate=t3 - t2)
    return g, steps, resid, t3 - t0


def targets_from(g, count, n, masks):
    """(count, n, n, 3) targets: u / lid, v / lid and (rho - mean rho) *
    RHO_SCALE over the fluid nodes (lid in code units), from the context's
    current macros.  The mean density is left out: with a moving lid the
    wall scheme changes the total mass over time, and the flow does not
    depend on the density level."""
    g.update_macros()
    lid = cases.CS * LID
    y = np.zeros((count, n, n, 3), np.float32)
    for b in range(count):
        fluid = masks[b] == 0
        rho = g.get_macro(b, 0)
        y[b, :, :, 0] = g.get_macro(b, 1) / lid
        y[b, :, :, 1] = g.get_macro(b, 2) / lid
        y[b, :, :, 2] = np.where(fluid, (rho - rho[fluid].mean()) * RHO_SCALE, 0.0)
    return y


def prediction_to_macros(pred, masks):
    """(rho, u, v) in code units from network outputs (count, n, n, >=3)."""
    lid = cases.CS * LID
    solid = np.stack(masks).astype(bool)
    u = np.where(solid, 0.0, pred[..., 0] * lid)
    v = np.where(solid, 0.0, pred[..., 1] * lid)
    rho = np.where(solid, 1.0, 1.0 + pred[..., 2] / RHO_SCALE)
    return rho, u, v


# ---- the pipeline: data, training, evaluation ---------------------------------
def predict(net, x):
    """Network outputs (count, n, n, 3) for inputs x (count, n, n, IN_CH)."""
    B = net.batch
    out = np.zeros(x.shape[:3] + (3,), np.float32)
    for k0 in range(0, len(x), B):
        xb = np.zeros((B,) + x.shape[1:], np.float32)
        k1 = min(len(x), k0 + B)
        xb[: k1 - k0] = x[k0:k1]
        y = net.forward(xb).numpy()
        out[k0:k1] = y[: k1 - k0, :, :, :3]
    return out


def train(net, x, y, epochs, lr=2e-3, seed=0, log=None):
    """Adam on the masked MSE; x, y host arrays; returns the loss history
    (one mean per epoch).  The learning rate follows a cosine schedule.
    The training set stays on the device: each step gathers its mini-batch
    there, and the losses are read back once per epoch."""
    rng = np.random.default_rng(seed)
    lib, B, n = net.lib, net.batch, net.n
    count = len(x)
    steps_per_epoch = count // B
    total = epochs * steps_per_epoch
    dx = DeviceArray.from_numpy(x)
    dy = DeviceArray.from_numpy(y)
    dm = DeviceArray.from_numpy(x[..., 0])
    fluid = (x[..., 0] == 0).sum(axis=(1, 2)).astype(np.float64)
    dperm = DeviceArray((max(count, 1),))
    parts = DeviceArray((max(steps_per_epoch, 1) * LOSS_PARTIALS,), zero=True)
    hist = []
    it = 0
    for ep in range(epochs):
        perm = rng.permutation(count).astype(np.int32)
        dperm.upload(perm)
        invs = []
        for s in range(steps_per_epoch):
            idx = _at(dperm, s * B)
            _ck(lib.mplbws_gather(_p(dx), _p(net.act["x"]), idx, B, n * n * IN_CH), "gather")
            _ck(lib.mplbws_gather(_p(dy), _p(net._tgt), idx, B, n * n * 3), "gather")
            _ck(lib.mplbws_gather(_p(dm), _p(net._msk), idx, B, n * n), "gather")
            inv = 1.0 / max(fluid[perm[s * B:(s + 1) * B]].sum() * 3.0, 1.0)
            invs.append(inv)
            net.forward(None)
            net.loss_grad_device(inv, _at(parts, s * LOSS_PARTIALS))
            net.backward(net.gout)
            rate = lr * 0.5 * (1 + math.cos(math.pi * it / max(total, 1)))
            net.adam(lr=rate)
            it += 1
        sums = parts.numpy().astype(np.float64).reshape(-1, LOSS_PARTIALS).sum(axis=1)
        losses = sums[:steps_per_epoch] * np.array(invs)
        hist.append(float(losses.mean()) if steps_per_epoch else 0.0)
        if log:
            log("epoch %d loss %.4e" % (ep, hist[-1]))
    return hist


def velocity_l2(g_a, g_b, count):
    """Per case relative L2 distance of the velocity fields of two contexts."""
    out = []
    for b in range(count):
        ua, va = g_a.get_macro(b, 1), g_a.get_macro(b, 2)
        ub, vb = g_b.get_macro(b, 1), g_b.get_macro(b, 2)
        num = np.sqrt(((ua - ub) ** 2 + (va - vb) ** 2).sum())
        den = np.sqrt((ub ** 2 + vb ** 2).sum())
        out.append(float(num / den))
    return np.array(out)


def run(n=128, n_train=4096, n_test=512, epochs=200, criteria=1e-8, check_period=1000,
        seed=0x4D504C42, batch=32, chunk=512, storage=hip.STORE_F32, restart=1e3,
        enc=ENC, bott=BOTTLENECK, coarse=COARSE, diagnose=False, log=print):
    """Time-to-residual of learned versus rest initialisation.

    1. Training targets: n_train cavities (Re log-uniform in [100, 1000],
       0-3 rectangles) solved from rest to `criteria` by the batched solver.
    2. Training of the U-Net.
    3. n_test held-out cavities, solved from rest and from the network's
       prediction lifted to populations; both wall times include set-up,
       and the learned one includes inference.  Fallback: an instance whose
       residual rises above `restart` times its lowest restarts at rest
       inside the batch; one still unconverged when the budget ends is
       re-run from rest and that time is added.
    With diagnose, two more solves split the learned start's error: the
    converged velocity with the predicted density, and the predicted
    velocity with the converged density."""
    t_all = time.perf_counter()
    res, rects = case_params(n_train + n_test, seed, n=n)
    x_train = inputs(n, res[:n_train], rects[:n_train])
    # 1. data
    t0 = time.perf_counter()
    ys, data_steps = [], []
    for k0 in range(0, n_train, chunk):
        k1 = min(n_train, k0 + chunk)
        g, steps, _, _ = solve(n, res[k0:k1], rects[k0:k1], criteria, check_period,
                               storage=storage)
        ys.append(targets_from(g, k1 - k0, n, [cases.rect_mask(n, r) for r in rects[k0:k1]]))
        data_steps.append(steps)
        g.close()
        log("data %d/%d" % (k1, n_train))
    y_train = np.concatenate(ys)
    data_steps = np.concatenate(data_steps)
    t_data = time.perf_counter() - t0
    # 2. training
    t0 = time.perf_counter()
    net = UNet(n=n, batch=batch, seed=seed & 0xFFFF, enc=enc, bott=bott, coarse=coarse)
    hist = train(net, x_train, y_train, epochs, log=log)
    mplbws_sync = load().mplbws_sync
    mplbws_sync()
    t_train = time.perf_counter() - t0
    # 3. evaluation on the held-out cases
    tr, tq = res[n_train:], rects[n_train:]
    ev = evaluate(net, n, tr, tq, criteria, check_period, storage, restart, diagnose, log)
    out = {
        "n": n, "train_cases": n_train, "test_cases": n_test, "criteria": criteria,
        "check_period": check_period, "params": param_count(layer_specs(enc, bott)),
        "epochs": epochs,
        "loss_first": hist[0], "loss_last": hist[-1],
        "data_seconds": round(t_data, 2), "data_steps_median": float(np.median(data_steps)),
        "train_seconds": round(t_train, 2),
        "network": {"encoder": list(enc), "bottleneck": bott, "coarse_loss": list(coarse),
                    "params": param_count(layer_specs(enc, bott))},
    }
    out.update(ev)
    out["total_seconds"] = round(time.perf_counter() - t_all, 1)
    return out, net


def evaluate(net, n, tr, tq, criteria=1e-8, check_period=1000, storage=hip.STORE_F32,
             restart=1e3, diagnose=False, log=print, fp64_cases=0):
    """Time-to-residual of the held-out cases (Reynolds numbers tr,
    rectangles tq) from rest and from the network's prediction lifted to
    populations (wall times include set-up; the learned one includes its
    inputs and inference).  Fallback: an instance whose residual rises above
    `restart` times its lowest restarts at rest inside the batch; one still
    unconverged when the budget ends is re-run from rest and that time is
    added.  fp64_cases > 0: the first that many cases are also solved from
    rest in FP64 storage (bit-exact with the CPU oracle) to the same
    criterion, and both starts' converged velocities are compared with it."""
    n_test = len(tr)
    log("evaluation: %d held-out cases from rest" % n_test)
    tm_rest, tm_warm = {}, {}
    g_rest, s_rest, r_rest, t_rest = solve(n, tr, tq, criteria, check_period, storage=storage,
                                           timings=tm_rest)
    log("rest: %.2f s, steps median %.0f max %d" % (t_rest, np.median(s_rest), s_rest.max()))
    g_rest.update_macros()
    # the learned start's own cost: its inputs, inference and conversion
    t0 = time.perf_counter()
    pred = predict(net, inputs(n, tr, tq))
    masks = [cases.rect_mask(n, r) for r in tq]
    init = prediction_to_macros(pred, masks)
    t_inf = time.perf_counter() - t0
    # a warm start that needs longer than twice the slowest rest start has
    # failed even counting one restart
    budget = 2 * int(s_rest.max()) + check_period
    g_warm, s_warm, r_warm, t_solve = solve(n, tr, tq, criteria, check_period, budget,
                                            storage=storage, init=init, restart=restart,
                                            timings=tm_warm)
    restarts = g_warm.batch_restarts()
    t_warm = t_inf + t_solve
    log("learned: %.2f s (inference %.3f s), steps median %.0f max %d"
        % (t_warm, t_inf, np.median(s_warm), s_warm.max()))
    # fallback: an instance that did not converge restarts from rest
    bad = [b for b in range(n_test) if s_warm[b] < 0]
    t_fb = 0.0
    if bad:
        _, s_fb, _, t_fb = solve(n, [tr[b] for b in bad], [tq[b] for b in bad], criteria,
                                 check_period, storage=storage)
        for k, b in enumerate(bad):
            s_warm[b] = s_fb[k]
    g_warm.update_macros()
    l2 = velocity_l2(g_warm, g_rest, n_test)
    pred_err, rho_err = [], []
    lid = cases.CS * LID
    exact = [np.stack([g_rest.get_macro(b, v) for b in range(n_test)]) for v in range(3)]
    for b in range(n_test):
        u, v = exact[1][b] / lid, exact[2][b] / lid
        num = np.sqrt(((pred[b, :, :, 0] - u) ** 2 + (pred[b, :, :, 1] - v) ** 2).sum())
        pred_err.append(float(num / np.sqrt((u ** 2 + v ** 2).sum())))
        fluid = masks[b] == 0
        dr = exact[0][b][fluid] - exact[0][b][fluid].mean()
        dp = init[0][b][fluid] - init[0][b][fluid].mean()
        rho_err.append(float(np.sqrt(((dp - dr) ** 2).sum() / (dr ** 2).sum())))
    diag = {}
    if diagnose:
        for name, start in (("exact_velocity_predicted_density", (init[0], exact[1], exact[2])),
                            ("predicted_velocity_exact_density", (exact[0], init[1], init[2]))):
            gd, sd, _, td = solve(n, tr, tq, criteria, check_period, budget, storage=storage,
                                  init=start)
            gd.close()
            diag[name] = {"seconds": round(td, 3), "steps_median": float(np.median(sd)),
                          "unconverged": int((sd < 0).sum()),
                          "per_case_step_ratio_median":
                              float(np.median(s_rest / np.maximum(sd, 1)))}
            log("diagnosis %s: %s" % (name, diag[name]))
    out = {
        "test_cases": n_test, "criteria": criteria, "check_period": check_period,
        "rest_seconds": round(t_rest, 3), "learned_seconds": round(t_warm + t_fb, 3),
        "rest_phases_seconds": {k: round(v, 3) for k, v in tm_rest.items()},
        "learned_phases_seconds": {k: round(v, 3) for k, v in tm_warm.items()},
        "inference_seconds": round(t_inf, 4), "fallback_cases": len(bad),
        "restart_factor": restart, "restarted_cases": int((restarts > 0).sum()),
        "speedup": round(t_rest / (t_warm + t_fb), 3),
        "rest_steps_median": float(np.median(s_rest)), "rest_steps_max": int(s_rest.max()),
        "learned_steps_median": float(np.median(s_warm)),
        "learned_steps_max": int(s_warm.max()),
        "per_case_step_ratio_median": float(np.median(s_rest / np.maximum(s_warm, 1))),
        "work_ratio_sum_steps": float(s_rest.sum() / max(s_warm.sum(), 1)),
        "prediction_rel_l2_median": float(np.median(pred_err)),
        "prediction_density_rel_l2_median": float(np.median(rho_err)),
        "velocity_rel_l2_learned_vs_rest_median": float(np.median(l2)),
        "velocity_rel_l2_learned_vs_rest_max": float(l2.max()),
        # the cases that set the batch times
        "slowest_cases": [{"Re": round(tr[b], 1), "rects": len(tq[b]),
                           "rest_steps": int(s_rest[b]), "learned_steps": int(s_warm[b]),
                           "restarted": bool(restarts[b]),
                           "prediction_rel_l2": round(pred_err[b], 4)}
                          for b in np.argsort(-s_rest)[:5]],
    }
    if diag:
        out["diagnosis"] = diag
    if fp64_cases > 0:
        k = min(fp64_cases, n_test)
        g64, s64, _, t64 = solve(n, tr[:k], tq[:k], criteria, check_period,
                                 storage=hip.STORE_F64)
        g64.update_macros()
        lw = velocity_l2(g_warm, g64, k)
        lr = velocity_l2(g_rest, g64, k)
        out["fp64_reference"] = {
            "cases": k, "seconds": round(t64, 3),
            "velocity_rel_l2_learned_vs_fp64_median": float(np.median(lw)),
            "velocity_rel_l2_learned_vs_fp64_max": float(lw.max()),
            "velocity_rel_l2_rest_vs_fp64_median": float(np.median(lr)),
            "velocity_rel_l2_rest_vs_fp64_max": float(lr.max())}
        log("FP64 reference on %d cases: %s" % (k, out["fp64_reference"]))
        g64.close()
    g_rest.close()
    g_warm.close()
    return out


def evaluate_saved(path, n_test=512, seed=0x4D504C42, log=print, fp64_cases=0):
    """The held-out evaluation of the network saved at `path` (trained on
    the first n_train cases of the seeded stream, as run() draws them): the
    n_test cases after them."""
    net, n_train = load_weights(path)
    n = net.n
    res, rects = case_params(n_train + n_test, seed, n=n)
    out = evaluate(net, n, res[n_train:], rects[n_train:], log=log, fp64_cases=fp64_cases)
    out.update({"n": n, "train_cases": n_train, "weights": os.path.basename(path)})
    return out

