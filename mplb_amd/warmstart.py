"""Learned warm start for batched 2D cavities with obstacles.

A U-Net (about 212k parameters) maps (obstacle mask, log Re, lid speed, lid
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
RHO_SCALE = 1.0 / (3 * LID * LID)  # scale of the density-deviation target
IN_CH = 8                      # input channels (6 used, padded to 8)
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
        a = np.ascontiguousarray(a, dtype=np.float32)
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


# ---- the network -----------------------------------------------------------
# (name, cin, cout, relu); the forward graph is fixed in UNet.forward
LAYERS = [("e1a", IN_CH, 16, 1), ("e1b", 16, 16, 1), ("e2a", 16, 32, 1), ("e2b", 32, 32, 1),
          ("e3a", 32, 64, 1), ("e3b", 64, 64, 1), ("ba", 64, 64, 1), ("bb", 64, 64, 1),
          ("d3a", 128, 32, 1), ("d3b", 32, 32, 1), ("d2a", 64, 16, 1), ("d2b", 16, 16, 1),
          ("d1a", 32, 16, 1), ("d1b", 16, 16, 1), ("out", 16, OUT_CH, 0)]


def param_count():
    return sum(9 * ci * co + co for _, ci, co, _ in LAYERS)


class UNet:
    """Three-level U-Net on n x n images (n divisible by 8), batch `batch`;
    parameters in FP32 with Adam state; activations kept for backward."""

    def __init__(self, n=128, batch=32, seed=0):
        self.lib = load()
        self.n, self.batch = n, batch
        rng = np.random.default_rng(seed)
        self.w, self.b = {}, {}
        self.gw, self.gb, self.state = {}, {}, {}
        for name, ci, co, relu in LAYERS:
            std = math.sqrt(2.0 / (9 * ci)) if relu else math.sqrt(1.0 / (9 * ci))
            w = (rng.standard_normal((9, ci, co)) * std).astype(np.float32)
            self.w[name] = DeviceArray.from_numpy(w)
            self.b[name] = DeviceArray((co,), zero=True)
            self.gw[name] = DeviceArray(w.shape, zero=True)
            self.gb[name] = DeviceArray((co,), zero=True)
            self.state[name] = [DeviceArray(w.shape, zero=True), DeviceArray(w.shape, zero=True),
                                DeviceArray((co,), zero=True), DeviceArray((co,), zero=True)]
        self.spec = {name: (ci, co, relu) for name, ci, co, relu in LAYERS}
        self.step_count = 0
        self._alloc(batch)

    def _alloc(self, batch):
        n = self.n
        self.batch = batch
        e = lambda s, c: DeviceArray((batch, s, s, c))
        self.act = {
            "x": e(n, IN_CH), "e1a": e(n, 16), "e1b": e(n, 16), "p1": e(n // 2, 16),
            "e2a": e(n // 2, 32), "e2b": e(n // 2, 32), "p2": e(n // 4, 32),
            "e3a": e(n // 4, 64), "e3b": e(n // 4, 64), "p3": e(n // 8, 64),
            "ba": e(n // 8, 64), "bb": e(n // 8, 64), "c3": e(n // 4, 128),
            "d3a": e(n // 4, 32), "d3b": e(n // 4, 32), "c2": e(n // 2, 64),
            "d2a": e(n // 2, 16), "d2b": e(n // 2, 16), "c1": e(n, 32),
            "d1a": e(n, 16), "d1b": e(n, 16), "out": e(n, OUT_CH)}
        self.grad = {k: DeviceArray(v.shape) for k, v in self.act.items()}
        self.gout = DeviceArray(self.act["out"].shape)
        ws = 0
        for name, (ci, co, _) in self.spec.items():
            s = self._size(name)
            ws = max(ws, self.lib.mplbws_wgrad_workspace(batch, s, s, ci, co))
        self.workspace = DeviceArray((int(ws),))
        self.wflip = DeviceArray((9 * 128 * 64,))
        self.lossbuf = DeviceArray((1024,), zero=True)

    def _size(self, name):
        n = self.n
        return {"e1": n, "e2": n // 2, "e3": n // 4, "ba": n // 8, "bb": n // 8,
                "d3": n // 4, "d2": n // 2, "d1": n, "ou": n}[name[:2]]

    # -- building blocks
    def _conv(self, name, src, dst):
        ci, co, relu = self.spec[name]
        s = self._size(name)
        _ck(self.lib.mplbws_conv3x3_fwd(_p(src), _p(self.w[name]), _p(self.b[name]), _p(dst),
                                        self.batch, s, s, ci, co, relu), name)

    def _conv_bwd(self, name, src, out, dout, din):
        """dout: gradient at the layer output (overwritten with the
        pre-activation gradient for ReLU layers); din: input gradient (or
        None for the first layer)."""
        ci, co, relu = self.spec[name]
        s = self._size(name)
        if relu:
            _ck(self.lib.mplbws_relu_bwd(_p(dout), _p(out), _p(dout), dout.numel()), "relu")
        _ck(self.lib.mplbws_conv3x3_wgrad(_p(src), _p(dout), _p(self.gw[name]), _p(self.gb[name]),
                                          _p(self.workspace), self.batch, s, s, ci, co), "wgrad")
        if din is not None:
            _ck(self.lib.mplbws_flip_weights(_p(self.w[name]), _p(self.wflip), ci, co), "flip")
            _ck(self.lib.mplbws_conv3x3_fwd(_p(dout), _p(self.wflip), None, _p(din), self.batch,
                                            s, s, co, ci, 0), "dgrad")

    def forward(self, x):
        """x: (batch, n, n, IN_CH) host array; returns the output activation
        (device, (batch, n, n, OUT_CH), channels 0..2 used)."""
        a, L, n, B = self.act, self.lib, self.n, self.batch
        a["x"].upload(x)
        self._conv("e1a", a["x"], a["e1a"])
        self._conv("e1b", a["e1a"], a["e1b"])
        _ck(L.mplbws_pool2_fwd(_p(a["e1b"]), _p(a["p1"]), B, n, n, 16), "pool")
        self._conv("e2a", a["p1"], a["e2a"])
        self._conv("e2b", a["e2a"], a["e2b"])
        _ck(L.mplbws_pool2_fwd(_p(a["e2b"]), _p(a["p2"]), B, n // 2, n // 2, 32), "pool")
        self._conv("e3a", a["p2"], a["e3a"])
        self._conv("e3b", a["e3a"], a["e3b"])
        _ck(L.mplbws_pool2_fwd(_p(a["e3b"]), _p(a["p3"]), B, n // 4, n // 4, 64), "pool")
        self._conv("ba", a["p3"], a["ba"])
        self._conv("bb", a["ba"], a["bb"])
        s = n // 8
        _ck(L.mplbws_up2_fwd(_p(a["bb"]), _p(a["c3"]), B, s, s, 64, 128, 0), "up")
        _ck(L.mplbws_channels(_p(a["e3b"]), _p(a["c3"]), B * 4 * s * s, 64, 128, 64, 0), "cat")
        self._conv("d3a", a["c3"], a["d3a"])
        self._conv("d3b", a["d3a"], a["d3b"])
        s = n // 4
        _ck(L.mplbws_up2_fwd(_p(a["d3b"]), _p(a["c2"]), B, s, s, 32, 64, 0), "up")
        _ck(L.mplbws_channels(_p(a["e2b"]), _p(a["c2"]), B * 4 * s * s, 32, 64, 32, 0), "cat")
        self._conv("d2a", a["c2"], a["d2a"])
        self._conv("d2b", a["d2a"], a["d2b"])
        s = n // 2
        _ck(L.mplbws_up2_fwd(_p(a["d2b"]), _p(a["c1"]), B, s, s, 16, 32, 0), "up")
        _ck(L.mplbws_channels(_p(a["e1b"]), _p(a["c1"]), B * 4 * s * s, 16, 32, 16, 0), "cat")
        self._conv("d1a", a["c1"], a["d1a"])
        self._conv("d1b", a["d1a"], a["d1b"])
        self._conv("out", a["d1b"], a["out"])
        return a["out"]

    def backward(self, gout):
        """gout: gradient at the output activation; fills self.gw/self.gb."""
        a, g, L, n, B = self.act, self.grad, self.lib, self.n, self.batch
        self._conv_bwd("out", a["d1b"], a["out"], gout, g["d1b"])
        self._conv_bwd("d1b", a["d1a"], a["d1b"], g["d1b"], g["d1a"])
        self._conv_bwd("d1a", a["c1"], a["d1a"], g["d1a"], g["c1"])
        s = n // 2
        _ck(L.mplbws_up2_bwd(_p(g["c1"]), _p(g["d2b"]), B, s, s, 16, 32, 0), "up")
        self._conv_bwd("d2b", a["d2a"], a["d2b"], g["d2b"], g["d2a"])
        self._conv_bwd("d2a", a["c2"], a["d2a"], g["d2a"], g["c2"])
        s = n // 4
        _ck(L.mplbws_up2_bwd(_p(g["c2"]), _p(g["d3b"]), B, s, s, 32, 64, 0), "up")
        self._conv_bwd("d3b", a["d3a"], a["d3b"], g["d3b"], g["d3a"])
        self._conv_bwd("d3a", a["c3"], a["d3a"], g["d3a"], g["c3"])
        s = n // 8
        _ck(L.mplbws_up2_bwd(_p(g["c3"]), _p(g["bb"]), B, s, s, 64, 128, 0), "up")
        self._conv_bwd("bb", a["ba"], a["bb"], g["bb"], g["ba"])
        self._conv_bwd("ba", a["p3"], a["ba"], g["ba"], g["p3"])
        # e3b feeds the pool and the skip into c3
        _ck(L.mplbws_pool2_bwd(_p(g["p3"]), _p(g["e3b"]), B, n // 4, n // 4, 64), "pool")
        _ck(L.mplbws_channels(_p(g["c3"]), _p(g["e3b"]), B * (n // 4) ** 2, 64, 128, 64, 2),
            "cat")
        self._conv_bwd("e3b", a["e3a"], a["e3b"], g["e3b"], g["e3a"])
        self._conv_bwd("e3a", a["p2"], a["e3a"], g["e3a"], g["p2"])
        _ck(L.mplbws_pool2_bwd(_p(g["p2"]), _p(g["e2b"]), B, n // 2, n // 2, 32), "pool")
        _ck(L.mplbws_channels(_p(g["c2"]), _p(g["e2b"]), B * (n // 2) ** 2, 32, 64, 32, 2),
            "cat")
        self._conv_bwd("e2b", a["e2a"], a["e2b"], g["e2b"], g["e2a"])
        self._conv_bwd("e2a", a["p1"], a["e2a"], g["e2a"], g["p1"])
        _ck(L.mplbws_pool2_bwd(_p(g["p1"]), _p(g["e1b"]), B, n, n, 16), "pool")
        _ck(L.mplbws_channels(_p(g["c1"]), _p(g["e1b"]), B * n * n, 16, 32, 16, 2), "cat")
        self._conv_bwd("e1b", a["e1a"], a["e1b"], g["e1b"], g["e1a"])
        self._conv_bwd("e1a", a["x"], a["e1a"], g["e1a"], None)

    def loss_grad(self, target, mask):
        """Masked MSE over the 3 used output channels.  target (batch, n, n,
        3) and mask (batch, n, n) are host arrays.  Returns (loss, gradient
        at the output, device)."""
        pixels = self.batch * self.n * self.n
        fluid = float((np.asarray(mask) == 0).sum())
        inv = 1.0 / max(fluid * 3.0, 1.0)
        if not hasattr(self, "_tgt"):
            self._tgt = DeviceArray((self.batch, self.n, self.n, 3))
            self._msk = DeviceArray((self.batch, self.n, self.n))
        self._tgt.upload(target)
        self._msk.upload(mask)
        nb = self.lossbuf.numel()
        _ck(self.lib.mplbws_mse(_p(self.act["out"]), _p(self._tgt), _p(self._msk), _p(self.gout),
                                _p(self.lossbuf), nb, pixels, 3, OUT_CH, inv), "mse")
        loss = float(self.lossbuf.numpy().astype(np.float64).sum()) * inv
        return loss, self.gout

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


def inputs(n, res, rects, re_range=(100.0, 1000.0)):
    """Network inputs (count, n, n, IN_CH): mask, normalised log Re, lid
    speed / 0.05, lid row, x and y coordinates in [0, 1]."""
    cnt = len(res)
    x = np.zeros((cnt, n, n, IN_CH), np.float32)
    lo, hi = math.log(re_range[0]), math.log(re_range[1])
    yy, xx = np.meshgrid(np.arange(n) / (n - 1), np.arange(n) / (n - 1), indexing="ij")
    for k in range(cnt):
        x[k, :, :, 0] = cases.rect_mask(n, rects[k])
        x[k, :, :, 1] = (math.log(res[k]) - lo) / (hi - lo)
        x[k, :, :, 2] = 1.0
        x[k, n - 1, :, 3] = 1.0
        x[k, :, :, 4] = xx
        x[k, :, :, 5] = yy
    return x


def solve(n, res, rects, criteria, check_period=1000, max_steps=2000000,
          storage=hip.STORE_F32, init=None, device=0):
    """Run the cavities (one block each) to the velocity residual criteria.
    init: None (rest) or (rho, u, v) arrays (count, n, n) in code units,
    lifted to populations.  Returns (case, steps per case, residual per case,
    wall seconds from case set-up to convergence)."""
    t0 = time.perf_counter()
    g = hip.LatticeCase(2, storage=storage, device=device)
    masks = [cases.rect_mask(n, r) for r in rects]
    cases.batched_cavities(g, [n] * len(res), res, ulid=LID, masks=masks)
    if init is not None:
        rho, u, v = init
        for b in range(len(res)):
            g.set_macros(b, rho[b], u[b], v[b])
        g.init_lift()
    steps, resid, _ = g.iterate_batch(criteria, check_period, max_steps)
    g.sync()
    return g, steps, resid, time.perf_counter() - t0


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
    (one mean per epoch).  The learning rate follows a cosine schedule."""
    rng = np.random.default_rng(seed)
    B = net.batch
    steps_per_epoch = len(x) // B
    total = epochs * steps_per_epoch
    hist = []
    it = 0
    for ep in range(epochs):
        perm = rng.permutation(len(x))
        acc = 0.0
        for s in range(steps_per_epoch):
            idx = np.sort(perm[s * B:(s + 1) * B])
            xb = x[idx]
            net.forward(xb)
            loss, g = net.loss_grad(y[idx], xb[..., 0])
            net.backward(g)
            rate = lr * 0.5 * (1 + math.cos(math.pi * it / max(total, 1)))
            net.adam(lr=rate)
            acc += loss
            it += 1
        hist.append(acc / max(steps_per_epoch, 1))
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


def run(n=128, n_train=4096, n_test=512, epochs=30, criteria=1e-8, check_period=1000,
        seed=0x4D504C42, batch=32, chunk=512, storage=hip.STORE_F32, log=print):
    """Time-to-residual of learned versus rest initialisation.

    1. Training targets: n_train cavities (Re log-uniform in [100, 1000],
       0-3 rectangles) solved from rest to `criteria` by the batched solver.
    2. Training of the U-Net.
    3. n_test held-out cavities, solved from rest and from the network's
       prediction lifted to populations; both wall times include set-up,
       and the learned one includes inference.  Instances whose residual
       rises after the warm start are re-run from rest (the fallback) and
       that time is added."""
    t_all = time.perf_counter()
    res, rects = case_params(n_train + n_test, seed, n=n)
    x_all = inputs(n, res, rects)
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
    net = UNet(n=n, batch=batch, seed=seed & 0xFFFF)
    hist = train(net, x_all[:n_train], y_train, epochs, log=log)
    mplbws_sync = load().mplbws_sync
    mplbws_sync()
    t_train = time.perf_counter() - t0
    # 3. evaluation on the held-out cases
    tr, tq = res[n_train:], rects[n_train:]
    log("evaluation: %d held-out cases from rest" % n_test)
    g_rest, s_rest, r_rest, t_rest = solve(n, tr, tq, criteria, check_period, storage=storage)
    log("rest: %.2f s, steps median %.0f max %d" % (t_rest, np.median(s_rest), s_rest.max()))
    g_rest.update_macros()
    t0 = time.perf_counter()
    pred = predict(net, x_all[n_train:])
    masks = [cases.rect_mask(n, r) for r in tq]
    init = prediction_to_macros(pred, masks)
    t_inf = time.perf_counter() - t0
    # a warm start that needs longer than the slowest rest start has failed
    budget = int(s_rest.max()) + check_period
    g_warm, s_warm, r_warm, t_solve = solve(n, tr, tq, criteria, check_period, budget,
                                            storage=storage, init=init)
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
    pred_err = []
    lid = cases.CS * LID
    for b in range(n_test):
        u, v = g_rest.get_macro(b, 1) / lid, g_rest.get_macro(b, 2) / lid
        num = np.sqrt(((pred[b, :, :, 0] - u) ** 2 + (pred[b, :, :, 1] - v) ** 2).sum())
        pred_err.append(float(num / np.sqrt((u ** 2 + v ** 2).sum())))
    out = {
        "n": n, "train_cases": n_train, "test_cases": n_test, "criteria": criteria,
        "check_period": check_period, "params": param_count(), "epochs": epochs,
        "loss_first": hist[0], "loss_last": hist[-1],
        "data_seconds": round(t_data, 2), "data_steps_median": float(np.median(data_steps)),
        "train_seconds": round(t_train, 2),
        "rest_seconds": round(t_rest, 3), "learned_seconds": round(t_warm + t_fb, 3),
        "inference_seconds": round(t_inf, 4), "fallback_cases": len(bad),
        "speedup": round(t_rest / (t_warm + t_fb), 3),
        "rest_steps_median": float(np.median(s_rest)), "rest_steps_max": int(s_rest.max()),
        "learned_steps_median": float(np.median(s_warm)),
        "learned_steps_max": int(s_warm.max()),
        "prediction_rel_l2_median": float(np.median(pred_err)),
        "velocity_rel_l2_learned_vs_rest_median": float(np.median(l2)),
        "velocity_rel_l2_learned_vs_rest_max": float(l2.max()),
        "total_seconds": round(time.perf_counter() - t_all, 1),
    }
    g_rest.close()
    g_warm.close()
    return out, net
