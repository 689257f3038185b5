"""CPU oracle of the warm-start network layers (test infrastructure only:
the tests compare libmplbws.so against it; nothing in the product imports
it).  The warm start has no reference counterpart; these are the textbook
definitions of the layers, computed in float64 from BF16-rounded operands
where the GPU rounds them, so that the only expected difference is the FP32
accumulation order.

Layout: channels last, x[n][y][x][c]; 3x3 weights w[tap = 3*ky + kx][cin][cout],
zero padding, stride 1.
"""
import numpy as np


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))


def conv3x3(x, w, b=None, relu=False, round_bf16=True):
    """out[n,y,x,co] = sum_{ky,kx,ci} x[n,y+ky-1,x+kx-1,ci] w[3ky+kx,ci,co] (+ b)."""
    xx = bf16(x) if round_bf16 else np.asarray(x, np.float64)
    ww = bf16(w) if round_bf16 else np.asarray(w, np.float64)
    n, h, wd, cin = xx.shape
    xp = _pad(xx)
    out = np.zeros((n, h, wd, ww.shape[2]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        out += xp[:, ky:ky + h, kx:kx + wd, :] @ ww[tap]
    if b is not None:
        out += np.asarray(b, np.float64)
    if relu:
        out = np.maximum(out, 0.0)
    return out


def conv3x3_wgrad(x, dout, round_bf16=True):
    """dw[tap,ci,co] = sum x[n,y+ky-1,x+kx-1,ci] dout[n,y,x,co]; db = sum dout."""
    xx = bf16(x) if round_bf16 else np.asarray(x, np.float64)
    dd = bf16(dout) if round_bf16 else np.asarray(dout, np.float64)
    n, h, wd, cin = xx.shape
    xp = _pad(xx)
    dw = np.zeros((9, cin, dd.shape[3]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        sl = xp[:, ky:ky + h, kx:kx + wd, :].reshape(-1, cin)
        dw[tap] = sl.T @ dd.reshape(-1, dd.shape[3])
    return dw, np.asarray(dout, np.float64).sum(axis=(0, 1, 2))


def conv3x3_dgrad(dout, w, round_bf16=True):
    """The adjoint of conv3x3 in x: din[n,y,x,ci] = sum_{tap,co}
    dout[n,y-ky+1,x-kx+1,co] w[tap,ci,co] (scatter form, independent of the
    weight flip the GPU uses)."""
    dd = bf16(dout) if round_bf16 else np.asarray(dout, np.float64)
    ww = bf16(w) if round_bf16 else np.asarray(w, np.float64)
    n, h, wd, cout = dd.shape
    cin = ww.shape[1]
    dinp = np.zeros((n, h + 2, wd + 2, cin))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        dinp[:, ky:ky + h, kx:kx + wd, :] += dd @ ww[tap].T
    return dinp[:, 1:h + 1, 1:wd + 1, :]


def pool2(x):
    n, h, wd, c = x.shape
    return np.asarray(x, np.float64).reshape(n, h // 2, 2, wd // 2, 2, c).mean(axis=(2, 4))


def up2(x):
    return np.repeat(np.repeat(np.asarray(x, np.float64), 2, axis=1), 2, axis=2)
