"""Lattice velocity sets with the reference's ordering and weights:
D2Q9 Src/model.cpp:99-112, D3Q19 Src/model.cpp:114-131, D3Q15
Src/model.cpp:133-150 (opposites found by FindReverseXi,
Src/model.cpp:82-97).  CS = sqrt(3) (Src/model.cpp:67-68)."""
import math

import numpy as np

CS = math.sqrt(3.0)

D2Q9_C = np.array([[0, 0], [-1, 1], [-1, 0], [-1, -1], [0, -1], [1, -1],
                   [1, 0], [1, 1], [0, 1]], dtype=np.int64)
D2Q9_W = np.array([4 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9,
                   1 / 36, 1 / 9])

D3Q19_C = np.array([
    [0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0],
    [-1, 0, -1], [-1, 0, 1], [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0],
    [0, 0, 1], [1, 1, 0], [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1],
    [0, 1, -1]], dtype=np.int64)
D3Q19_W = np.array([1 / 3] + [1 / 18] * 3 + [1 / 36] * 6 + [1 / 18] * 3 +
                   [1 / 36] * 6)


D3Q15_C = np.array([
    [0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, -1], [-1, -1, 1],
    [-1, 1, -1], [-1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1],
    [1, 1, -1], [1, -1, 1], [1, -1, -1]], dtype=np.int64)
D3Q15_W = np.array([2 / 9] + [1 / 9] * 3 + [1 / 72] * 4 + [1 / 9] * 3 + [1 / 72] * 4)


def opposite(c):
    q = len(c)
    opp = np.empty(q, dtype=np.int64)
    for i in range(q):
        for j in range(q):
            if np.all(c[i] + c[j] == 0):
                opp[i] = j
                break
    return opp


def lattice(dim, q=None):
    """(velocities, weights) of D2Q9 (dim 2), D3Q19 (dim 3) or D3Q15
    (dim 3, q = 15)."""
    if dim == 2 and q in (None, 9):
        return D2Q9_C, D2Q9_W
    if dim == 3 and q in (None, 19):
        return D3Q19_C, D3Q19_W
    if dim == 3 and q == 15:
        return D3Q15_C, D3Q15_W
    raise ValueError((dim, q))


def feq(dim, rho, u, q=None):
    """Second-order equilibrium in code units, as CalcBGKFeq with T = 1
    (Src/model_host_device.h:34-83); u has shape (dim, ...)."""
    c, w = lattice(dim, q)
    u = np.asarray(u, dtype=np.float64)
    cu = np.tensordot(CS * c.astype(np.float64), u, axes=(1, 0))
    u2 = np.sum(u * u, axis=0)
    return w.reshape((-1,) + (1,) * u2.ndim) * rho * (1.0 + cu + 0.5 * (cu * cu - u2))
