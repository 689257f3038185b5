"""Lattice velocity sets with the reference's ordering and weights:
D2Q9 Src/model.cpp:99-112, D3Q19 Src/model.cpp:114-131 (opposites found by
FindReverseXi, Src/model.cpp:82-97).  CS = sqrt(3) (Src/model.cpp:67-68)."""
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


def opposite(c):
    q = len(c)
    opp = np.empty(q, dtype=np.int64)
    for i in range(q):
        for j in range(q):
            if np.all(c[i] + c[j] == 0):
                opp[i] = j
                break
    return opp


def lattice(dim):
    if dim == 2:
        return D2Q9_C, D2Q9_W
    if dim == 3:
        return D3Q19_C, D3Q19_W
    raise ValueError(dim)


def feq(dim, rho, u):
    """Second-order equilibrium in code units, as CalcBGKFeq with T = 1
    (Src/model_host_device.h:34-83); u has shape (dim, ...)."""
    c, w = lattice(dim)
    u = np.asarray(u, dtype=np.float64)
    cu = np.tensordot(CS * c.astype(np.float64), u, axes=(1, 0))
    u2 = np.sum(u * u, axis=0)
    return w.reshape((-1,) + (1,) * u2.ndim) * rho * (1.0 + cu + 0.5 * (cu * cu - u2))
