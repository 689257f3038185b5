"""ctypes binding of the CPU oracle (oracle/mplb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and
by the benchmark's cpu_baseline leg, as the checker.  The product package
never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libmplb_oracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    D = ctypes.c_double
    DP = ctypes.POINTER(ctypes.c_double)
    IP = ctypes.POINTER(ctypes.c_int)
    sig = {
        "orc_create": (P, [I, I]),
        "orc_destroy": (None, [P]),
        "orc_error": (ctypes.c_char_p, [P]),
        "orc_add_block": (I, [P, I, I, I, I]),
        "orc_add_connection": (I, [P, I, I, I, I, I]),
        "orc_add_rotated_connection": (I, [P, I, I, I, I, I]),
        "orc_set_obstacles": (I, [P, I, ctypes.c_void_p]),
        "orc_set_collision": (I, [P, D, I, I]),
        "orc_add_boundary": (I, [P, I, I, I, DP, I, I]),
        "orc_prepare": (I, [P]),
        "orc_set_macros": (I, [P, I, DP, DP, DP, DP]),
        "orc_init_feq": (I, [P]),
        "orc_set_dt": (None, [P, D]),
        "orc_set_accel": (I, [P, I, DP]),
        "orc_step": (I, [P, I]),
        "orc_apply_boundaries": (I, [P]),
        "orc_set_scheme": (I, [P, I]),
        "orc_set_collision_type": (I, [P, I]),
        "orc_set_temperature": (I, [P, I, DP]),
        "orc_feq_t": (D, [P, I, D, D, D, D, D, I]),
        "orc_update_macros": (None, [P]),
        "orc_get_f": (I, [P, I, DP]),
        "orc_set_f": (I, [P, I, DP]),
        "orc_get_macro": (I, [P, I, I, DP]),
        "orc_get_int": (I, [P, I, I, IP]),
        "orc_mass": (D, [P]),
        "orc_residual": (D, [P, I, DP]),
        "orc_set_threads": (None, [I]),
        "orc_get_threads": (I, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleCase:
    """One case on the CPU oracle; the method set matches the HIP-side
    mplb_amd.LatticeCase so a test can drive either with the same code."""

    def __init__(self, space_dim, lattice=None):
        """lattice: the number of velocities, 9 (D2Q9), 19 (D3Q19, the 3D
        default) or 15 (D3Q15)."""
        self.lib = load()
        if lattice is None:
            lattice = 9 if space_dim == 2 else 19
        self.dim = space_dim
        self.Q = int(lattice)
        self.h = self.lib.orc_create(space_dim, {9: 0, 19: 1, 15: 2}.get(self.Q, -1))
        if not self.h:
            raise ValueError("unsupported dimension/lattice pair")
        self.sizes = {}

    def __del__(self):
        if getattr(self, "h", None) and self.lib is not None:
            self.lib.orc_destroy(self.h)
            self.h = None

    def _ck(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.orc_error(self.h).decode())

    def add_block(self, bid, size):
        size = list(size) + [1] * (3 - len(size))
        self._ck(self.lib.orc_add_block(self.h, bid, *size[:3]))
        self.sizes[bid] = tuple(size[: self.dim])

    def add_connection(self, frm, fsurf, to, tsurf, vtype):
        self._ck(self.lib.orc_add_connection(self.h, frm, int(fsurf), to,
                                             int(tsurf), int(vtype)))

    def set_obstacles(self, block, solid):
        m = np.ascontiguousarray(np.asarray(solid, dtype=np.uint8))
        self._ck(self.lib.orc_set_obstacles(self.h, block, m.ctypes.data))

    def add_rotated_connection(self, frm, fsurf, to, tsurf, vtype):
        self._ck(self.lib.orc_add_rotated_connection(self.h, frm, int(fsurf), to,
                                                     int(tsurf), int(vtype)))

    def set_collision(self, tau_ref, force_type=0, force_macros=False):
        self._ck(self.lib.orc_set_collision(self.h, tau_ref, int(force_type),
                                            int(bool(force_macros))))

    def add_boundary(self, block, surface, scheme, values=(), vtype=1000):
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
        if v.size == 0:
            v = np.zeros(1)
            n = 0
        else:
            n = v.size
        self._ck(self.lib.orc_add_boundary(self.h, block, int(surface),
                                           int(scheme), _dp(v), n,
                                           int(vtype)))

    def prepare(self):
        self._ck(self.lib.orc_prepare(self.h))

    def _shape(self, block):
        return tuple(reversed(self.sizes[block]))

    def set_macros(self, block, rho, u, v, w=None):
        shp = self._shape(block)
        arrs = []
        for a in (rho, u, v, w if w is not None else np.zeros(shp)):
            arrs.append(np.ascontiguousarray(np.broadcast_to(
                np.asarray(a, dtype=np.float64), shp)))
        self._ck(self.lib.orc_set_macros(self.h, block, *[_dp(a) for a in arrs]))

    def init_feq(self):
        self._ck(self.lib.orc_init_feq(self.h))

    def set_dt(self, dt):
        self.lib.orc_set_dt(self.h, dt)

    def set_accel(self, block, g):
        g = np.ascontiguousarray(np.asarray(list(g) + [0.0] * (3 - len(g)),
                                            dtype=np.float64))
        self._ck(self.lib.orc_set_accel(self.h, block, _dp(g)))

    def step(self, n=1):
        self._ck(self.lib.orc_step(self.h, int(n)))

    def update_macros(self):
        self.lib.orc_update_macros(self.h)

    def set_scheme(self, scheme):
        """SchemeType: 10 StreamCollision (default), 11 StreamCollision_Swap."""
        self._ck(self.lib.orc_set_scheme(self.h, int(scheme)))

    def set_collision_type(self, ctype):
        """0 Collision_BGKIsothermal2nd, 1 Collision_BGKThermal4th"""
        self._ck(self.lib.orc_set_collision_type(self.h, int(ctype)))

    def set_temperature(self, bid, T):
        T = np.ascontiguousarray(T, dtype=np.float64)
        assert T.shape == tuple(reversed(self.sizes[bid])), T.shape
        self._ck(self.lib.orc_set_temperature(self.h, bid, _dp(T)))

    def feq_t(self, q, rho, u, v, w, T, order):
        """CalcBGKFeq at temperature T and polynomial order `order`"""
        return self.lib.orc_feq_t(self.h, int(q), rho, u, v, w, T, int(order))

    def apply_boundaries(self):
        """The boundary stage of a step alone, on the current f."""
        self._ck(self.lib.orc_apply_boundaries(self.h))

    def get_f(self, block):
        out = np.empty((self.Q,) + self._shape(block), dtype=np.float64)
        self._ck(self.lib.orc_get_f(self.h, block, _dp(out)))
        return out

    def set_f(self, block, f):
        f = np.ascontiguousarray(np.asarray(f, dtype=np.float64))
        self._ck(self.lib.orc_set_f(self.h, block, _dp(f)))

    def get_macro(self, block, var):
        out = np.empty(self._shape(block), dtype=np.float64)
        self._ck(self.lib.orc_get_macro(self.h, block, int(var), _dp(out)))
        return out

    def get_int(self, block, which):
        out = np.empty(self._shape(block), dtype=np.int32)
        self._ck(self.lib.orc_get_int(self.h, block, int(which),
                                      out.ctypes.data_as(
                                          ctypes.POINTER(ctypes.c_int))))
        return out

    def mass(self):
        return self.lib.orc_mass(self.h)

    def residual(self, check_period):
        per = np.zeros(4)
        r = self.lib.orc_residual(self.h, int(check_period), _dp(per))
        return r, per[: 1 + self.dim]


def set_threads(n):
    load().orc_set_threads(int(n))


def get_threads():
    return load().orc_get_threads()
