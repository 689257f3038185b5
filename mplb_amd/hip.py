"""ctypes binding of libmplbhip.so (include/mplbhip.h).

LatticeCase mirrors the HiLeMMS call sequence (DefineBlocks,
DefineBlockConnection, DefineComponents/Collision/BodyForce,
DefineBlockBoundary, Partition, SetInitialMacrosVars,
PreDefinedInitialCondition, SetTimeStep, StreamCollision / Iterate,
UpdateMacroVars, CalcResidualError).  Errors raise MplbError with the
library's message.  There is no fallback: without the built library (or
without a GPU for compute calls) every call raises.
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmplbhip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "mplbhip.h")

STORE_F64, STORE_F32, STORE_F16 = 0, 1, 2
_lib = None


class MplbError(RuntimeError):
    pass


def header_functions():
    """Names of the entry points include/mplbhip.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mplbhip_[a-z0-9_]+)\s*\(", text)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MplbError("libmplbhip.so is not built (run __graft_entry__.build() "
                        "or python -m mplb_amd.build)")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    PP = ctypes.POINTER(ctypes.c_void_p)
    I = ctypes.c_int
    D = ctypes.c_double
    DP = ctypes.POINTER(ctypes.c_double)
    IP = ctypes.POINTER(ctypes.c_int)
    LLP = ctypes.POINTER(ctypes.c_longlong)
    sig = {
        "mplbhip_create": (I, [PP, I, I, I]),
        "mplbhip_destroy": (None, [P]),
        "mplbhip_last_error": (ctypes.c_char_p, [P]),
        "mplbhip_version": (ctypes.c_char_p, []),
        "mplbhip_define_block": (I, [P, I, IP]),
        "mplbhip_define_connection": (I, [P, I, I, I, I, I]),
        "mplbhip_define_rotated_connection": (I, [P, I, I, I, I, I]),
        "mplbhip_define_obstacles": (I, [P, I, ctypes.c_void_p]),
        "mplbhip_define_collision": (I, [P, D, I, I]),
        "mplbhip_define_scheme": (I, [P, I]),
        "mplbhip_define_lattice": (I, [P, I]),
        "mplbhip_define_collision_type": (I, [P, I]),
        "mplbhip_set_temperature": (I, [P, I, DP]),
        "mplbhip_define_boundary": (I, [P, I, I, I, DP, I, I]),
        "mplbhip_prepare": (I, [P]),
        "mplbhip_node_info": (I, [P, I, I, IP]),
        "mplbhip_set_macros": (I, [P, I, DP, DP, DP, DP]),
        "mplbhip_init_feq": (I, [P]),
        "mplbhip_init_lift": (I, [P]),
        "mplbhip_set_f": (I, [P, I, DP]),
        "mplbhip_set_timestep": (I, [P, D]),
        "mplbhip_set_accel": (I, [P, I, DP]),
        "mplbhip_step": (I, [P, I]),
        "mplbhip_step_timed": (I, [P, I, DP, DP]),
        "mplbhip_kernel_times": (I, [P, DP, DP]),
        "mplbhip_update_macros": (I, [P]),
        "mplbhip_get_f": (I, [P, I, DP]),
        "mplbhip_get_macro": (I, [P, I, I, DP]),
        "mplbhip_residual": (I, [P, I, DP, DP]),
        "mplbhip_mass": (I, [P, DP]),
        "mplbhip_mass_noalloc": (I, [P, DP]),
        "mplbhip_time_step": (I, [P, LLP]),
        "mplbhip_sync": (I, [P]),
        "mplbhip_copy_probe": (I, [I, ctypes.c_longlong, I, DP]),
        "mplbhip_iterate": (I, [P, D, I, ctypes.c_longlong, LLP, DP]),
        "mplbhip_last_residual": (I, [P, DP, DP]),
        "mplbhip_define_block_collision": (I, [P, I, D, D]),
        "mplbhip_iterate_batch": (I, [P, D, I, ctypes.c_longlong, LLP, DP, LLP]),
        "mplbhip_set_batch_restart": (I, [P, D]),
        "mplbhip_launch_shape": (I, [P, IP, IP, IP, LLP]),
        "mplbhip_batch_restarts": (I, [P, IP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def copy_probe(nbytes, iters=20, device=0):
    """HIP-event average of a plain nbytes device copy (ms)."""
    ms = ctypes.c_double()
    rc = load().mplbhip_copy_probe(device, int(nbytes), int(iters), ctypes.byref(ms))
    if rc != 0:
        raise MplbError("copy probe failed (%d)" % rc)
    return ms.value


class LatticeCase:
    """One lattice case (one component) on one GPU, or a plan-only case
    with device=-1 (node typing and boundary programs, no compute)."""

    def __init__(self, space_dim, storage=STORE_F64, device=0, lattice=None):
        """lattice: the number of velocities, 9 (D2Q9), 19 (D3Q19, the 3D
        default) or 15 (D3Q15)."""
        self.lib = load()
        self.dim = space_dim
        self.Q = 9 if space_dim == 2 else 19
        self.storage = storage
        h = ctypes.c_void_p()
        rc = self.lib.mplbhip_create(ctypes.byref(h), device, space_dim, storage)
        if rc != 0:
            raise MplbError("mplbhip_create failed (%d): no usable device %d?"
                            % (rc, device))
        self.h = h
        self.sizes = {}
        if lattice is not None:
            self._ck(self.lib.mplbhip_define_lattice(self.h, int(lattice)))
            self.Q = int(lattice)

    def close(self):
        if getattr(self, "h", None):
            self.lib.mplbhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc != 0:
            raise MplbError(self.lib.mplbhip_last_error(self.h).decode())

    # ---- case definition ------------------------------------------------
    def add_block(self, bid, size):
        size = list(size)[: self.dim]
        arr = (ctypes.c_int * 3)(*(size + [1] * (3 - len(size))))
        self._ck(self.lib.mplbhip_define_block(self.h, bid, arr))
        self.sizes[bid] = tuple(size)

    def add_connection(self, frm, fsurf, to, tsurf, vtype):
        self._ck(self.lib.mplbhip_define_connection(self.h, frm, int(fsurf), to,
                                                    int(tsurf), int(vtype)))

    def set_obstacles(self, block, solid):
        m = np.ascontiguousarray(np.asarray(solid, dtype=np.uint8))
        self._ck(self.lib.mplbhip_define_obstacles(self.h, block, m.ctypes.data))

    def add_rotated_connection(self, frm, fsurf, to, tsurf, vtype):
        self._ck(self.lib.mplbhip_define_rotated_connection(self.h, frm, int(fsurf), to,
                                                            int(tsurf), int(vtype)))

    def set_scheme(self, scheme):
        """DefineScheme: 10 Scheme_StreamCollision (default), 11
        Scheme_StreamCollision_Swap (D3Q19, FP64 storage)."""
        self._ck(self.lib.mplbhip_define_scheme(self.h, int(scheme)))

    def set_collision_type(self, ctype):
        """DefineCollision: 0 Collision_BGKIsothermal2nd (default), 1
        Collision_BGKThermal4th (FP64 storage)."""
        self._ck(self.lib.mplbhip_define_collision_type(self.h, int(ctype)))

    def set_collision(self, tau_ref, force_type=0, force_macros=False):
        self._ck(self.lib.mplbhip_define_collision(self.h, tau_ref, int(force_type),
                                                   int(bool(force_macros))))

    def add_boundary(self, block, surface, scheme, values=(), vtype=1000):
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).ravel())
        n = v.size
        if n == 0:
            v = np.zeros(1)
        self._ck(self.lib.mplbhip_define_boundary(self.h, block, int(surface),
                                                  int(scheme), _dp(v), n, int(vtype)))

    def prepare(self):
        self._ck(self.lib.mplbhip_prepare(self.h))

    def _shape(self, block):
        return tuple(reversed(self.sizes[block]))

    def node_info(self, block, which):
        out = np.empty(self._shape(block), dtype=np.int32)
        self._ck(self.lib.mplbhip_node_info(
            self.h, block, int(which), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    # ---- state ------------------------------------------------------------
    def set_macros(self, block, rho, u, v, w=None):
        shp = self._shape(block)
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shp))
                for a in (rho, u, v, w if w is not None else np.zeros(shp))]
        self._ck(self.lib.mplbhip_set_macros(self.h, block, *[_dp(a) for a in arrs]))

    def set_temperature(self, block, T):
        """Variable_T of one block, interior [k][j][i] (thermal collision)."""
        T = np.ascontiguousarray(np.broadcast_to(np.asarray(T, dtype=np.float64),
                                                 self._shape(block)))
        self._ck(self.lib.mplbhip_set_temperature(self.h, block, _dp(T)))

    def init_feq(self):
        self._ck(self.lib.mplbhip_init_feq(self.h))

    def init_lift(self):
        """feq of the uploaded macros plus the first-order non-equilibrium
        part (warm start from a predicted flow)."""
        self._ck(self.lib.mplbhip_init_lift(self.h))

    def set_f(self, block, f):
        f = np.ascontiguousarray(np.asarray(f, dtype=np.float64))
        if f.shape != (self.Q,) + self._shape(block):
            raise ValueError("f must have shape %s" % ((self.Q,) + self._shape(block),))
        self._ck(self.lib.mplbhip_set_f(self.h, block, _dp(f)))

    def set_dt(self, dt):
        self._ck(self.lib.mplbhip_set_timestep(self.h, dt))

    def set_accel(self, block, g):
        g = np.ascontiguousarray(np.asarray(list(g) + [0.0] * (3 - len(g)),
                                            dtype=np.float64))
        self._ck(self.lib.mplbhip_set_accel(self.h, block, _dp(g)))

    # ---- evolution ----------------------------------------------------------
    def step(self, n=1):
        self._ck(self.lib.mplbhip_step(self.h, int(n)))

    def step_timed(self, n):
        tot = ctypes.c_double()
        ker = ctypes.c_double()
        self._ck(self.lib.mplbhip_step_timed(self.h, int(n), ctypes.byref(tot),
                                             ctypes.byref(ker)))
        return tot.value, ker.value

    def kernel_times(self):
        """(bulk ms, boundary ms) summed over the last step_timed call."""
        b = ctypes.c_double()
        d = ctypes.c_double()
        self._ck(self.lib.mplbhip_kernel_times(self.h, ctypes.byref(b), ctypes.byref(d)))
        return b.value, d.value

    def iterate(self, criteria, check_period, max_steps):
        """Iterate(convergenceCriteria, checkPeriod) when criteria >= 0,
        Iterate(steps, checkPeriod) when criteria < 0 (max_steps steps).
        Returns (steps taken, last residual or -1)."""
        done = ctypes.c_longlong()
        res = ctypes.c_double()
        self._ck(self.lib.mplbhip_iterate(self.h, float(criteria), int(check_period),
                                          int(max_steps), ctypes.byref(done),
                                          ctypes.byref(res)))
        return done.value, res.value

    def set_block_collision(self, block, tau_ref, dt):
        self._ck(self.lib.mplbhip_define_block_collision(self.h, block, float(tau_ref),
                                                         float(dt)))

    def iterate_batch(self, criteria, check_period, max_steps):
        """Advance all instances, freezing each at its own convergence.
        Returns (steps per block (-1: not converged), residual per block,
        steps taken)."""
        nb = len(self.sizes)
        steps = np.zeros(nb, dtype=np.int64)
        res = np.zeros(nb)
        done = ctypes.c_longlong()
        self._ck(self.lib.mplbhip_iterate_batch(
            self.h, float(criteria), int(check_period), int(max_steps),
            steps.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), _dp(res),
            ctypes.byref(done)))
        return steps, res, done.value

    def launch_shape(self):
        """(fold level, bulk workgroup shape, two-node FP16 kernel, nodes one
        bulk launch updates) of the step in force."""
        f, x, h = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        nodes = ctypes.c_longlong()
        self._ck(self.lib.mplbhip_launch_shape(self.h, ctypes.byref(f), ctypes.byref(x),
                                               ctypes.byref(h), ctypes.byref(nodes)))
        return f.value, x.value, h.value, nodes.value

    def set_batch_restart(self, factor):
        """Restart, once, a batched instance whose residual climbs above
        factor times its lowest (0: off)."""
        self._ck(self.lib.mplbhip_set_batch_restart(self.h, float(factor)))

    def batch_restarts(self):
        out = np.zeros(len(self.sizes), dtype=np.int32)
        self._ck(self.lib.mplbhip_batch_restarts(
            self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def update_macros(self):
        self._ck(self.lib.mplbhip_update_macros(self.h))

    def get_f(self, block):
        out = np.empty((self.Q,) + self._shape(block), dtype=np.float64)
        self._ck(self.lib.mplbhip_get_f(self.h, block, _dp(out)))
        return out

    def get_macro(self, block, var):
        out = np.empty(self._shape(block), dtype=np.float64)
        self._ck(self.lib.mplbhip_get_macro(self.h, block, int(var), _dp(out)))
        return out

    def residual(self, check_period):
        """CalcResidualError + GetMaximumResidual: (max residual, per-variable
        ratios sum (m - m_copy)^2 / sum m^2)."""
        per = np.zeros(4)
        mx = ctypes.c_double()
        self._ck(self.lib.mplbhip_residual(self.h, int(check_period), _dp(per),
                                           ctypes.byref(mx)))
        return mx.value, per[: 1 + self.dim]

    def mass(self):
        m = ctypes.c_double()
        self._ck(self.lib.mplbhip_mass(self.h, ctypes.byref(m)))
        return m.value

    def mass_noalloc(self):
        """The FP64 total mass through the entry point that allocates no
        device memory (partial sums in the library's static scratch)."""
        m = ctypes.c_double()
        self._ck(self.lib.mplbhip_mass_noalloc(self.h, ctypes.byref(m)))
        return m.value

    def time_step(self):
        t = ctypes.c_longlong()
        self._ck(self.lib.mplbhip_time_step(self.h, ctypes.byref(t)))
        return t.value

    def sync(self):
        self._ck(self.lib.mplbhip_sync(self.h))
