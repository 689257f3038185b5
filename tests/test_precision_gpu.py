"""Reduced-precision storage forms against the FP64 path, and full-size,
full-length checks through properties that do not depend on an oracle run.

Tolerances (stated per test) follow from the storage: FP32 keeps g = f - w_i
with ~2^-24 relative rounding per step; FP16 keeps g/s_i with an 11-bit
significand, so its velocity error is ~1e-3 of the lid speed.
"""
import os
import sys

import numpy as np
import pytest

from mplb_amd import cases, hip
from oracle.oracle import OracleCase

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(storage, build, steps):
    g = hip.LatticeCase(2 if build[0] == 2 else 3, storage=storage, device=0)
    p = build[1](g)
    g.step(steps)
    g.update_macros()
    return g, p


@pytest.mark.parametrize("storage,tol_u,tol_rho", [
    (hip.STORE_F32, 1e-4, 1e-5),
    (hip.STORE_F16, 5e-3, 5e-4),
])
def test_reduced_storage_cavity2d_vs_oracle(storage, tol_u, tol_rho):
    """64^2 cavity, Re=100, 3000 steps: max|u - u_FP64| <= tol_u * lid and
    max|rho - rho_FP64| <= tol_rho."""
    name = {1: "FP32", 2: "FP16"}[storage]
    log("%s 2D cavity 64^2 3000 steps vs oracle" % name)
    build = (2, lambda be: cases.cavity2d(be, 64, re=100.0, ulid=0.05))
    g, p = _run(storage, build, 3000)
    o = OracleCase(2)
    build[1](o)
    o.step(3000)
    o.update_macros()
    du = max(float(np.max(np.abs(g.get_macro(0, v) - o.get_macro(0, v))))
             for v in (1, 2)) / p["lid"]
    dr = float(np.max(np.abs(g.get_macro(0, 0) - o.get_macro(0, 0))))
    log("%s: max|du|/lid = %.3e, max|drho| = %.3e" % (name, du, dr))
    assert du <= tol_u and dr <= tol_rho


def test_fp32_cavity3d_vs_oracle():
    log("FP32 3D cavity 32^3 noise 400 steps vs oracle")
    build = (3, lambda be: cases.cavity3d(be, 32, re=400.0, noise=True))
    g, p = _run(hip.STORE_F32, build, 400)
    o = OracleCase(3)
    build[1](o)
    o.step(400)
    o.update_macros()
    du = max(float(np.max(np.abs(g.get_macro(0, v) - o.get_macro(0, v))))
             for v in (1, 2, 3)) / p["lid"]
    df = float(np.max(np.abs(g.get_f(0) - o.get_f(0))))
    log("FP32 3D: max|du|/lid = %.3e, max|df| = %.3e" % (du, df))
    assert du <= 2e-5 and df <= 5e-6


def test_fp32_mass_drift_periodic():
    """A periodic box conserves mass exactly in exact arithmetic; in FP32
    deviation storage the FP64 mass sum may drift only by rounding."""
    log("FP32 periodic mass drift")
    rng = np.random.default_rng(3)
    size = (40, 36, 32)
    shp = tuple(reversed(size))
    g = hip.LatticeCase(3, storage=hip.STORE_F32, device=0)
    cases.periodic_box(g, size, 0.6, 1 + 0.01 * rng.standard_normal(shp),
                       *[0.03 * rng.standard_normal(shp) for _ in range(3)])
    m0 = g.mass()
    g.step(2000)
    m1 = g.mass()
    log("FP32 periodic: relative mass drift %.3e" % ((m1 - m0) / m0))
    assert abs(m1 - m0) / m0 < 1e-6


def _mirror_z(field):
    return field[::-1, :, :]


@pytest.mark.parametrize("storage,tol", [(hip.STORE_F64, 1e-13), (hip.STORE_F32, 1e-6)])
def test_c2_full_size_mirror_symmetry(storage, tol):
    """Config 2 size and length (256^3, 2000 steps) without density noise:
    the cavity with stationary Front and Back walls and the lid moving along
    x is mirror symmetric in z (u, v even; w odd), which the FP64 path keeps
    to round-off and the FP32 path to its storage precision."""
    n = int(os.environ.get("MPLB_C2_N", "256"))
    steps = int(os.environ.get("MPLB_C2_LONG_STEPS", "2000"))
    log("C2 %d^3 storage %d, %d steps, z-mirror symmetry" % (n, storage, steps))
    g = hip.LatticeCase(3, storage=storage, device=0)
    p = cases.cavity3d(g, n, re=400.0, ulid=0.05, noise=False)
    g.step(steps)
    g.update_macros()
    lid = p["lid"]
    u, v, w = (g.get_macro(0, k) for k in (1, 2, 3))
    assert np.max(np.abs(u)) > 0.1 * lid  # the flow did develop
    e = max(float(np.max(np.abs(u - _mirror_z(u)))), float(np.max(np.abs(v - _mirror_z(v)))),
            float(np.max(np.abs(w + _mirror_z(w))))) / lid
    log("mirror asymmetry / lid = %.3e" % e)
    assert e <= tol


@pytest.mark.parametrize("dim,size", [(3, (67, 40, 31)), (3, (128, 36, 20)), (2, (201, 201))])
def test_fp16_two_node_kernel_matches_one_node(monkeypatch, dim, size):
    """FP16 storage: the bulk kernel that moves two neighbouring nodes per
    lane as one half pair (x shifts through lane shuffles) against the
    one-node kernel, on odd and even row lengths.  Both do the same FP32
    arithmetic on the same inputs, up to how the compiler fuses it, so they
    agree to a few FP16 steps of the stored values, and both stay as close
    to an FP64 run."""
    log("FP16 two-node vs one-node kernel, %dD %s" % (dim, size))
    monkeypatch.setenv("MPLBHIP_FUSED", "0")
    monkeypatch.setenv("MPLBHIP_XFACE_FOLD", "0")

    def run(storage, h2):
        monkeypatch.setenv("MPLBHIP_H2", h2)
        g = hip.LatticeCase(dim, storage=storage, device=0)
        if dim == 3:
            cases.cavity3d(g, size[0], re=300.0, ulid=0.05, noise=True, size=size)
        else:
            cases.cavity2d(g, size[0], re=300.0, ulid=0.05)
        g.step(300)
        f = g.get_f(0)
        g.close()
        return f

    f_two, f_one, f_64 = run(hip.STORE_F16, "1"), run(hip.STORE_F16, "0"), run(hip.STORE_F64, "0")
    d = float(np.max(np.abs(f_two - f_one)))
    e_two = float(np.max(np.abs(f_two - f_64)))
    e_one = float(np.max(np.abs(f_one - f_64)))
    log("max |two - one| = %.3e; error vs FP64: two %.3e, one %.3e" % (d, e_two, e_one))
    assert d <= 1e-4
    assert e_two <= 1.5 * e_one + 1e-6
