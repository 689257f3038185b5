"""Reduced-precision storage (FP32 f - w_i, FP16 (f - w_i)/s_i) through every
form of the step that large cases take, against the FP64 oracle.

Cases below 2^18 nodes normally run the whole step as one launch; these
tests force the split step (MPLBHIP_FUSED=0) so that small cases go through
the bulk kernel's raw-buffer pulls and stores, its folded boundary lanes
(fold levels 0-2, where a lane's unpulled source is moved back into the
block), the three bulk workgroup shapes (MPLBHIP_XGROUP 0/1/2) and, for
linked blocks, the FP16/FP32 halo copies with and without the side-stream
boundary kernel.

Tolerances are those of the whole-lattice launch for the same cases
(tests/test_precision_gpu.py): FP32 rounds g = f - w_i to 2^-24 relative per
step, FP16 keeps 11 significant bits of g/s_i, s_i = w_i/2^10.
"""
import sys

import numpy as np
import pytest

from mplb_amd import cases, hip
from oracle.oracle import OracleCase

pytestmark = pytest.mark.gpu

NAME = {hip.STORE_F32: "FP32", hip.STORE_F16: "FP16"}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _errors(g, o, blocks, lid):
    """max |u - u_FP64| / lid over the velocity components and max |rho -
    rho_FP64| over every block, after UpdateMacroVars on both."""
    g.update_macros()
    o.update_macros()
    du = dr = 0.0
    for b in blocks:
        for v in range(1, 1 + g.dim):
            du = max(du, float(np.max(np.abs(g.get_macro(b, v) - o.get_macro(b, v)))))
        dr = max(dr, float(np.max(np.abs(g.get_macro(b, 0) - o.get_macro(b, 0)))))
    return du / lid, dr


SPLIT = [("0", "0"), ("1", "0"), ("2", "0"), ("2", "1"), ("1", "2"), ("2", "2")]


@pytest.mark.parametrize("fold,xgroup", SPLIT)
@pytest.mark.parametrize("storage,tol_u,tol_rho", [
    (hip.STORE_F32, 1e-4, 1e-5),
    (hip.STORE_F16, 5e-3, 5e-4),
])
def test_split_step_cavity2d_vs_oracle(monkeypatch, fold, xgroup, storage, tol_u, tol_rho):
    """64^2 cavity (Re 100, 3000 steps) on the split step."""
    monkeypatch.setenv("MPLBHIP_FUSED", "0")
    monkeypatch.setenv("MPLBHIP_XFACE_FOLD", fold)
    monkeypatch.setenv("MPLBHIP_XGROUP", xgroup)
    monkeypatch.setenv("MPLBHIP_H2", "0")
    log("%s 2D cavity 64^2 split fold=%s xgroup=%s, 3000 steps" % (NAME[storage], fold, xgroup))
    g = hip.LatticeCase(2, storage=storage, device=0)
    p = cases.cavity2d(g, 64, re=100.0, ulid=0.05)
    o = OracleCase(2)
    cases.cavity2d(o, 64, re=100.0, ulid=0.05)
    g.step(3000)
    o.step(3000)
    du, dr = _errors(g, o, [0], p["lid"])
    log("  max|du|/lid = %.3e, max|drho| = %.3e" % (du, dr))
    assert du <= tol_u and dr <= tol_rho
    g.close()


@pytest.mark.parametrize("fold,xgroup", SPLIT)
@pytest.mark.parametrize("storage,tol_u,tol_f", [
    (hip.STORE_F32, 2e-5, 5e-6),
    (hip.STORE_F16, 2e-3, 5e-5),
])
def test_split_step_cavity3d_vs_oracle(monkeypatch, fold, xgroup, storage, tol_u, tol_f):
    """32^3 cavity with density noise (Re 400, 400 steps) on the split step;
    the f tolerance bounds every population of every node."""
    monkeypatch.setenv("MPLBHIP_FUSED", "0")
    monkeypatch.setenv("MPLBHIP_XFACE_FOLD", fold)
    monkeypatch.setenv("MPLBHIP_XGROUP", xgroup)
    monkeypatch.setenv("MPLBHIP_H2", "0")
    log("%s 3D cavity 32^3 split fold=%s xgroup=%s, 400 steps" % (NAME[storage], fold, xgroup))
    g = hip.LatticeCase(3, storage=storage, device=0)
    p = cases.cavity3d(g, 32, re=400.0, noise=True)
    o = OracleCase(3)
    cases.cavity3d(o, 32, re=400.0, noise=True)
    g.step(400)
    o.step(400)
    df = float(np.max(np.abs(g.get_f(0) - o.get_f(0))))
    du, dr = _errors(g, o, [0], p["lid"])
    log("  max|du|/lid = %.3e, max|drho| = %.3e, max|df| = %.3e" % (du, dr, df))
    assert du <= tol_u and df <= tol_f
    g.close()


@pytest.mark.parametrize("storage,tol_u,tol_f", [
    (hip.STORE_F32, 2e-5, 5e-6),
    (hip.STORE_F16, 2e-3, 5e-5),
])
@pytest.mark.parametrize("fold", ["0", "1", "2"])
def test_row_workgroups_reduced_vs_oracle(monkeypatch, storage, tol_u, tol_f, fold):
    """The row-workgroup bulk shape (four 64-node tiles of one row per
    workgroup; rows of a multiple of 256 nodes) in reduced precision."""
    monkeypatch.setenv("MPLBHIP_FUSED", "0")
    monkeypatch.setenv("MPLBHIP_XFACE_FOLD", fold)
    monkeypatch.setenv("MPLBHIP_XGROUP", "2")
    monkeypatch.setenv("MPLBHIP_H2", "0")
    log("%s box 256x8x12 row workgroups fold=%s, 60 steps" % (NAME[storage], fold))
    build = lambda be: cases.cavity3d(be, 12, re=150.0, noise=True, size=(256, 8, 12))  # noqa: E731
    g = hip.LatticeCase(3, storage=storage, device=0)
    p = build(g)
    o = OracleCase(3)
    build(o)
    g.step(60)
    o.step(60)
    assert g.launch_shape()[1] == 2
    df = float(np.max(np.abs(g.get_f(0) - o.get_f(0))))
    du, dr = _errors(g, o, [0], p["lid"])
    log("  max|du|/lid = %.3e, max|drho| = %.3e, max|df| = %.3e" % (du, dr, df))
    assert du <= tol_u and df <= tol_f
    g.close()


# ---- linked blocks: the L-channel of the C3 configuration ----------------------
@pytest.mark.parametrize("overlap", ["0", "1"])
@pytest.mark.parametrize("fold", ["0", "2"])
@pytest.mark.parametrize("storage,tol_u,tol_rho,tol_mass", [
    (hip.STORE_F32, 2e-4, 2e-5, 2e-5),
    (hip.STORE_F16, 2e-3, 2e-5, 2e-5),
])
def test_lchannel_reduced_vs_oracle(monkeypatch, overlap, fold, storage, tol_u, tol_rho,
                                    tol_mass):
    """The closed three-block L-channel (VirtualBoundary halos, the rotated
    periodic link, per-block body force, FP16/FP32 halo copies) at a = 9,
    length 14, g = 1e-4, 300 steps, split step with the boundary kernel on
    the side stream or in line.  Velocity error relative to the largest
    FP64 speed; FP64 mass drift of the reduced run relative to the oracle's
    mass.  At a = 9 most nodes are wall nodes, whose scheme runs in FP32 on
    full values (f, not f - w_i): its constants and wall density round to
    FP32 at every step, which sets the mass difference (measured 4.6e-6
    after 300 steps in both storage forms, FP32 and FP16 alike) and the
    density error."""
    monkeypatch.setenv("MPLBHIP_FUSED", "0")
    monkeypatch.setenv("MPLBHIP_OVERLAP", overlap)
    monkeypatch.setenv("MPLBHIP_XFACE_FOLD", fold)
    monkeypatch.setenv("MPLBHIP_H2", "0")
    log("%s closed L-channel a=9 fold=%s overlap=%s, 300 steps" % (NAME[storage], fold, overlap))
    g = hip.LatticeCase(3, storage=storage, device=0)
    cases.lchannel_periodic(g, a=9, length=14, g=1e-4)
    o = OracleCase(3)
    cases.lchannel_periodic(o, a=9, length=14, g=1e-4)
    g.step(300)
    o.step(300)
    o.update_macros()
    umax = max(float(np.max(np.abs(o.get_macro(b, v)))) for b in range(3) for v in (1, 2, 3))
    assert umax > 1e-3  # the loop carries flow
    du, dr = _errors(g, o, range(3), umax)
    dm = abs(g.mass() - o.mass()) / o.mass()
    log("  umax %.3e: max|du|/umax = %.3e, max|drho| = %.3e, mass rel %.3e" % (umax, du, dr, dm))
    assert du <= tol_u and dr <= tol_rho and dm <= tol_mass
    g.close()


def test_c3_lchannel_fp16_full_size_vs_fp64():
    """The C3 configuration as configured: blocks 128 x 256 x 128,
    128^3 and 256 x 128 x 128 (10,485,760 nodes), tau_L = 0.8, g = 1e-6,
    FP16 storage, against the FP64 GPU run of the same case (bit-exact with
    the oracle elsewhere), 500 steps, FP64 mass taken every 100 steps.
    Apps/3DLChannel/lbm3d_L.cpp:72-190 is the reference shape."""
    log("C3 L-channel 128/256 FP16 vs FP64 GPU, 500 steps")
    runs = {}
    for st in (hip.STORE_F64, hip.STORE_F16):
        g = hip.LatticeCase(3, storage=st, device=0)
        p = cases.lchannel_periodic(g, a=128, length=256, tau_l=0.8, g=1e-6)
        assert p["nodes"] == 10485760
        masses = [g.mass()]
        for _ in range(5):
            g.step(100)
            masses.append(g.mass())
        g.update_macros()
        runs[st] = (g, masses)
        log("  storage %d: masses %s" % (st, ["%.12g" % m for m in masses]))
    g64, m64 = runs[hip.STORE_F64]
    g16, m16 = runs[hip.STORE_F16]
    umax = max(float(np.max(np.abs(g64.get_macro(b, v)))) for b in range(3) for v in (1, 2, 3))
    du = max(float(np.max(np.abs(g16.get_macro(b, v) - g64.get_macro(b, v))))
             for b in range(3) for v in (1, 2, 3))
    dr = max(float(np.max(np.abs(g16.get_macro(b, 0) - g64.get_macro(b, 0)))) for b in range(3))
    drift = max(abs(m - m16[0]) / m16[0] for m in m16)
    log("  umax %.3e, max|du| %.3e (%.3e of umax), max|drho| %.3e, FP16 mass drift %.3e, "
        "FP64 mass drift %.3e" % (umax, du, du / umax, dr, drift,
                                  max(abs(m - m64[0]) / m64[0] for m in m64)))
    assert umax > 1e-5
    assert du <= 0.05 * umax and dr <= 1e-4 and drift <= 1e-5
    for g, _ in runs.values():
        g.close()
