"""D3Q15 (SetupD3Q15Latt, Src/model.cpp:133-150) and Collision_BGKThermal4th
(KerCollideBGKThermal[3D], Src/model_kernel.inc:267-310, 457-502) on the GPU
against the CPU oracle: FP64 storage bit for bit (populations and macros;
mass to 1e-13, its summation order differs), reduced storage within stated
tolerances."""
import math
import sys

import numpy as np
import pytest

from mplb_amd import cases, hip
from mplb_amd.enums import BoundaryScheme as BSch
from mplb_amd.enums import BoundarySurface as BS
from mplb_amd.enums import VertexType as VT
from oracle.oracle import OracleCase

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare(g, o, what):
    for b in g.sizes:
        d = float(np.max(np.abs(g.get_f(b) - o.get_f(b))))
        assert d == 0.0, (what, b, d)
    g.update_macros()
    o.update_macros()
    for b in g.sizes:
        for v in range(g.dim + 1):
            assert np.array_equal(g.get_macro(b, v), o.get_macro(b, v)), (what, b, v)
    assert abs(g.mass() - o.mass()) <= 1e-13 * abs(o.mass()), what


def temperature_field(shape, amp=0.25):
    """A smooth positive Variable_T field: 1 + amp sin(...) products."""
    axes = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    t = np.ones(shape)
    for k, a in enumerate(axes):
        t *= 1 + amp * np.sin(2 * np.pi * (a + 0.5) / shape[k] + k)
    return t


# ---- D3Q15 ---------------------------------------------------------------------
@pytest.mark.parametrize("order", [("Left", "Right", "Top", "Bottom", "Front", "Back"),
                                   ("Top", "Bottom", "Left", "Right", "Front", "Back")])
def test_d3q15_cavity_bitexact(order, step_form):
    log("D3Q15 cavity 11^3 F64 order %s form %s" % (order[0], step_form))
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0, lattice=15)
    cases.cavity3d(g, 11, re=100.0, ulid=0.05, noise=True, order=order)
    o = OracleCase(3, lattice=15)
    cases.cavity3d(o, 11, re=100.0, ulid=0.05, noise=True, order=order)
    g.step(30)
    o.step(30)
    compare(g, o, "d3q15 cavity")


def test_d3q15_periodic_forced_box_bitexact(step_form):
    log("D3Q15 periodic forced box F64 form %s" % step_form)
    rng = np.random.default_rng(3)
    size = (9, 7, 8)
    shp = tuple(reversed(size))
    rho = 1 + 0.01 * rng.standard_normal(shp)
    u = [0.02 * rng.standard_normal(shp) for _ in range(3)]
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0, lattice=15)
    cases.periodic_box(g, size, 0.8, rho, *u, force=(2e-4, -1e-4, 5e-5))
    o = OracleCase(3, lattice=15)
    cases.periodic_box(o, size, 0.8, rho, *u, force=(2e-4, -1e-4, 5e-5))
    g.step(40)
    o.step(40)
    compare(g, o, "d3q15 box")


def test_d3q15_lchannel_bitexact(step_form):
    """Three D3Q15 blocks joined by VirtualBoundary faces and the rotated
    periodic link, walls on the rest, body force: the halo copies carry
    D3Q15's inward populations."""
    log("D3Q15 closed L-channel a=9 form %s" % step_form)
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0, lattice=15)
    cases.lchannel_periodic(g, a=9, length=14, g=1e-4)
    o = OracleCase(3, lattice=15)
    cases.lchannel_periodic(o, a=9, length=14, g=1e-4)
    g.step(60)
    o.step(60)
    compare(g, o, "d3q15 l-channel")


def test_d3q15_iterate_residual_matches_oracle():
    log("D3Q15 iterate to 1e-3")
    period, crit = 25, 1e-3
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0, lattice=15)
    cases.cavity3d(g, 10, re=50.0, noise=False)
    steps, rg = g.iterate(crit, period, 20000)
    o = OracleCase(3, lattice=15)
    cases.cavity3d(o, 10, re=50.0, noise=False)
    it, ro = 0, 1.0
    while ro >= crit:
        o.step(period)
        it += period
        o.update_macros()
        ro, _ = o.residual(period)
    assert steps == it, (steps, it)
    assert math.isclose(rg, ro, rel_tol=1e-9), (rg, ro)


@pytest.mark.parametrize("storage,tol_u", [(hip.STORE_F32, 2e-6), (hip.STORE_F16, 2e-4)])
def test_d3q15_reduced_storage_vs_oracle(storage, tol_u, step_form):
    """FP32 f - w and FP16 (f - w)/s_q on D3Q15 (scales w_q / 2^10 of its
    own weights) against the FP64 oracle after 60 steps of a 12^3 cavity, in
    every step form: max-abs velocity error below tol_u (lid speed 0.087)."""
    log("D3Q15 cavity storage %d form %s" % (storage, step_form))
    g = hip.LatticeCase(3, storage=storage, device=0, lattice=15)
    cases.cavity3d(g, 12, re=100.0, ulid=0.05, noise=True)
    o = OracleCase(3, lattice=15)
    cases.cavity3d(o, 12, re=100.0, ulid=0.05, noise=True)
    g.step(60)
    o.step(60)
    g.update_macros()
    o.update_macros()
    du = max(float(np.max(np.abs(g.get_macro(0, v) - o.get_macro(0, v)))) for v in (1, 2, 3))
    assert du < tol_u, du
    assert abs(g.mass() - o.mass()) < 1e-5 * o.mass()


# ---- Collision_BGKThermal4th ----------------------------------------------------
def _thermal_cavity(be, dim, n):
    be.set_collision_type(1)
    if dim == 2:
        cases.cavity2d(be, n, re=100.0, ulid=0.05)
    else:
        cases.cavity3d(be, n, re=100.0, ulid=0.05, noise=True)
    shp = (n, n) if dim == 2 else (n, n, n)
    be.set_temperature(0, temperature_field(shp))


@pytest.mark.parametrize("dim,q", [(2, 9), (3, 19), (3, 15)])
def test_thermal_cavity_bitexact(dim, q, step_form):
    """A cavity with a fixed, nonuniform temperature field: per-node
    relaxation tau = tauRef / (rho sqrt(T)) and the 4th-order equilibrium,
    bit for bit in every step form (whole-lattice launch, bulk plus boundary
    kernels at each fold level)."""
    n = 12 if dim == 2 else 10
    log("thermal cavity %dD Q%d form %s" % (dim, q, step_form))
    g = hip.LatticeCase(dim, storage=hip.STORE_F64, device=0, lattice=q)
    _thermal_cavity(g, dim, n)
    o = OracleCase(dim, lattice=q)
    _thermal_cavity(o, dim, n)
    g.step(25)
    o.step(25)
    compare(g, o, "thermal cavity")


@pytest.mark.parametrize("q", [19, 15])
def test_thermal_periodic_forced_box_bitexact(q):
    """Thermal collision with BodyForce_1st on a periodic box: the force
    term tau w F uses the per-node tau."""
    log("thermal forced box Q%d" % q)
    rng = np.random.default_rng(11)
    size = (8, 7, 6)
    shp = tuple(reversed(size))
    rho = 1 + 0.01 * rng.standard_normal(shp)
    u = [0.02 * rng.standard_normal(shp) for _ in range(3)]
    T = temperature_field(shp, amp=0.15)
    backends = []
    for be in (hip.LatticeCase(3, storage=hip.STORE_F64, device=0, lattice=q),
               OracleCase(3, lattice=q)):
        be.set_collision_type(1)
        cases.periodic_box(be, size, 0.8, rho, *u, force=(1e-4, 2e-4, -1e-4))
        be.set_temperature(0, T)
        be.step(30)
        backends.append(be)
    compare(backends[0], backends[1], "thermal box")


def test_thermal_and_lattice_errors():
    """What the library refuses: the thermal collision with reduced storage,
    a step before the temperature is given, an undefined lattice, D3Q15 in
    the swap scheme."""
    r = hip.LatticeCase(3, storage=hip.STORE_F32, device=0)
    with pytest.raises(hip.MplbError, match="FP64"):
        r.set_collision_type(1)
    g = hip.LatticeCase(3, storage=hip.STORE_F64, device=0)
    g.set_collision_type(1)
    cases.cavity3d(g, 8, noise=False)
    with pytest.raises(hip.MplbError, match="set_temperature"):
        g.step(1)
    with pytest.raises(hip.MplbError, match="d3q15"):
        hip.LatticeCase(3, storage=hip.STORE_F64, device=0, lattice=27)
    s = hip.LatticeCase(3, storage=hip.STORE_F64, device=0, lattice=15)
    s.set_scheme(11)
    s.add_block(0, (6, 6, 6))
    s.set_collision(0.1)
    with pytest.raises(hip.MplbError, match="D3Q19"):
        s.prepare()
