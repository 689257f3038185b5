"""Case definitions that drive a lattice backend through the HiLeMMS call
sequence of the reference Apps (DefineBlocks ... Partition,
SetInitialMacrosVars, PreDefinedInitialCondition, SetTimeStep).

A backend is any object with the method set of mplb_amd.LatticeCase
(add_block, add_connection, set_collision, add_boundary, prepare,
set_macros, init_feq, set_dt, set_accel, step, ...).  The same builder
therefore sets up the HIP path and the CPU oracle identically.

Unit mapping (as in Apps/2DCavity/lbm2d_cavity.cpp:76-77, 138 and
Src/model_kernel.inc:429): meshSize = 1/(N-1) for N nodes per side,
dt = meshSize/sqrt(3), a lattice relaxation time tau_L maps to
TauRef = (tau_L - 1/2) dt, and a lattice speed U maps to the code
velocity sqrt(3) U (CalcBGKFeq takes cu = CS xi . u with CS = sqrt(3)).
"""
import math

import numpy as np

from .enums import BoundaryScheme as BSch
from .enums import BoundarySurface as BS
from .enums import VertexType as VT

CS = math.sqrt(3.0)
NOISE_SEED = 0x4D504C42
SCHEME_SWAP = 11  # SchemeType Scheme_StreamCollision_Swap (Src/scheme.h:71)
_MASK = (1 << 64) - 1


def splitmix64(x):
    """Counter-based splitmix64 (Steele, Lea & Flood 2014) on uint64 arrays."""
    z = (np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def density_noise(shape, seed=NOISE_SEED, amplitude=1e-4):
    """rho0 = 1 + amplitude*(2u-1), u = splitmix64(seed ^ linear_index) as a
    53-bit fraction; linear index i + nx*(j + ny*k), i fastest."""
    n = int(np.prod(shape))
    with np.errstate(over="ignore"):
        idx = np.arange(n, dtype=np.uint64) ^ np.uint64(seed)
        h = splitmix64(idx)
    u = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (1.0 + amplitude * (2.0 * u - 1.0)).reshape(shape)


def tau_ref_for(tau_lattice, dt):
    return (tau_lattice - 0.5) * dt


def cavity_params(n, re, ulid):
    mesh = 1.0 / (n - 1)
    dt = mesh / CS
    tau_l = 0.5 + 3.0 * ulid * n / re
    return mesh, dt, tau_l


def cavity2d(be, n, re=100.0, ulid=0.05,
             order=("Left", "Right", "Top", "Bottom"), rho0=None):
    """Apps/2DCavity/lbm2d_cavity.cpp:72-135: one D2Q9 block, EQMDiffuseRefl
    walls (stationary) and a moving lid on Top, defined in `order`."""
    mesh, dt, tau_l = cavity_params(n, re, ulid)
    be.add_block(0, (n, n))
    be.set_collision(tau_ref_for(tau_l, dt), force_type=0)
    lid = CS * ulid
    for s in order:
        vel = (lid, 0.0) if s == "Top" else (0.0, 0.0)
        be.add_boundary(0, BS[s], BSch.EQMDiffuseRefl, vel, VT.Wall)
    be.prepare()
    shp = (n, n)
    be.set_macros(0, rho0 if rho0 is not None else np.ones(shp),
                  np.zeros(shp), np.zeros(shp))
    be.init_feq()
    be.set_dt(dt)
    return dict(dt=dt, tau_l=tau_l, mesh=mesh, lid=lid)


def cavity3d(be, n, re=400.0, ulid=0.05, noise=True,
             order=("Left", "Right", "Top", "Bottom", "Front", "Back"),
             size=None):
    """Apps/3DCavity/lbm3d_cavity.cpp:72-180 with the hard-coded BC order
    (Left, Right, Top, Bottom, Front, Back); the lid moves along x on Top.
    With noise, rho0 = 1 +- 1e-4 from splitmix64 (seeded, counter based)."""
    mesh, dt, tau_l = cavity_params(n, re, ulid)
    size = tuple(size) if size is not None else (n, n, n)
    be.add_block(0, size)
    be.set_collision(tau_ref_for(tau_l, dt), force_type=0)
    lid = CS * ulid
    for s in order:
        vel = (lid, 0.0, 0.0) if s == "Top" else (0.0, 0.0, 0.0)
        be.add_boundary(0, BS[s], BSch.EQMDiffuseRefl, vel, VT.Wall)
    be.prepare()
    shp = tuple(reversed(size))
    rho = density_noise(shp) if noise else np.ones(shp)
    z = np.zeros(shp)
    be.set_macros(0, rho, z, z, z)
    be.init_feq()
    be.set_dt(dt)
    return dict(dt=dt, tau_l=tau_l, mesh=mesh, lid=lid)


def periodic_box(be, size, tau_l, rho, u, v, w=None, force=None):
    """A block whose every face is a periodic self-link (Field::CreateHalos,
    field.h:207-214: "the periodic boundary can be treated by setting the
    neighbor to the block itself") with MDPeriodic nodes on the envelope."""
    dim = len(size)
    be.add_block(0, size)
    pairs = [(BS.Left, BS.Right), (BS.Right, BS.Left), (BS.Top, BS.Bottom),
             (BS.Bottom, BS.Top)]
    if dim == 3:
        pairs += [(BS.Front, BS.Back), (BS.Back, BS.Front)]
    for a, b in pairs:
        be.add_connection(0, a, 0, b, VT.MDPeriodic)
    dt = 1.0
    be.set_collision(tau_ref_for(tau_l, dt), force_type=1 if force else 0,
                     force_macros=bool(force))
    for a, _ in pairs:
        be.add_boundary(0, a, BSch.None_, (), VT.MDPeriodic)
    be.prepare()
    if force:
        be.set_accel(0, force)
    be.set_macros(0, rho, u, v, w)
    be.init_feq()
    be.set_dt(dt)
    return dict(dt=dt)


def channel2d(be, nx, ny, tau_l, g, order=("Top", "Bottom")):
    """Body-force driven channel: periodic in x, EQMDiffuseRefl walls at
    Top and Bottom, BodyForce_1st with the *_Force velocity macros."""
    be.add_block(0, (nx, ny))
    be.add_connection(0, BS.Left, 0, BS.Right, VT.MDPeriodic)
    be.add_connection(0, BS.Right, 0, BS.Left, VT.MDPeriodic)
    dt = 1.0
    be.set_collision(tau_ref_for(tau_l, dt), force_type=1, force_macros=True)
    be.add_boundary(0, BS.Left, BSch.None_, (), VT.MDPeriodic)
    be.add_boundary(0, BS.Right, BSch.None_, (), VT.MDPeriodic)
    for s in order:
        be.add_boundary(0, BS[s], BSch.EQMDiffuseRefl, (0.0, 0.0), VT.Wall)
    be.prepare()
    be.set_accel(0, (g, 0.0))
    shp = (ny, nx)
    be.set_macros(0, np.ones(shp), np.zeros(shp), np.zeros(shp))
    be.init_feq()
    be.set_dt(dt)
    return dict(dt=dt)


def lchannel(be, n=33, top_lid=-0.001, outlet_rho=1.0, tau_ref=0.01,
             mesh=None):
    """Apps/3DLChannel/lbm3d_L.cpp:72-190: three D3Q19 blocks (Top, Middle,
    Right) joined by VirtualBoundary faces, a moving wall on Top of block 0,
    an ExtrapolPressure1ST outlet on the Right of block 2 and
    EQMDiffuseRefl walls elsewhere, in the App's definition order."""
    mesh = mesh if mesh is not None else 1.0 / (n - 1)
    m = n - 1
    be.add_block(0, (n, m, n))
    be.add_block(1, (n, n, n))
    be.add_block(2, (m, n, n))
    for frm, fs, to, ts in ((0, BS.Bottom, 1, BS.Top), (1, BS.Top, 0, BS.Bottom),
                            (1, BS.Right, 2, BS.Left), (2, BS.Left, 1, BS.Right)):
        be.add_connection(frm, fs, to, ts, VT.VirtualBoundary)
    be.set_collision(tau_ref, force_type=0)
    wall = (0.0, 0.0, 0.0)
    E = BSch.EQMDiffuseRefl
    VB = (BSch.None_, (), VT.VirtualBoundary)
    be.add_boundary(0, BS.Left, E, wall, VT.Wall)
    be.add_boundary(0, BS.Right, E, wall, VT.Wall)
    be.add_boundary(0, BS.Top, E, (0.0, top_lid, 0.0), VT.Wall)
    be.add_boundary(0, BS.Bottom, *VB)
    be.add_boundary(0, BS.Front, E, wall, VT.Wall)
    be.add_boundary(0, BS.Back, E, wall, VT.Wall)
    be.add_boundary(1, BS.Left, E, wall, VT.Wall)
    be.add_boundary(1, BS.Right, *VB)
    be.add_boundary(1, BS.Top, *VB)
    be.add_boundary(1, BS.RightTop, E, wall, VT.Wall)
    be.add_boundary(1, BS.Bottom, E, wall, VT.Wall)
    be.add_boundary(1, BS.Front, E, wall, VT.Wall)
    be.add_boundary(1, BS.Back, E, wall, VT.Wall)
    be.add_boundary(2, BS.Left, *VB)
    be.add_boundary(2, BS.Right, BSch.ExtrapolPressure1ST, (outlet_rho,),
                    VT.OutLet)
    be.add_boundary(2, BS.Top, E, wall, VT.Wall)
    be.add_boundary(2, BS.Bottom, E, wall, VT.Wall)
    be.add_boundary(2, BS.Front, E, wall, VT.Wall)
    be.add_boundary(2, BS.Back, E, wall, VT.Wall)
    be.prepare()
    for b, size in ((0, (n, m, n)), (1, (n, n, n)), (2, (m, n, n))):
        shp = tuple(reversed(size))
        z = np.zeros(shp)
        be.set_macros(b, np.ones(shp), z, z, z)
    be.init_feq()
    dt = mesh / CS
    be.set_dt(dt)
    return dict(dt=dt)


def lchannel_periodic(be, a=128, length=256, tau_l=0.8, g=1e-6):
    """The L-channel of Apps/3DLChannel/lbm3d_L.cpp:72-190 (blocks Top,
    Middle, Right joined by VirtualBoundary faces, EQMDiffuseRefl walls) run
    as a closed loop: the outlet face (Right of block 2) and the inlet face
    (Top of block 0) are joined by a rotated periodic connection, and
    BodyForce_1st drives the flow with |g| along each arm's axis (-y in the
    Top arm, +x in the Right arm, the bisector in the Middle block).

    Blocks: 0 = a x length x a, 1 = a x a x a, 2 = length x a x a.  Lattice
    units: dt = 1, TauRef = tau_l - 1/2.  The periodic faces are defined
    before each block's walls so that the walls own the edges."""
    be.add_block(0, (a, length, a))
    be.add_block(1, (a, a, a))
    be.add_block(2, (length, a, a))
    for frm, fs, to, ts in ((0, BS.Bottom, 1, BS.Top), (1, BS.Top, 0, BS.Bottom),
                            (1, BS.Right, 2, BS.Left), (2, BS.Left, 1, BS.Right)):
        be.add_connection(frm, fs, to, ts, VT.VirtualBoundary)
    be.add_rotated_connection(2, BS.Right, 0, BS.Top, VT.MDPeriodic)
    be.add_rotated_connection(0, BS.Top, 2, BS.Right, VT.MDPeriodic)
    dt = 1.0
    be.set_collision(tau_ref_for(tau_l, dt), force_type=1, force_macros=True)
    wall = (0.0, 0.0, 0.0)
    E = BSch.EQMDiffuseRefl
    VB = (BSch.None_, (), VT.VirtualBoundary)
    MP = (BSch.None_, (), VT.MDPeriodic)
    be.add_boundary(0, BS.Top, *MP)
    be.add_boundary(0, BS.Left, E, wall, VT.Wall)
    be.add_boundary(0, BS.Right, E, wall, VT.Wall)
    be.add_boundary(0, BS.Bottom, *VB)
    be.add_boundary(0, BS.Front, E, wall, VT.Wall)
    be.add_boundary(0, BS.Back, E, wall, VT.Wall)
    be.add_boundary(1, BS.Left, E, wall, VT.Wall)
    be.add_boundary(1, BS.Right, *VB)
    be.add_boundary(1, BS.Top, *VB)
    be.add_boundary(1, BS.RightTop, E, wall, VT.Wall)
    be.add_boundary(1, BS.Bottom, E, wall, VT.Wall)
    be.add_boundary(1, BS.Front, E, wall, VT.Wall)
    be.add_boundary(1, BS.Back, E, wall, VT.Wall)
    be.add_boundary(2, BS.Right, *MP)
    be.add_boundary(2, BS.Left, *VB)
    be.add_boundary(2, BS.Top, E, wall, VT.Wall)
    be.add_boundary(2, BS.Bottom, E, wall, VT.Wall)
    be.add_boundary(2, BS.Front, E, wall, VT.Wall)
    be.add_boundary(2, BS.Back, E, wall, VT.Wall)
    be.prepare()
    s2 = g / math.sqrt(2.0)
    be.set_accel(0, (0.0, -g, 0.0))
    be.set_accel(1, (s2, -s2, 0.0))
    be.set_accel(2, (g, 0.0, 0.0))
    sizes = ((a, length, a), (a, a, a), (length, a, a))
    for b, size in enumerate(sizes):
        shp = tuple(reversed(size))
        z = np.zeros(shp)
        be.set_macros(b, np.ones(shp), z, z, z)
    be.init_feq()
    be.set_dt(dt)
    return dict(dt=dt, nodes=sum(int(np.prod(x)) for x in sizes), sizes=sizes)


# ---- HiLeMMS JSON case files (keys of the reference's configuration.cpp) ----
_SCHEME_STRINGS = {"EQMDiffuseREfl": BSch.EQMDiffuseRefl,
                   "ExtrapolPressure1ST": BSch.ExtrapolPressure1ST,
                   "MDPeriodic": BSch.MDPeriodic, "FDPeriodic": BSch.FDPeriodic,
                   "BounceBack": BSch.BounceBack, "None": BSch.None_}


def from_config(be, cfg):
    """Drive a backend through the call sequence of an App's
    simulate(config) for a parsed HiLeMMS JSON case; returns
    (steps_to_run, check_period, dt)."""
    dim = int(cfg["SpaceDim"])
    sizes = cfg["BlockSize"]
    ids = cfg["BlockIds"]
    for k, bid in enumerate(ids):
        be.add_block(bid, tuple(sizes[dim * k: dim * (k + 1)]))
    for frm, fs, to, ts, ty in zip(cfg.get("FromBlockIds") or [],
                                   cfg.get("FromBoundarySurface") or [],
                                   cfg.get("ToBlockIds") or [],
                                   cfg.get("ToBoundarySurface") or [],
                                   cfg.get("BlockConnectionType") or []):
        be.add_connection(frm, BS[fs], to, BS[ts], VT[ty])
    if cfg.get("SchemeType") == "Scheme_StreamCollision_Swap":
        be.set_scheme(SCHEME_SWAP)
    if cfg["CollisionType"][0] == "Collision_BGKThermal4th":
        be.set_collision_type(1)
    force = cfg["BodyForceType"][0]
    force_macros = any(t.endswith("_Force") for t in cfg["MacroVarTypes"])
    be.set_collision(cfg["TauRef"][0],
                     force_type=1 if force in ("BodyForce_1st", "BodyForce_1st_Swap") else 0,
                     force_macros=force_macros)
    k = 0
    while ("BoundaryCondition%d" % k) in cfg and cfg["BoundaryCondition%d" % k]:
        bc = cfg["BoundaryCondition%d" % k]
        be.add_boundary(bc["BlockIndex"], BS[bc["BoundarySurface"]],
                        _SCHEME_STRINGS[bc["BoundaryScheme"]], bc["GivenVars"],
                        VT[bc["BoundaryType"]])
        k += 1
    be.prepare()
    for k, bid in enumerate(ids):
        shp = tuple(reversed(sizes[dim * k: dim * (k + 1)]))
        z = np.zeros(shp)
        be.set_macros(bid, np.ones(shp), z, z, z if dim == 3 else None)
    be.init_feq()
    dt = cfg["MeshSize"] / CS
    be.set_dt(dt)
    steps = cfg.get("TimeStepsToRun", 0) if cfg.get("Transient", True) else None
    return steps, cfg["CheckPeriod"], dt


# ---- batched independent cavities -------------------------------------------
BATCH_SIZES = (64, 96, 128, 192, 256)


def batch_spec(count=256, seed=NOISE_SEED, sizes=BATCH_SIZES, re_range=(100.0, 1000.0)):
    """Sizes and Reynolds numbers of a batch, from the counter-based
    splitmix64 stream: instance i draws N from `sizes` with
    splitmix64(seed ^ 2i) and Re uniform in re_range with
    splitmix64(seed ^ (2i+1))."""
    i = np.arange(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h1 = splitmix64((2 * i) ^ np.uint64(seed))
        h2 = splitmix64((2 * i + np.uint64(1)) ^ np.uint64(seed))
    n = np.array(sizes)[(h1 % np.uint64(len(sizes))).astype(np.int64)]
    u = (h2 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    re = re_range[0] + (re_range[1] - re_range[0]) * u
    return [int(x) for x in n], [float(x) for x in re]


def rect_mask(n, rects):
    """Obstacle mask [y][x] of an n x n cavity from rectangles (x0, y0, w, h)."""
    m = np.zeros((n, n), np.uint8)
    for x0, y0, w, h in rects:
        m[y0:y0 + h, x0:x0 + w] = 1
    return m


def obstacle_rects(count, n=128, seed=NOISE_SEED, max_rects=3, side=(8, 32)):
    """Per case: 0..max_rects axis-aligned rectangles with sides in
    [side[0], side[1]] nodes, kept two nodes clear of the cavity walls (the
    bounce-back neighbours must be bulk Fluid nodes).  Case i draws from
    numpy's PCG64 seeded with (seed, i)."""
    out = []
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        k = int(rng.integers(0, max_rects + 1))
        rects = []
        for _ in range(k):
            w, h = (int(v) for v in rng.integers(side[0], side[1] + 1, size=2))
            x0 = int(rng.integers(2, n - 2 - w + 1))
            y0 = int(rng.integers(2, n - 2 - h + 1))
            rects.append((x0, y0, w, h))
        out.append(rects)
    return out


def cavity2d_obstacles(be, n, rects, re=100.0, ulid=0.05):
    """The 2D cavity (Apps/2DCavity) with in-domain obstacles (half-way
    bounce-back, an extension): Left, Right, Top (lid), Bottom walls."""
    mesh, dt, tau_l = cavity_params(n, re, ulid)
    be.add_block(0, (n, n))
    be.set_collision(tau_ref_for(tau_l, dt), force_type=0)
    lid = CS * ulid
    for s_ in ("Left", "Right", "Top", "Bottom"):
        vel = (lid, 0.0) if s_ == "Top" else (0.0, 0.0)
        be.add_boundary(0, BS[s_], BSch.EQMDiffuseRefl, vel, VT.Wall)
    solid = rect_mask(n, rects)
    be.set_obstacles(0, solid)
    be.prepare()
    shp = (n, n)
    be.set_macros(0, np.ones(shp), np.zeros(shp), np.zeros(shp))
    be.init_feq()
    be.set_dt(dt)
    return dict(dt=dt, tau_l=tau_l, mesh=mesh, lid=lid, solid=solid)


def batched_cavities(be, sizes, res, ulid=0.05, masks=None):
    """One D2Q9 cavity per block (the 2D cavity App's boundary order), each
    with its own relaxation time and time step; blocks are unconnected."""
    for b, n in enumerate(sizes):
        be.add_block(b, (n, n))
    be.set_collision(1.0, force_type=0)
    lid = CS * ulid
    params = []
    for b, (n, re) in enumerate(zip(sizes, res)):
        mesh, dt, tau_l = cavity_params(n, re, ulid)
        be.set_block_collision(b, tau_ref_for(tau_l, dt), dt)
        for s in ("Left", "Right", "Top", "Bottom"):
            vel = (lid, 0.0) if s == "Top" else (0.0, 0.0)
            be.add_boundary(b, BS[s], BSch.EQMDiffuseRefl, vel, VT.Wall)
        if masks is not None and masks[b] is not None and masks[b].any():
            be.set_obstacles(b, masks[b])
        params.append(dict(n=n, re=re, dt=dt, tau_l=tau_l))
    be.prepare()
    for b, n in enumerate(sizes):
        shp = (n, n)
        be.set_macros(b, np.ones(shp), np.zeros(shp), np.zeros(shp))
    be.init_feq()
    return params


def _channel_blocks(be, first, sizes, ny, tau_l, g, per_block_collision):
    """A force-driven channel periodic in x, made of len(sizes) blocks of
    sizes[k] x ny nodes joined end to end (VirtualBoundary junctions, the
    last block linked back to the first through MDPeriodic faces), walls
    Top and Bottom; block ids first .. first + len(sizes) - 1."""
    nb = len(sizes)
    ids = [first + k for k in range(nb)]
    for k, nx in enumerate(sizes):
        be.add_block(ids[k], (nx, ny))
    for k in range(nb):
        a, b = ids[k], ids[(k + 1) % nb]
        vt = VT.MDPeriodic if k == nb - 1 else VT.VirtualBoundary
        be.add_connection(a, BS.Right, b, BS.Left, vt)
        be.add_connection(b, BS.Left, a, BS.Right, vt)
    for k in range(nb):
        right = VT.MDPeriodic if k == nb - 1 else VT.VirtualBoundary
        left = VT.MDPeriodic if k == 0 else VT.VirtualBoundary
        be.add_boundary(ids[k], BS.Left, BSch.None_, (), left)
        be.add_boundary(ids[k], BS.Right, BSch.None_, (), right)
        for s in ("Top", "Bottom"):
            be.add_boundary(ids[k], BS[s], BSch.EQMDiffuseRefl, (0.0, 0.0), VT.Wall)
        if per_block_collision:
            be.set_block_collision(ids[k], tau_ref_for(tau_l, 1.0), 1.0)
    return ids


def channel_multi(be, sizes, ny, tau_l, g):
    """One force-driven periodic channel of one or more blocks (as one
    instance of batched_mixed builds it), on its own: dt = 1,
    BodyForce_1st with the *_Force velocity macros."""
    ids = _channel_blocks(be, 0, sizes, ny, tau_l, g, False)
    be.set_collision(tau_ref_for(tau_l, 1.0), force_type=1, force_macros=True)
    be.prepare()
    for b in ids:
        be.set_accel(b, (g, 0.0))
    for b, nx in zip(ids, sizes):
        shp = (ny, nx)
        be.set_macros(b, np.ones(shp), np.zeros(shp), np.zeros(shp))
    be.init_feq()
    be.set_dt(1.0)
    return dict(dt=1.0, blocks=ids)


def batched_mixed(be, instances, ulid=0.05):
    """Independent instances of different kinds in one context, each with
    its own collision constants: ("cavity", n, re) is the 2D cavity of
    cavity2d; ("channel", [nx, ...], ny, tau_l, g) a force-driven periodic
    channel of one or more blocks (channel_multi).  The context runs
    BodyForce_1st with the *_Force macros throughout; the cavities' blocks
    have zero acceleration, so their arithmetic is that of a cavity without
    force.  Returns, per instance, its block ids and dt."""
    out = []
    nb = 0
    lid = CS * ulid
    accel = []
    for inst in instances:
        if inst[0] == "cavity":
            _, n, re = inst
            mesh, dt, tau_l = cavity_params(n, re, ulid)
            be.add_block(nb, (n, n))
            be.set_block_collision(nb, tau_ref_for(tau_l, dt), dt)
            for s in ("Left", "Right", "Top", "Bottom"):
                vel = (lid, 0.0) if s == "Top" else (0.0, 0.0)
                be.add_boundary(nb, BS[s], BSch.EQMDiffuseRefl, vel, VT.Wall)
            out.append(dict(kind="cavity", blocks=[nb], dt=dt, sizes=[(n, n)]))
            nb += 1
        else:
            _, sizes, ny, tau_l, g = inst
            ids = _channel_blocks(be, nb, sizes, ny, tau_l, g, True)
            accel += [(b, g) for b in ids]
            out.append(dict(kind="channel", blocks=ids, dt=1.0,
                            sizes=[(nx, ny) for nx in sizes]))
            nb += len(ids)
    be.set_collision(1.0, force_type=1, force_macros=True)
    be.prepare()
    for b, g in accel:
        be.set_accel(b, (g, 0.0))
    for inst in out:
        for b, (nx, ny) in zip(inst["blocks"], inst["sizes"]):
            shp = (ny, nx)
            be.set_macros(b, np.ones(shp), np.zeros(shp), np.zeros(shp))
    be.init_feq()
    return out


def cavity3d_swap(be, n, re=100.0, ulid=0.05, noise=True, size=None,
                  order=("Left", "Right", "Top", "Bottom", "Front", "Back")):
    """Apps/3DCavity-Swap/lbm3d_cavity_swap.cpp:72-110: the 3D cavity on the
    swap scheme (Scheme_StreamCollision_Swap, Collision_BGKIsothermal2nd_Swap,
    BodyForce_None_Swap), EQMDiffuseRefl walls, the lid along x on Top."""
    mesh, dt, tau_l = cavity_params(n, re, ulid)
    size = tuple(size) if size is not None else (n, n, n)
    be.add_block(0, size)
    be.set_scheme(SCHEME_SWAP)
    be.set_collision(tau_ref_for(tau_l, dt), force_type=0)
    lid = CS * ulid
    for s in order:
        vel = (lid, 0.0, 0.0) if s == "Top" else (0.0, 0.0, 0.0)
        be.add_boundary(0, BS[s], BSch.EQMDiffuseRefl, vel, VT.Wall)
    be.prepare()
    shp = tuple(reversed(size))
    rho = density_noise(shp, amplitude=1e-3) if noise else np.ones(shp)
    z = np.zeros(shp)
    be.set_macros(0, rho, z, z, z)
    be.init_feq()
    be.set_dt(dt)
    return dict(dt=dt, tau_l=tau_l, lid=lid)


def swap_box(be, size, tau_l, rho, u, v, w, force=None, fd_axis=None):
    """A 3D box on the swap scheme, self-linked on every face (MDPeriodic)
    or, along fd_axis, joined to itself by FDPeriodic connections with
    FDPeriodic boundaries on those two faces and stationary EQMDiffuseRefl
    walls on the others."""
    be.add_block(0, size)
    be.set_scheme(SCHEME_SWAP)
    faces = [(BS.Left, BS.Right), (BS.Bottom, BS.Top), (BS.Back, BS.Front)]
    dt = 1.0
    be.set_collision(tau_ref_for(tau_l, dt), force_type=1 if force else 0,
                     force_macros=bool(force))
    for ax, (lo, hi) in enumerate(faces):
        if fd_axis is None:
            be.add_connection(0, lo, 0, hi, VT.MDPeriodic)
            be.add_connection(0, hi, 0, lo, VT.MDPeriodic)
        elif ax == fd_axis:
            be.add_connection(0, lo, 0, hi, VT.FDPeriodic)
            be.add_connection(0, hi, 0, lo, VT.FDPeriodic)
    for ax, (lo, hi) in enumerate(faces):
        for s in (lo, hi):
            if fd_axis is None:
                be.add_boundary(0, s, BSch.None_, (), VT.MDPeriodic)
            elif ax == fd_axis:
                be.add_boundary(0, s, BSch.FDPeriodic, (), VT.FDPeriodic)
            else:
                be.add_boundary(0, s, BSch.EQMDiffuseRefl, (0.0, 0.0, 0.0), VT.Wall)
    be.prepare()
    if force:
        be.set_accel(0, force)
    be.set_macros(0, rho, u, v, w)
    be.init_feq()
    be.set_dt(dt)
    return dict(dt=dt)
