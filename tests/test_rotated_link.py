"""Rotated block connections (joined perpendicular faces; the reference
links same-axis faces only, field.h:207-214).

The oracle's rotated link is pinned by rotation equivariance: a two-block
case joined Right->Left, and the same case with the second block stored
turned by a quarter rotation R and joined through perpendicular faces, must
evolve identically up to R (fields moved, populations permuted, velocities
turned).  A wrong sign, axis, offset or population map in the link breaks
this at O(1); only summation order differs, at round-off.  Every quarter
rotation that can take +x to a perpendicular axis is exercised, so both
signs of each transverse map occur.
"""
import numpy as np
import pytest

from mplb_amd.enums import BoundaryScheme as BSch
from mplb_amd.enums import BoundarySurface as BS
from mplb_amd.enums import VertexType as VT
from mplb_amd.lattice import lattice
from oracle.oracle import OracleCase

CS = 3 ** 0.5
E = BSch.EQMDiffuseRefl
FACES = {(0, -1): "Left", (0, 1): "Right", (1, -1): "Bottom", (1, 1): "Top",
         (2, -1): "Back", (2, 1): "Front"}
NORMAL = {v: k for k, v in FACES.items()}


def quarter(src_axis, dst_axis, sign):
    """Rotation taking e_src to sign*e_dst, fixing the third axis."""
    R = np.zeros((3, 3), dtype=np.int64)
    a3 = 3 - src_axis - dst_axis
    R[dst_axis, src_axis] = sign
    R[src_axis, dst_axis] = -sign
    R[a3, a3] = 1
    assert round(np.linalg.det(R)) == 1
    return R


# (dim, R): quarter turns taking +x to -y, +y (about z) and to -z, +z (about y)
ROTATIONS = [(2, quarter(0, 1, -1)), (2, quarter(0, 1, 1)),
             (3, quarter(0, 1, -1)), (3, quarter(0, 1, 1)),
             (3, quarter(0, 2, -1)), (3, quarter(0, 2, 1))]


def face_image(R, face):
    ax, s = NORMAL[face]
    v = R[:, ax] * s
    a = int(np.flatnonzero(v)[0])
    return FACES[(a, int(v[a]))]


def rot_size(R, size):
    size = list(size) + [1] * (3 - len(size))
    out = [0, 0, 0]
    for j in range(3):
        k = int(np.flatnonzero(R[j])[0])
        out[j] = size[k]
    return out


def rot_field(a, R, dim):
    """Scalar field [k][j][i] (or [j][i]) of B -> the same field on B_R:
    p'_j = p_k if R[j,k] = 1, n_k - 1 - p_k if R[j,k] = -1."""
    a3 = a if dim == 3 else a[None]
    n = a3.shape[::-1]
    nr = rot_size(R, n)
    grid = np.meshgrid(*[np.arange(m) for m in nr], indexing="ij")
    src = [None, None, None]
    for j in range(3):
        k = int(np.flatnonzero(R[j])[0])
        src[k] = grid[j] if R[j, k] == 1 else n[k] - 1 - grid[j]
    out = a3[src[2], src[1], src[0]]  # indexed [i'][j'][k']
    out = np.ascontiguousarray(np.transpose(out, (2, 1, 0)))
    return out if dim == 3 else out[0]


def build_pair(be, dim, na, nb, ny, nz, R, tau_ref, dt, lid_a, lid_b):
    """A (na x ny [x nz]) joined at its Right face to B (nb x ny [x nz]);
    walls elsewhere; moving lids on A.Top and B.Top.  With a rotation R, B is
    stored as B_R, turned by R, and joined through B_R's image of B.Left."""
    szA = (na, ny, nz)[:dim]
    szB = (nb, ny, nz)[:dim]
    rotated = R is not None
    Rm = R if rotated else np.eye(3, dtype=np.int64)
    be.add_block(0, szA)
    be.add_block(1, tuple(rot_size(Rm, szB)[:dim]))
    bface = lambda f: BS[face_image(Rm, f)]
    if rotated:
        be.add_rotated_connection(0, BS.Right, 1, bface("Left"), VT.VirtualBoundary)
        be.add_rotated_connection(1, bface("Left"), 0, BS.Right, VT.VirtualBoundary)
    else:
        be.add_connection(0, BS.Right, 1, BS.Left, VT.VirtualBoundary)
        be.add_connection(1, BS.Left, 0, BS.Right, VT.VirtualBoundary)
    be.set_collision(tau_ref, force_type=0)
    zero = (0.0,) * dim
    faces = ["Left", "Right", "Top", "Bottom"] + (["Front", "Back"] if dim == 3 else [])
    for f in faces:
        if f == "Right":
            be.add_boundary(0, BS.Right, BSch.None_, (), VT.VirtualBoundary)
        else:
            be.add_boundary(0, BS[f], E, lid_a if f == "Top" else zero, VT.Wall)
    for f in faces:
        if f == "Left":
            be.add_boundary(1, bface("Left"), BSch.None_, (), VT.VirtualBoundary)
            continue
        vel = np.array(list(lid_b if f == "Top" else zero) + [0.0] * (3 - dim))
        be.add_boundary(1, bface(f), E, tuple((Rm @ vel)[:dim]), VT.Wall)
    be.prepare()
    rng = np.random.default_rng(11)
    shpA = tuple(reversed(szA))
    shpB = tuple(reversed(szB))
    rhoA = 1 + 1e-3 * rng.standard_normal(shpA)
    rhoB = 1 + 1e-3 * rng.standard_normal(shpB)
    uB = [1e-2 * rng.standard_normal(shpB) for _ in range(dim)]
    be.set_macros(0, rhoA, *[np.zeros(shpA) for _ in range(dim)])
    if rotated:
        uB3 = uB + [np.zeros(shpB)] * (3 - dim)
        uR = [sum(Rm[j, k] * uB3[k] for k in range(3)) for j in range(dim)]
        be.set_macros(1, rot_field(rhoB, Rm, dim), *[rot_field(u, Rm, dim) for u in uR])
    else:
        be.set_macros(1, rhoB, *uB)
    be.init_feq()
    be.set_dt(dt)


def qmap(R, dim):
    xi = lattice(dim)[0]
    m = np.zeros(len(xi), dtype=np.int64)
    for q in range(len(xi)):
        v = (R @ np.array(list(xi[q]) + [0] * (3 - dim)))[:dim]
        m[q] = [r for r in range(len(xi)) if np.array_equal(xi[r], v)][0]
    return m


def pair_cases(dim, R, backend=OracleCase, **kw):
    na, nb, ny, nz = 7, 6, 5, 5
    dt = 0.1
    lid_a = (0.05 * CS,) + (0.0,) * (dim - 1)
    lid_b = (0.0, 0.0, 0.03 * CS) if dim == 3 else (0.03 * CS, 0.0)
    out = []
    for rr in (None, R):
        be = backend(dim, **kw)
        build_pair(be, dim, na, nb, ny, nz, rr, 0.2 * dt, dt, lid_a, lid_b)
        out.append(be)
    return out


def equivariance_error(u, r, dim, R):
    err = float(np.max(np.abs(u.get_f(0) - r.get_f(0))))
    fb_u, fb_r = u.get_f(1), r.get_f(1)
    m = qmap(R, dim)
    for q in range(fb_u.shape[0]):
        err = max(err, float(np.max(np.abs(rot_field(fb_u[q], R, dim) - fb_r[m[q]]))))
    return err


@pytest.mark.parametrize("dim,R", ROTATIONS)
def test_rotated_link_is_rotation_equivariant(dim, R):
    u, r = pair_cases(dim, R)
    u.step(40)
    r.step(40)
    err = equivariance_error(u, r, dim, R)
    assert err < 1e-13, err
    fb = u.get_f(1)
    assert np.ptp(fb[1]) > 1e-4  # the flow crossed the link


def test_rotated_connection_rejects_parallel_faces():
    o = OracleCase(3)
    o.add_block(0, (4, 4, 4))
    o.add_block(1, (4, 4, 4))
    with pytest.raises(Exception):
        o.add_rotated_connection(0, BS.Right, 1, BS.Left, VT.VirtualBoundary)


def test_lchannel_periodic_plan_only():
    """The closed L-channel (rotated periodic inlet/outlet) plans without a
    device; a rotated link with mismatched transverse sizes is refused."""
    from mplb_amd import cases, hip
    g = hip.LatticeCase(3, storage=hip.STORE_F32, device=-1)
    with pytest.raises(hip.MplbError):
        cases.lchannel_periodic(g, a=8, length=12)  # set_accel needs a device
    bad = hip.LatticeCase(3, storage=hip.STORE_F32, device=-1)
    bad.add_block(0, (8, 12, 8))
    bad.add_block(1, (12, 9, 8))
    bad.add_rotated_connection(1, BS.Right, 0, BS.Top, VT.MDPeriodic)
    bad.add_boundary(0, BS.Top, BSch.None_, (), VT.MDPeriodic)
    with pytest.raises(hip.MplbError):
        bad.prepare()


# ---- the HIP path against the oracle (FP64 arithmetic is bit-exact) ------------
def _gpu(dim, **kw):
    from mplb_amd import hip
    return hip.LatticeCase(dim, storage=hip.STORE_F64, device=0, **kw)


@pytest.mark.gpu
@pytest.mark.usefixtures("step_form")
@pytest.mark.parametrize("dim,R", ROTATIONS)
def test_gpu_rotated_pair_bitexact(dim, R):
    print("rotated pair dim=%d R=%s" % (dim, R.tolist()), flush=True)
    _, o = pair_cases(dim, R)
    _, g = pair_cases(dim, R, backend=_gpu)
    o.step(40)
    g.step(40)
    for b in (0, 1):
        assert np.array_equal(g.get_f(b), o.get_f(b)), b


@pytest.mark.gpu
@pytest.mark.usefixtures("step_form")
def test_gpu_lchannel_periodic_bitexact():
    from mplb_amd import cases
    print("closed L-channel a=9 length=14, 80 steps", flush=True)
    o = OracleCase(3)
    cases.lchannel_periodic(o, a=9, length=14, g=1e-4)
    g = _gpu(3)
    cases.lchannel_periodic(g, a=9, length=14, g=1e-4)
    o.step(80)
    g.step(80)
    for b in range(3):
        assert np.array_equal(g.get_f(b), o.get_f(b)), b
    g.update_macros()
    o.update_macros()
    for b in range(3):
        for v in range(4):
            assert np.array_equal(g.get_macro(b, v), o.get_macro(b, v)), (b, v)
    assert g.mass() == pytest.approx(o.mass(), rel=1e-13)
    # the loop carries flow through the rotated link
    assert np.max(np.abs(g.get_macro(0, 2))) > 0


@pytest.mark.gpu
def test_gpu_c3_lchannel_full_size_fp64_few_steps():
    """The L-channel at its benchmark size (blocks 128 x 256 x 128, 128^3,
    256 x 128 x 128: 10.5M nodes, the merged halo launch, the autotuned
    launch shape, the boundary kernel beside the bulk kernel) in FP64 for as
    many steps as the oracle finishes quickly: every population and macro
    bit for bit."""
    from mplb_amd import cases
    steps = 40  # one replay of the 32-step graph, then single steps
    print("closed L-channel a=128 FP64, %d steps against the oracle" % steps, flush=True)
    o = OracleCase(3)
    cases.lchannel_periodic(o, a=128, length=256, tau_l=0.8, g=1e-6)
    g = _gpu(3)
    cases.lchannel_periodic(g, a=128, length=256, tau_l=0.8, g=1e-6)
    g.step(steps)
    o.step(steps)
    for b in range(3):
        assert np.array_equal(g.get_f(b), o.get_f(b)), b
    g.update_macros()
    o.update_macros()
    for b in range(3):
        for v in range(4):
            assert np.array_equal(g.get_macro(b, v), o.get_macro(b, v)), (b, v)
    assert g.mass() == pytest.approx(o.mass(), rel=1e-13)
    g.close()
