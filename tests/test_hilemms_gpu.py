"""The HiLeMMS C++ drivers on the GPU against the CPU oracle: a JSON case
read by the C++ layer and run through DefineCase ... Iterate must reproduce
the oracle's macroscopic fields bit for bit (FP64 storage)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from mplb_amd import build as hipbuild
from mplb_amd import cases
from oracle.oracle import OracleCase

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(exe, *args, cwd=None):
    """Run a HiLeMMS driver; its checkpoint files go to cwd (a temporary
    directory, never the repository)."""
    print("run", os.path.basename(exe), *args, flush=True)
    r = subprocess.run([exe] + list(args), capture_output=True, text=True, timeout=240,
                       env=dict(os.environ, MPLB_STORAGE="f64"), cwd=cwd)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def split_dump(path, sizes, nvars):
    data = np.fromfile(path, dtype=np.float64)
    out, p = {}, 0
    for b in sorted(sizes):
        shp = tuple(reversed(sizes[b]))
        n = int(np.prod(shp))
        out[b] = [data[p + v * n: p + (v + 1) * n].reshape(shp) for v in range(nvars)]
        p += nvars * n
    assert p == data.size
    return out


@pytest.mark.parametrize("name,dim", [("case_cavity2d_16.json", 2),
                                      ("case_cavity3d_10_toporder.json", 3),
                                      ("case_lchannel_9.json", 3)])
def test_case_driver_json_matches_oracle(name, dim, tmp_path):
    path = os.path.join(GOLDEN, name)
    cfg = json.load(open(path))
    dump = str(tmp_path / "macros.bin")
    run(hipbuild.app_path("case_driver%dd" % dim), "Config=" + path, "Dump=" + dump,
        cwd=str(tmp_path))
    o = OracleCase(dim)
    steps, period, _ = cases.from_config(o, cfg)
    o.step(steps)
    o.update_macros()
    got = split_dump(dump, o.sizes, dim + 1)
    for b in o.sizes:
        for v in range(dim + 1):
            d = float(np.max(np.abs(got[b][v] - o.get_macro(b, v))))
            assert d == 0.0, (name, b, v, d)


def test_cavity3d_driver_matches_oracle(tmp_path):
    dump = str(tmp_path / "macros.bin")
    run(hipbuild.app_path("cavity3d"), "N=12", "Re=100", "Lid=0.05", "Steps=20",
        "Check=10", "Noise=1", "Dump=" + dump)
    o = OracleCase(3)
    cases.cavity3d(o, 12, re=100.0, ulid=0.05, noise=True)
    o.step(20)
    o.update_macros()
    got = split_dump(dump, o.sizes, 4)
    for v in range(4):
        assert np.array_equal(got[0][v], o.get_macro(0, v)), v


def test_cavity3d_driver_converges_with_residual(tmp_path):
    """Iterate(criteria, checkPeriod) through the C++ layer stops once the
    printed residual is below the criterion."""
    out = run(hipbuild.app_path("cavity3d"), "N=16", "Re=20", "Lid=0.05",
              "Criteria=1e-3", "Check=100")
    res = [float(l.split("=")[1]) for l in out.splitlines() if l.startswith("Residual of")]
    assert res, out[-500:]
    assert max(res[-4:]) < 1e-3


def test_force_hook_constant_and_time_dependent(tmp_path):
    """StreamCollision calls the App's UpdateMacroscopicBodyForce(iter * dt)
    before every step (Src/evolution.cpp:137-151).  A hook that sets a
    constant acceleration (case_driver Force=g0) runs through Iterate and
    matches the oracle with that acceleration bit for bit; a hook whose
    acceleration changes inside Iterate (ForceSchedule=g0: g0 (1 +
    floor(k/7) mod 3) at step k) stops the run with an error rather than
    being ignored."""
    path = os.path.join(GOLDEN, "case_channel2d_force.json")
    cfg = json.load(open(path))
    dump = str(tmp_path / "macros.bin")
    g0 = 2e-3
    run(hipbuild.app_path("case_driver2d"), "Config=" + path, "Dump=" + dump,
        "Force=%r" % g0, cwd=str(tmp_path))
    o = OracleCase(2)
    steps, period, _ = cases.from_config(o, cfg)
    o.set_accel(0, (g0, 0.0))
    o.step(steps)
    o.update_macros()
    got = split_dump(dump, o.sizes, 3)
    for v in range(3):
        d = float(np.max(np.abs(got[0][v] - o.get_macro(0, v))))
        assert d == 0.0, (v, d)
    assert float(np.max(np.abs(got[0][1]))) > 0  # the force drives a flow
    r = subprocess.run([hipbuild.app_path("case_driver2d"), "Config=" + path,
                        "ForceSchedule=%r" % g0], capture_output=True, text=True, timeout=240,
                       env=dict(os.environ, MPLB_STORAGE="f64", MPLB_HDF5="0"),
                       cwd=str(tmp_path))
    assert r.returncode != 0
    assert "changed the body force at time step 7" in r.stderr, r.stderr[-2000:]


@pytest.mark.parametrize("name,force", [("case_cavity3d_swap_12.json", 0.0),
                                        ("case_channel3d_fdperiodic.json", 3e-3)])
def test_case_driver_swap_scheme_matches_oracle(name, force, tmp_path):
    """A JSON case with SchemeType Scheme_StreamCollision_Swap runs
    Iterate(SwapStreamCollision, ...) as Apps/3DCavity-Swap does
    (lbm3d_cavity_swap.cpp:101-107): the swap cavity and a channel joined to
    itself through FDPeriodic faces, driven by BodyForce_1st_Swap, match the
    oracle's swap scheme bit for bit."""
    path = os.path.join(GOLDEN, name)
    cfg = json.load(open(path))
    dump = str(tmp_path / "macros.bin")
    args = ["Config=" + path, "Dump=" + dump]
    if force:
        args.append("Force=%r" % force)
    out = run(hipbuild.app_path("case_driver3d"), *args, cwd=str(tmp_path))
    assert "swap stream-collision scheme" in out
    o = OracleCase(3)
    steps, period, _ = cases.from_config(o, cfg)
    if force:
        o.set_accel(0, (force, 0.0, 0.0))
    o.step(steps)
    o.update_macros()
    got = split_dump(dump, o.sizes, 4)
    for v in range(4):
        d = float(np.max(np.abs(got[0][v] - o.get_macro(0, v))))
        assert d == 0.0, (name, v, d)
    assert float(np.max(np.abs(got[0][1]))) > 0


def test_case_driver_d3q15_thermal_matches_oracle(tmp_path):
    """A JSON case with LatticeName d3q15 and Collision_BGKThermal4th, its
    Variable_T macro set by the App hook (Temperature=1: the cavity's
    EQMDiffuseRefl walls are built on the T = 1 equilibrium and add mass at
    other temperatures, in the reference as here): DefineComponents selects
    the D3Q15 tables, PrepareFlowField the thermal collision (per-node tau
    = tauRef / rho, the 3rd- and 4th-order terms), and the temperature
    reaches the device with the initial macros; the result matches the
    oracle bit for bit."""
    path = os.path.join(GOLDEN, "case_cavity3d_d3q15_thermal.json")
    cfg = json.load(open(path))
    dump = str(tmp_path / "macros.bin")
    run(hipbuild.app_path("case_driver3d"), "Config=" + path, "Dump=" + dump,
        "Temperature=1.0", cwd=str(tmp_path))
    o = OracleCase(3, lattice=15)
    steps, period, _ = cases.from_config(o, cfg)
    n = cfg["BlockSize"]
    o.set_temperature(0, np.full(tuple(reversed(n)), 1.0))
    o.step(steps)
    o.update_macros()
    got = split_dump(dump, o.sizes, 5)  # rho, u, v, w and the untouched T
    for v in range(4):
        d = float(np.max(np.abs(got[0][v] - o.get_macro(0, v))))
        assert d == 0.0, (v, d)
    assert np.all(got[0][4] == 1.0)


def test_swap_cycle_must_match_scheme(tmp_path):
    """StreamCollision handed to a swap case (or the reverse) is an error,
    not a silent run of the other scheme."""
    cfg = json.load(open(os.path.join(GOLDEN, "case_cavity3d_10_toporder.json")))
    cfg["CollisionType"] = ["Collision_BGKIsothermal2nd_Swap"]
    path = str(tmp_path / "mixed.json")
    json.dump(cfg, open(path, "w"))
    r = subprocess.run([hipbuild.app_path("case_driver3d"), "Config=" + path],
                       capture_output=True, text=True, timeout=240,
                       env=dict(os.environ, MPLB_STORAGE="f64", MPLB_HDF5="0"),
                       cwd=str(tmp_path))
    assert r.returncode != 0
    assert "are implemented for Scheme_StreamCollision" in r.stderr, \
        r.stderr[-2000:]


H5DUMP = "/opt/conda/bin/h5dump"


@pytest.mark.parametrize("name,dim", [("case_cavity2d_16.json", 2),
                                      ("case_lchannel_9.json", 3)])
def test_hdf5_checkpoint_and_restart(name, dim, tmp_path):
    """WriteFlowfieldToHdf5 / WriteDistributionsToHdf5 /
    WriteNodePropertyToHdf5 at every checkpoint (Src/flowfield.cpp:118-137)
    into <case>_<block>_T<step>.h5 with the OPS layout that
    Src/PostProcess.py:107-196 reads (group per block, dataset
    <field>_<block> with its halo, components fastest); a second run started
    from the step-20 files (DefineComponents/DefineMacroVars with
    CurrentTimeStep = 20, Src/model.cpp:299-311, 379-388) must end exactly
    where the uninterrupted run ends."""
    cfg = json.load(open(os.path.join(GOLDEN, name)))
    assert cfg["CheckPeriod"] == 20 and cfg["TimeStepsToRun"] == 40
    full = tmp_path / "full"
    full.mkdir()
    dump_full = str(tmp_path / "full.bin")
    run(hipbuild.app_path("case_driver%dd" % dim), "Config=" + os.path.join(GOLDEN, name),
        "Dump=" + dump_full, cwd=str(full))
    case = cfg["CaseName"]
    sizes = cfg["BlockSize"]
    for k, bname in enumerate(cfg["BlockNames"]):
        n = sizes[dim * k: dim * (k + 1)]
        for step in (20, 40):
            fn = full / ("%s_%s_T%d.h5" % (case, bname, step))
            assert fn.exists(), fn
        if os.path.exists(H5DUMP):
            # -A: the header with attribute values, no dataset data
            hdr = subprocess.run([H5DUMP, "-A", str(full / ("%s_%s_T20.h5" % (case, bname)))],
                                 capture_output=True, text=True, timeout=60).stdout
            shp = ", ".join(str(x + 2) for x in reversed(n[1:]))
            for field, comps in (("f", 9 if dim == 2 else 19), ("rho", 1),
                                 ("CoordinateXYZ", dim), ("GeometryProperty", 1),
                                 ("NodeType_Fluid", 1)):
                want = 'DATASET "%s_%s"' % (field, bname)
                assert want in hdr, (want, hdr[:2000])
                block = hdr.split(want)[1].split("DATASET")[0]
                assert "( %s, %d )" % (shp, (n[0] + 2) * comps) in block, (field, block[:500])
            assert 'GROUP "%s"' % bname in hdr
            assert '(0): "ops_block"' in hdr and '(0): "ops_dat"' in hdr
    # restart from the step-20 files in a copy of the directory
    rst = tmp_path / "restart"
    rst.mkdir()
    for fn in full.iterdir():
        if "_T20.h5" in fn.name:
            (rst / fn.name).write_bytes(fn.read_bytes())
    cfg2 = dict(cfg, CurrentTimeStep=20, TimeStepsToRun=20)
    path2 = str(tmp_path / "restart.json")
    json.dump(cfg2, open(path2, "w"))
    dump_rst = str(tmp_path / "restart.bin")
    run(hipbuild.app_path("case_driver%dd" % dim), "Config=" + path2, "Dump=" + dump_rst,
        cwd=str(rst))
    a = np.fromfile(dump_full, dtype=np.float64)
    b = np.fromfile(dump_rst, dtype=np.float64)
    assert a.size == b.size and np.array_equal(a, b)
    assert (rst / ("%s_%s_T40.h5" % (case, cfg["BlockNames"][0]))).exists()
