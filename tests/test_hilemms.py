"""The C++ HiLeMMS layer (mplb_amd/hilemms): the reference's own Apps
compile and link against it unchanged, the example drivers build, and the
JSON reader parses case files as the reference's ParseJson does."""
import json
import os
import subprocess

import pytest

from mplb_amd import build as hipbuild

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HILEMMS = os.path.join(ROOT, "mplb_amd", "hilemms")
REF_APPS = "/root/reference/Apps"
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.mark.skipif(not os.path.isdir(REF_APPS), reason="reference Apps not mounted")
@pytest.mark.parametrize("app,dim", [("Apps/2DCavity/lbm2d_cavity.cpp", 2),
                                     ("Apps/3DCavity/lbm3d_cavity.cpp", 3),
                                     ("Apps/3DLChannel/lbm3d_L.cpp", 3),
                                     ("Apps/3DCavity-Swap/lbm3d_cavity_swap.cpp", 3),
                                     ("Tests/ConservationTest3D/conservation3d.cpp", 3)])
def test_reference_apps_link_unchanged(app, dim, tmp_path):
    """The unmodified App sources (and the conservation test program, which
    never steps and defines no body-force hook) build against the HiLeMMS
    headers and link against libmplbhip.so (compile and link only: nothing
    of the reference is run or shipped)."""
    out = tmp_path / "app"
    subprocess.check_call(["make", "-s", "-B", "-C", HILEMMS,
                           "APP=" + os.path.join(os.path.dirname(REF_APPS), app),
                           "DIM=%d" % dim, "OUT=" + str(out)])
    assert out.exists()


def test_example_drivers_build():
    for p in hipbuild.build_apps():
        assert os.access(p, os.X_OK)


@pytest.mark.parametrize("name,dim", [("case_cavity2d_16.json", 2),
                                      ("case_cavity3d_10_toporder.json", 3),
                                      ("case_lchannel_9.json", 3),
                                      ("case_cavity3d_swap_12.json", 3),
                                      ("case_channel3d_fdperiodic.json", 3),
                                      ("case_cavity3d_d3q15_thermal.json", 3)])
def test_json_reader_matches_python(name, dim):
    path = os.path.join(GOLDEN, name)
    cfg = json.load(open(path))
    exe = hipbuild.app_path("case_driver%dd" % dim)
    if not os.path.exists(exe):
        hipbuild.build_apps()
    out = subprocess.check_output([exe, "Config=" + path, "DryRun=1"], text=True)
    lines = [ln for ln in out.splitlines() if ln.startswith(("case=", "bc "))]
    head = dict(kv.split("=", 1) for kv in lines[0].split())
    assert int(head["dim"]) == cfg["SpaceDim"]
    assert int(head["blocks"]) == len(cfg["BlockIds"])
    nbc = sum(1 for k in cfg if k.startswith("BoundaryCondition"))
    assert int(head["bcs"]) == nbc
    assert float(head["tau"]) == cfg["TauRef"][0]
    assert float(head["mesh"]) == cfg["MeshSize"]
    assert int(head["steps"]) == cfg["TimeStepsToRun"]
    bcs = lines[1:]
    for k, line in enumerate(bcs):
        f = dict(kv.split("=", 1) for kv in line.split()[1:])
        bc = cfg["BoundaryCondition%d" % k]
        assert int(f["block"]) == bc["BlockIndex"]
        assert int(f["nvars"]) == len(bc["GivenVars"])
        if bc["BoundaryScheme"] == "EQMDiffuseREfl":
            assert int(f["scheme"]) == 24
