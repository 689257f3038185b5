"""Builds libmplbhip.so (the lattice library) and libmplbws.so (the warm
start network kernels) in-tree for gfx950 with hipcc (no JIT, no cache
outside the repository): `python -m mplb_amd.build`."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmplbhip.so")
ARCH = os.environ.get("MPLBHIP_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off: plain VALU FP32 only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
SOURCES = ["plan.cpp", "mplbhip.cpp", "lattice_kernels.hip", "swap_kernels.hip"]


WS_OUT = os.path.join(HERE, "libmplbws.so")
WS_SOURCES = ["warmstart.hip"]


def build(verbose=False):
    out = _build_lib(SOURCES, OUT, verbose)
    _build_lib(WS_SOURCES, WS_OUT, verbose)
    return out


def _build_lib(sources, target, verbose):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for src in sources:
        obj = os.path.join(HERE, "_build", os.path.splitext(src)[0] + ".o")
        cmd = [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = target + ".tmp"
    cmd = [hipcc, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, target)
    return target


APPS_DIR = os.path.join(HERE, "hilemms")
APPS = [("apps/cavity3d.cpp", 3, "cavity3d"), ("apps/case_driver.cpp", 2, "case_driver2d"),
        ("apps/case_driver.cpp", 3, "case_driver3d")]


def app_path(name):
    return os.path.join(APPS_DIR, "build", name)


def build_apps(verbose=False):
    """The HiLeMMS example drivers, linked against libmplbhip.so."""
    out = []
    for src, dim, name in APPS:
        cmd = ["make", "-s", "-B", "-C", APPS_DIR, "APP=" + os.path.join(APPS_DIR, src),
               "DIM=%d" % dim, "OUT=" + app_path(name)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        out.append(app_path(name))
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_apps(verbose="-v" in sys.argv))
