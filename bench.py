#!/usr/bin/env python3
"""Benchmark of the MI355X lattice update on the 3D lid-driven cavity:
D3Q19, 256^3 nodes, Re = 400, lid speed 0.05 (lattice units), rest state plus
seeded density noise +-1e-4, populations stored as f - w_i in FP32.

One step = one fused pull stream-collide sweep over the whole lattice (the
boundary programs run inside the same kernel).  Inputs are resident in HBM
before the timed region.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 2000] [--warmup 50]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mplb_amd import cases, hip  # noqa: E402

METRIC = ("MLUPS + % HBM roofline, D3Q19 256³; time-to-residual 1e-8 "
          "learned vs plain")
HBM_PEAK_GBS = 8000.0
STORE = {"f64": hip.STORE_F64, "f32": hip.STORE_F32, "f16": hip.STORE_F16}
ELEM = {"f64": 8, "f32": 4, "f16": 2}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel calibration)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as fh:
            d = json.load(fh)
        return d.get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(n, steps):
    """The FP64 CPU oracle (a port of the reference step) on the host cores,
    on the full 256^3 domain for a few steps."""
    from oracle.oracle import OracleCase, get_threads
    o = OracleCase(3)
    cases.cavity3d(o, n, re=400.0, ulid=0.05, noise=True)
    o.step(1)  # first touch of every array
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt / 1e6, "unit": "MLUPS",
            "cores": get_threads(), "kind": "port",
            "sample": "FP64 oracle, D3Q19 cavity %d^3, %d steps (%.1f s)"
                      % (n, steps, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--storage", choices=sorted(STORE), default="f32")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true",
                    help="skip the FP64 comparison run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the multi-GPU path "
              "(face-slab halos over xGMI) is a design, not a run. Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    n, K, W = args.n, args.steps, args.warmup
    st = args.storage
    log("bench: D3Q19 cavity %d^3 storage=%s warmup=%d steps=%d" % (n, st, W, K))
    g = hip.LatticeCase(3, storage=STORE[st], device=0)
    cases.cavity3d(g, n, re=400.0, ulid=0.05, noise=True)
    if W > 0:
        g.step(W)
    g.sync()
    t0 = time.perf_counter()
    ms_gpu_span, ms_kernel = g.step_timed(K)
    g.sync()
    wall = time.perf_counter() - t0
    nodes = n ** 3
    mlups = nodes * K / wall / 1e6
    kavg_s = ms_kernel / K / 1e3
    q = 19
    bytes_per_launch = nodes * 2 * q * ELEM[st]
    achieved = bytes_per_launch / kavg_s / 1e9
    log("bench: wall %.3f s, gpu span %.3f ms, kernel avg %.4f ms, %.0f MLUPS, "
        "%.0f GB/s" % (wall, ms_gpu_span, kavg_s * 1e3, mlups, achieved))

    copy_bytes = 2 << 30
    copy_ms = hip.copy_probe(copy_bytes, iters=20)
    copy_gbs = 2 * copy_bytes / (copy_ms / 1e3) / 1e9

    err = None
    if not args.no_error and st != "f64":
        steps_done = W + K
        log("bench: FP64 reference run, %d steps, for the max-abs error" % steps_done)
        ref = hip.LatticeCase(3, storage=hip.STORE_F64, device=0)
        cases.cavity3d(ref, n, re=400.0, ulid=0.05, noise=True)
        ref.step(steps_done)
        g.update_macros()
        ref.update_macros()
        err = {}
        for v, name in enumerate(("rho", "u", "v", "w")):
            err[name] = float(np.max(np.abs(g.get_macro(0, v) - ref.get_macro(0, v))))
        fa, fb = g.get_f(0), ref.get_f(0)
        err["f"] = float(np.max(np.abs(fa - fb)))
        del fa, fb
        mg, mr = g.mass(), ref.mass()
        err["mass_rel"] = abs(mg - mr) / mr
        ref.close()
        log("bench: max-abs error vs FP64: %s" % err)

    cpu = None
    if not args.no_cpu:
        log("bench: CPU oracle baseline (%d steps at %d^3)" % (args.cpu_steps, n))
        cpu = cpu_baseline(n, args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(wall / K * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": st,
        "data": "synthetic: rest state with seeded density noise +-1e-4",
        "config": {"workload": "3D lid-driven cavity, D3Q19 %d^3, Re=400, lid 0.05, "
                               "%s storage (f - w_i)" % (n, st),
                   "nodes": nodes, "lattice": "D3Q19", "storage": st,
                   "bytes_per_lattice_update": 2 * q * ELEM[st]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(
                         "k_lattice_d3q19_%d_%s_traffic.json" % (n, st)),
                     "kernel_avg_ms": round(kavg_s * 1e3, 5),
                     "copy_probe_GBs": round(copy_gbs, 1)},
        "cpu_baseline": cpu,
    }
    if err is not None:
        out["error_vs_fp64"] = err
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
