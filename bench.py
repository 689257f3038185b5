#!/usr/bin/env python3
"""Benchmark of the MI355X lattice update on the 3D lid-driven cavity:
D3Q19, 256^3 nodes, Re = 400, lid speed 0.05 (lattice units), rest state plus
seeded density noise +-1e-4, populations stored as f - w_i in FP32.

One step = the bulk kernel (fused pull stream + macros + BGK collision over
the interior nodes) and the boundary kernel (the region programs of the
block's faces, edges and corners).  Inputs are resident in HBM before the
timed region.  The roofline figure is that of the bulk kernel, which moves
(n-2)^3 nodes x 19 populations x (read + write).  Prints one JSON line.

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


def traffic_profile_name(workload, storage, fold, xgroup):
    """The committed rocprofv3 --pmc summary for a launch shape, e.g.
    k_bulk_d3q19_256_f32_fold2_xgroup1_traffic.json."""
    return "k_bulk_%s_%s_fold%d_xgroup%d_traffic.json" % (workload, storage, fold, xgroup)


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


def host_info():
    """What the CPU baseline ran on: the processor model, the cores this
    process may use, and the OpenMP placement settings."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "OMP_PLACES": os.environ.get("OMP_PLACES")}


def _time_oracle(o, nodes, seconds):
    from oracle.oracle import get_threads
    t0 = time.perf_counter()
    o.step(1)  # first touch of every array, and the step-time estimate
    one = time.perf_counter() - t0
    steps = max(1, min(500, int(seconds / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return {"value": round(nodes * steps / dt / 1e6, 2), "unit": "MLUPS",
            "cores": get_threads(), "kind": "port", "host": host_info()}, steps, dt


def cpu_baseline(n, seconds):
    """The FP64 CPU oracle (a port of the reference step) on the host cores,
    on the full n^3 domain, for as many steps as fill about `seconds`."""
    from oracle.oracle import OracleCase
    o = OracleCase(3)
    cases.cavity3d(o, n, re=400.0, ulid=0.05, noise=True)
    out, steps, dt = _time_oracle(o, n ** 3, seconds)
    out["sample"] = ("FP64 oracle, D3Q19 cavity %d^3 (the benchmark workload), "
                     "%d steps in %.1f s" % (n, steps, dt))
    return out


def bench_lchannel(args):
    """The closed three-block L-channel (blocks 128x256x128, 128^3,
    256x128x128 for a=128): rotated periodic inlet/outlet, body force 1e-6,
    tau 0.8, FP16 storage; FP64 mass every 100 steps; max-abs error against
    an FP64 run of the same setup."""
    a, length, K = args.lc_a, 2 * args.lc_a, args.steps
    st = args.storage
    log("bench: L-channel a=%d storage=%s steps=%d" % (a, st, K))
    g = hip.LatticeCase(3, storage=STORE[st], device=0)
    p = cases.lchannel_periodic(g, a=a, length=length, tau_l=0.8, g=1e-6)
    nodes = p["nodes"]
    # warm-up (the first step autotunes), then K steps as graph replays
    # with the FP64 mass every 100 steps, timed on the wall clock
    W = args.warmup
    if W > 0:
        g.step(W)
    m0 = g.mass()
    masses = [m0]
    g.sync()
    t0 = time.perf_counter()
    done = 0
    while done < K:
        k = min(100, K - done)
        g.step(k)
        done += k
        masses.append(g.mass())
    g.sync()
    wall = time.perf_counter() - t0
    drift = max(abs(m - m0) for m in masses) / m0
    mlups = nodes * K / wall / 1e6
    # kernel split from 100 more steps with events around each kernel
    kt = 100
    _, t_kernel = g.step_timed(kt)
    t_bulk, t_bnd = g.kernel_times()
    # the nodes one bulk launch updates (interior plus the boundary nodes
    # folded into it), as for the cavity line
    fold, xgroup, h2, bulk_nodes = g.launch_shape()
    bulk_bytes = bulk_nodes * 2 * 19 * ELEM[st]
    achieved = bulk_bytes / (t_bulk / kt / 1e3) / 1e9
    log("bench: L-channel %.0f MLUPS (wall, halos and mass included; %.4f ms/step), "
        "bulk %.4f ms = %.0f GB/s, boundary %.4f ms per step, mass drift %.3e"
        % (mlups, wall / K * 1e3, t_bulk / kt, achieved, t_bnd / kt, drift))
    err = None
    if not args.no_error and st != "f64":
        ref = hip.LatticeCase(3, storage=hip.STORE_F64, device=0)
        cases.lchannel_periodic(ref, a=a, length=length, tau_l=0.8, g=1e-6)
        ref.step(W + K + kt)
        g.update_macros()
        ref.update_macros()
        umax = max(float(np.max(np.abs(ref.get_macro(b, v)))) for b in range(3)
                   for v in (1, 2, 3))
        err = {"u_max_abs": max(float(np.max(np.abs(g.get_macro(b, v) - ref.get_macro(b, v))))
                                for b in range(3) for v in (1, 2, 3)),
               "rho_max_abs": max(float(np.max(np.abs(g.get_macro(b, 0) - ref.get_macro(b, 0))))
                                  for b in range(3)),
               "u_max_fp64": umax,
               "mass_rel": abs(g.mass() - ref.mass()) / ref.mass()}
        ref.close()
        log("bench: L-channel error vs FP64: %s" % err)
    cpu = None
    if not args.no_cpu:
        from oracle.oracle import OracleCase
        o = OracleCase(3)
        cases.lchannel_periodic(o, a=a, length=length, tau_l=0.8, g=1e-6)
        cpu, csteps, cdt = _time_oracle(o, nodes, args.cpu_seconds)
        cpu["sample"] = ("FP64 oracle, the same three-block L-channel, %d steps in %.1f s"
                         % (csteps, cdt))
        del o
        log("bench: CPU baseline %s" % cpu)
    tname = traffic_profile_name("lchannel", st, fold, xgroup)
    traffic = traffic_from_profiles(tname)
    return {
        "metric": "MLUPS, three-block L-channel",
        "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": round(wall / K * 1e3, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": st,
        "data": "synthetic: rest state, body-force driven",
        "config": {"workload": "closed L-channel, blocks %dx%dx%d, %d^3, %dx%dx%d, D3Q19, "
                               "tau 0.8, g 1e-6, rotated periodic inlet/outlet, %s storage"
                               % (a, length, a, a, length, a, a, st),
                   "nodes": nodes},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": ("profiles/%s (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE "
                                        "for this launch shape, scaled by the copy-kernel "
                                        "calibration; not measured in this run)" % tname)
                     if traffic else
                     "no committed counter summary for this launch shape (%s)" % tname,
                     "kernel": "k_bulk (D3Q19, %s, step)" % st,
                     "bulk_nodes_per_launch": bulk_nodes,
                     "launch_shape": {"fold": fold, "xgroup": xgroup, "h2": h2},
                     "boundary_and_bulk_ms_per_step": round(t_kernel / kt, 5),
                     "bulk_ms_per_step": round(t_bulk / kt, 5)},
        "cpu_baseline": cpu,
        "mass_drift_max_rel": drift,
        "mass_every_100": masses,
        "wall_s": round(wall, 3),
        "error_vs_fp64": err,
    }


def bench_batch(args):
    """256 independent D2Q9 cavities (N in {64,96,128,192,256}, Re in
    [100, 1000], seeded), FP32 storage, one launch per step for all of them:
    aggregate MLUPS over a timed run, then iterate_batch to the FP64
    velocity residual 1e-8 with per-instance freezing."""
    st = args.storage
    sizes, res = cases.batch_spec(args.batch)
    nodes = sum(n * n for n in sizes)
    log("bench: %d cavities, %.2fM nodes, storage %s" % (len(sizes), nodes / 1e6, st))
    g = hip.LatticeCase(2, storage=STORE[st], device=0)
    cases.batched_cavities(g, sizes, res)
    g.step(args.warmup)
    g.sync()
    # the value: K steps as graph replays (the path iterate_batch runs);
    # then the kernel split from K more steps with events around each kernel
    t0 = time.perf_counter()
    g.step(args.steps)
    g.sync()
    wall = time.perf_counter() - t0
    mlups = nodes * args.steps / wall / 1e6
    _, ker = g.step_timed(args.steps)
    b, d = g.kernel_times()
    fold, xgroup, h2, bulk_nodes = g.launch_shape()
    achieved = bulk_nodes * 2 * 9 * ELEM[st] / (b / args.steps / 1e3) / 1e9
    tname = traffic_profile_name("batch%d" % len(sizes), st, fold, xgroup)
    traffic = traffic_from_profiles(tname)
    log("bench: batch %.0f MLUPS (wall, graph replays), event-timed kernels %.4f ms/step "
        "(bulk %.4f = %.0f GB/s, boundary %.4f)"
        % (mlups, ker / args.steps, b / args.steps, achieved, d / args.steps))
    conv = None
    if args.converge:
        t0 = time.perf_counter()
        steps, resid, done = g.iterate_batch(1e-8, args.check_period, args.max_steps)
        tconv = time.perf_counter() - t0
        ok = steps > 0
        conv = {"criteria": 1e-8, "check_period": args.check_period,
                "converged": int(ok.sum()), "of": len(sizes),
                "steps_min": int(steps[ok].min()) if ok.any() else None,
                "steps_median": float(np.median(steps[ok])) if ok.any() else None,
                "steps_max": int(steps[ok].max()) if ok.any() else None,
                "steps_run": int(done), "wall_s": round(tconv, 2),
                "per_instance_steps": [int(x) for x in steps]}
        log("bench: batch convergence %s" % {k: v for k, v in conv.items()
                                              if k != "per_instance_steps"})
    return {
        "metric": "aggregate MLUPS, batched cavities",
        "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": st,
        "data": "synthetic: %d seeded cavities at rest" % len(sizes),
        "config": {"workload": "%d D2Q9 lid-driven cavities, N in {64..256}, Re in "
                               "[100,1000], lid 0.05, %s storage" % (len(sizes), st),
                   "nodes": nodes},
        "kernel_ms_per_step": round(ker / args.steps, 5),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": ("profiles/%s (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE "
                                        "for this launch shape, scaled by the copy-kernel "
                                        "calibration; not measured in this run)" % tname)
                     if traffic else
                     "no committed counter summary for this launch shape (%s)" % tname,
                     "kernel": "k_bulk (D2Q9, %s, step)" % st,
                     "bulk_nodes_per_launch": bulk_nodes,
                     "launch_shape": {"fold": fold, "xgroup": xgroup, "h2": h2},
                     "bulk_ms_per_step": round(b / args.steps, 5)},
        "timing": "value: K steps of graph replays (the path iterate_batch runs); kernel times: "
                  "HIP events around every kernel of K more steps launched one by one",
        "convergence": conv,
    }


def bench_cavity2d(args):
    """The 2D cavity: D2Q9 128^2, Re 100, lid 0.05, FP64 (bit-exact with the
    oracle), 10,000 steps from rest; small enough to live in the caches, so
    it measures launch and latency costs rather than HBM."""
    n, K = 128, args.steps
    g = hip.LatticeCase(2, storage=STORE[args.storage], device=0)
    cases.cavity2d(g, n, re=100.0, ulid=0.05)
    g.step(args.warmup)
    g.sync()
    t0 = time.perf_counter()
    g.step(K)
    g.sync()
    wall = time.perf_counter() - t0
    mlups = n * n * K / wall / 1e6
    log("bench: cavity2d %d^2 %s, %d steps in %.3f s, %.0f MLUPS" % (n, args.storage, K, wall,
                                                                     mlups))
    return {
        "metric": "MLUPS and wall time, 2D cavity", "value": round(mlups, 1), "unit": "MLUPS",
        "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": round(wall / K * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.storage, "data": "synthetic: rest state",
        "config": {"workload": "2D lid-driven cavity, D2Q9 128^2, Re=100, lid 0.05, %s"
                               % args.storage, "nodes": n * n},
        "wall_s": round(wall, 4),
    }


def bench_unet(args):
    """Training throughput of the warm-start U-Net alone (the C5 training
    path): mini-batches of 32 synthetic 128^2 inputs and targets resident
    on the device, forward + backward + Adam per step; the roofline is the
    convolutions' dense BF16 matrix work against the matrix-core peak."""
    from mplb_amd import warmstart
    n, B = args.ws_n, 32
    enc = tuple(int(c) for c in args.ws_enc.split(","))
    rng = np.random.default_rng(0)
    # a pool of at most 32 mini-batches, cycled as epochs
    pool = min(max(args.steps, 1), 32)
    epochs = max(1, round(args.steps / pool))
    steps = epochs * pool
    count = B * pool
    x = rng.standard_normal((count, n, n, warmstart.IN_CH)).astype(np.float32)
    x[..., 0] = (rng.random((count, n, n)) < 0.05).astype(np.float32)
    y = (0.1 * rng.standard_normal((count, n, n, 3))).astype(np.float32)
    net = warmstart.UNet(n=n, batch=B, seed=1, enc=enc, bott=args.ws_bott)
    if args.warmup > 0:
        w = min(args.warmup, pool)
        warmstart.train(net, x[:B * w], y[:B * w], 1)
    sync = warmstart.load().mplbws_sync
    sync()
    t0 = time.perf_counter()
    warmstart.train(net, x, y, epochs)
    sync()
    wall = time.perf_counter() - t0
    ms = wall / steps * 1e3
    layers = warmstart.layer_specs(enc, args.ws_bott)
    # forward, data gradient (not for the first layer) and weight gradient
    flops = sum(2.0 * 9 * ci * co * B * (n >> lev) ** 2 * (2 if name == "e0a" else 3)
                for name, ci, co, _, lev in layers)
    tflops = flops / (ms * 1e-3) / 1e12
    log("bench: U-Net training %.3f ms/step, %.1f TFLOP/s of convolution work" % (ms, tflops))
    return {
        "metric": "U-Net training throughput", "value": round(B / (ms * 1e-3), 1),
        "unit": "images/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (fp32 accumulation)", "data": "synthetic",
        "config": {"workload": "warm-start U-Net training step, 32 images of %d^2, "
                               "encoder %s, bottleneck %d, %d parameters"
                               % (n, "-".join(map(str, enc)), args.ws_bott,
                                  warmstart.param_count(layers))},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 5), "traffic": None,
                     "kernel": "whole step (convolutions, pooling, loss, Adam)"},
        "cpu_baseline": None,
    }


def secondary(args):
    """The other configurations, measured in the same process after the
    headline line's own timed region: the 2D cavity (FP64, 10,000 steps),
    the three-block L-channel (FP16, 500 steps), the 256 batched cavities
    (FP32, to the residual 1e-8 per instance) and the learned warm start's
    time-to-residual on the 512 held-out cases with the committed network
    (velocity against FP64 runs of the first 64)."""
    import copy
    from mplb_amd import warmstart
    out = {}

    def sub(**kw):
        a = copy.copy(args)
        for k, v in kw.items():
            setattr(a, k, v)
        return a

    t0 = time.perf_counter()
    c1 = bench_cavity2d(sub(steps=10000, warmup=100, storage="f64"))
    out["cavity2d_128_f64"] = {k: c1[k] for k in ("value", "unit", "ms_per_step", "wall_s")}
    c3 = bench_lchannel(sub(steps=500, warmup=20, storage="f16", lc_a=128, no_error=True,
                            no_cpu=True))
    out["lchannel_f16"] = {"value": c3["value"], "unit": c3["unit"],
                           "ms_per_step": c3["ms_per_step"],
                           "bulk_frac": c3["roofline"]["frac"],
                           "mass_drift_max_rel": c3["mass_drift_max_rel"]}
    c4 = bench_batch(sub(steps=1000, warmup=50, storage="f32", batch=256, converge=True,
                         check_period=1000, max_steps=400000))
    conv = dict(c4.get("convergence") or {})
    conv.pop("per_instance_steps", None)
    out["batch256_f32"] = {"value": c4["value"], "unit": c4["unit"],
                           "ms_per_step": c4["ms_per_step"], "convergence": conv}
    weights = os.path.join(ROOT, "mplb_amd", "data", "unet128_c5.npz")
    if os.path.exists(weights):
        ev = warmstart.evaluate_saved(weights, n_test=512, fp64_cases=64, log=log)
        keep = ("test_cases", "rest_seconds", "learned_seconds", "speedup",
                "per_case_step_ratio_median", "work_ratio_sum_steps", "rest_steps_max",
                "learned_steps_max", "inference_seconds", "fallback_cases", "restarted_cases",
                "prediction_rel_l2_median", "fp64_reference")
        out["warmstart_time_to_residual"] = {k: ev[k] for k in keep if k in ev}
        out["warmstart_time_to_residual"]["weights"] = "mplb_amd/data/unet128_c5.npz"
    out["seconds"] = round(time.perf_counter() - t0, 1)
    return out


def bench_warmstart(args):
    """Learned warm start on 128^2 cavities with 0-3 rectangular obstacles:
    training data from the batched solver, U-Net training on the matrix
    cores, then time-to-residual 1e-8 of learned versus rest starts on the
    held-out cases."""
    from mplb_amd import warmstart
    enc = tuple(int(c) for c in args.ws_enc.split(","))
    coarse = tuple((int(k), float(w)) for k, w in
                   (item.split(":") for item in args.ws_coarse.split(",") if item))
    res, net = warmstart.run(n=args.ws_n, n_train=args.ws_train, n_test=args.ws_test,
                             epochs=args.ws_epochs, enc=enc, bott=args.ws_bott, coarse=coarse,
                             diagnose=args.ws_diagnose, log=log)
    log("bench: warm start %s" % res)
    if args.ws_save:
        warmstart.save_weights(net, args.ws_save, n_train=args.ws_train)
        log("bench: network weights written to %s" % args.ws_save)
    return {
        "metric": "time-to-residual 1e-8, learned vs rest initialisation",
        "value": res["speedup"], "unit": "x (rest seconds / learned seconds)", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 lattice, bf16 network",
        "data": "synthetic: seeded cavities with rectangular obstacles",
        "config": {"workload": "%d held-out %d^2 D2Q9 cavities, Re log-uniform in "
                               "[100,1000], 0-3 rectangles; U-Net %d parameters trained "
                               "on %d cases" % (res["test_cases"], res["n"], res["params"],
                                                res["train_cases"])},
        "warmstart": res,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--storage", choices=sorted(STORE), default=None,
                    help="population storage (default: f32; f16 for --workload lchannel, "
                         "the three-block L-channel's configured form)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the other configurations measured after the headline")
    ap.add_argument("--no-error", action="store_true",
                    help="skip the FP64 comparison run")
    ap.add_argument("--ws-n", type=int, default=128)
    ap.add_argument("--ws-train", type=int, default=4096)
    ap.add_argument("--ws-test", type=int, default=512)
    ap.add_argument("--ws-epochs", type=int, default=200)
    ap.add_argument("--ws-enc", default="24,48,64,64", help="encoder widths per level")
    ap.add_argument("--ws-bott", type=int, default=64, help="bottleneck width")
    ap.add_argument("--ws-coarse", default="2:4,4:4",
                    help="large-scale loss terms, pooling-levels:weight pairs ('' for none)")
    ap.add_argument("--ws-save", default="",
                    help="write the trained network's weights (.npz) to this path")
    ap.add_argument("--ws-diagnose", action="store_true",
                    help="also solve from exact velocity / exact density splits")
    ap.add_argument("--workload", choices=("cavity3d", "cavity2d", "lchannel", "batch",
                                           "warmstart", "unet"),
                    default="cavity3d",
                    help="cavity3d: the headline line; lchannel and batch: the "
                         "multi-block and batched workloads")
    ap.add_argument("--lc-a", type=int, default=128, help="L-channel cross-section")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--converge", action="store_true")
    ap.add_argument("--check-period", type=int, default=1000)
    ap.add_argument("--max-steps", type=int, default=400000)
    args = ap.parse_args()
    if args.storage is None:
        args.storage = "f16" if args.workload == "lchannel" else "f32"
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the multi-GPU path "
              "(face-slab halos over xGMI) is a design, not a run. Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "lchannel":
        print(json.dumps(bench_lchannel(args)), flush=True)
        return
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "cavity2d":
        print(json.dumps(bench_cavity2d(args)), flush=True)
        return
    if args.workload == "warmstart":
        print(json.dumps(bench_warmstart(args)), flush=True)
        return
    if args.workload == "unet":
        print(json.dumps(bench_unet(args)), flush=True)
        return

    n, K, W = args.n, args.steps, args.warmup
    st = args.storage
    log("bench: D3Q19 cavity %d^3 storage=%s warmup=%d steps=%d" % (n, st, W, K))
    g = hip.LatticeCase(3, storage=STORE[st], device=0)
    cases.cavity3d(g, n, re=400.0, ulid=0.05, noise=True)
    if W > 0:
        g.step(W)
    # the timed region: K steps of the production path (mplbhip_step, the
    # same replays of captured 32-step graphs that Iterate runs)
    g.sync()
    t0 = time.perf_counter()
    g.step(K)
    g.sync()
    wall = time.perf_counter() - t0
    # then K more steps launched one by one with HIP events around each
    # kernel on the stream it runs on (a captured graph cannot carry
    # timing events): the kernels' average durations for the roofline
    g.sync()
    t1 = time.perf_counter()
    ms_gpu_span, ms_kernel = g.step_timed(K)
    g.sync()
    wall_events = time.perf_counter() - t1
    ms_bulk, ms_bnd = g.kernel_times()
    nodes = n ** 3
    mlups = nodes * K / wall / 1e6
    q = 19
    bulk_s = ms_bulk / K / 1e3
    # algorithmic bytes per bulk launch: 2 Q s for every node it updates
    # (the interior, plus the boundary nodes folded into it at fold level
    # 1 or 2); the interior-only figure is kept beside it
    fold, xgroup, h2, bulk_nodes = g.launch_shape()
    bulk_bytes = bulk_nodes * 2 * q * ELEM[st]
    interior_bytes = (n - 2) ** 3 * 2 * q * ELEM[st]
    achieved = bulk_bytes / bulk_s / 1e9
    log("bench: wall %.3f s (graph replays), %.3f s with per-kernel events (gpu span "
        "%.3f ms), bulk kernel avg %.4f ms (%.0f GB/s), boundary kernel avg %.4f ms, "
        "%.0f MLUPS" % (wall, wall_events, ms_gpu_span, bulk_s * 1e3, achieved, ms_bnd / K,
                        mlups))

    copy_bytes = 2 << 30
    copy_ms = hip.copy_probe(copy_bytes, iters=20)
    copy_gbs = 2 * copy_bytes / (copy_ms / 1e3) / 1e9

    err = None
    if not args.no_error and st != "f64":
        steps_done = W + 2 * K
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
        log("bench: CPU oracle baseline (about %.0f s at %d^3)" % (args.cpu_seconds, n))
        cpu = cpu_baseline(n, args.cpu_seconds)

    tname = traffic_profile_name("d3q19_%d" % n, st, fold, xgroup)
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
                     "traffic": traffic_from_profiles(tname),
                     "traffic_source": ("profiles/%s (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE "
                                        "for this launch shape, scaled by the copy-kernel "
                                        "calibration; not measured in this run)" % tname)
                     if traffic_from_profiles(tname) else
                     "no committed counter summary for this launch shape (%s)" % tname,
                     "kernel": "k_bulk (D3Q19, %s, step)" % st,
                     "algorithmic_bytes": bulk_bytes,
                     "bulk_nodes_per_launch": bulk_nodes,
                     "launch_shape": {"fold": fold, "xgroup": xgroup, "h2": h2},
                     "achieved_interior_only": round(interior_bytes / bulk_s / 1e9, 1),
                     "kernel_avg_ms": round(bulk_s * 1e3, 5),
                     "boundary_kernel_avg_ms": round(ms_bnd / K, 5),
                     "step_kernels_avg_ms": round(ms_kernel / K, 5),
                     "ms_per_step_event_timed": round(wall_events / K * 1e3, 5),
                     "timing": "value: K steps of graph replays (the production path); "
                               "kernel times: HIP events around every kernel of K more "
                               "steps launched one by one on the same streams",
                     "copy_probe_GBs": round(copy_gbs, 1),
                     "frac_of_copy": round(achieved / copy_gbs, 4)},
        "cpu_baseline": cpu,
    }
    if err is not None:
        out["error_vs_fp64"] = err
    if not args.no_secondary and args.gpus == 1 and n == 256 and st == "f32":
        out["secondary"] = secondary(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
