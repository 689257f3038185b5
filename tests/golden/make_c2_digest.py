"""Writes tests/golden/c2_fp64_digest.json: the FP64 oracle's state after the
full C2 run (3D lid-driven cavity, D3Q19 256^3, Re 400, lid 0.05, density
noise 1 +- 1e-4 from splitmix64 with seed 0x4D504C42, 2,000 steps), as a
SHA-256 of f in [q][k][j][i] order (little-endian float64) plus the total
mass and a few probe values.  The oracle's per-node arithmetic does not
depend on its thread count, so the digest is reproducible.

  python tests/golden/make_c2_digest.py [--n 256] [--steps 2000] [--threads 8]

About 20 minutes on 8 cores for the full size.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from mplb_amd import cases  # noqa: E402
from oracle.oracle import OracleCase, set_threads  # noqa: E402


def digest(f):
    """SHA-256 of a [q][k][j][i] float64 array in little-endian byte order."""
    a = np.ascontiguousarray(f, dtype="<f8")
    h = hashlib.sha256()
    flat = a.reshape(-1)
    step = 1 << 24
    for s in range(0, flat.size, step):
        h.update(flat[s:s + step].tobytes())
    return h.hexdigest()


PROBES = [(0, 0, 0, 0), (5, 128, 255, 17), (12, 255, 255, 255), (18, 64, 200, 3),
          (3, 1, 1, 1), (9, 200, 254, 128)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden",
                                                  "c2_fp64_digest.json"))
    a = ap.parse_args()
    set_threads(a.threads)
    o = OracleCase(3)
    cases.cavity3d(o, a.n, re=400.0, ulid=0.05, noise=True)
    t0 = time.time()
    done = 0
    while done < a.steps:
        k = min(100, a.steps - done)
        o.step(k)
        done += k
        print("oracle C2 step %d (%.0f s)" % (done, time.time() - t0), flush=True)
    f = o.get_f(0)
    out = {"config": "3D cavity D3Q19 %d^3 Re 400 lid 0.05 noise seed 0x4D504C42" % a.n,
           "n": a.n, "steps": a.steps, "sha256_f": digest(f),
           "mass": "%.17g" % o.mass(),
           "probes": [[list(p), "%.17g" % f[p]] for p in PROBES if max(p[1:]) < a.n]}
    with open(a.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
