"""mplbhip_mass_noalloc, the entry point that allocates no device memory (its
partial sums live in the library's file-scope scratch array), against
mplbhip_mass on the same state: two calls in a row on two cases, the second
smaller, each bit-equal to the allocating entry point; plus its refusals."""
import sys

import numpy as np
import pytest

from mplb_amd import cases, hip
from mplb_amd.enums import BoundaryScheme as BSch
from mplb_amd.enums import VertexType as VT


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_noalloc_refuses_cases_beyond_its_scratch():
    """A plan-only context of 1024 x 1024 x 32 nodes needs 16 x 256 x 32
    workgroups, more than the 8 MiB scratch holds: refused with a message,
    before any device is needed."""
    g = hip.LatticeCase(3, storage=hip.STORE_F32, device=-1)
    g.add_block(0, (1024, 1024, 32))
    g.set_collision(1.0)
    for s in ("Left", "Right", "Top", "Bottom", "Front", "Back"):
        g.add_boundary(0, cases.BS[s], BSch.EQMDiffuseRefl, (0.0, 0.0, 0.0), VT.Wall)
    g.prepare()
    with pytest.raises(hip.MplbError, match="doubles of scratch"):
        g.mass_noalloc()


@pytest.mark.gpu
def test_noalloc_mass_matches_mass_on_two_cases():
    log("no-alloc mass: 3D cavity 40^3 FP32, then 2D cavity 24^2 FP64")
    g = hip.LatticeCase(3, storage=hip.STORE_F32, device=0)
    cases.cavity3d(g, 40, re=300.0, noise=True)
    with pytest.raises(hip.MplbError, match="no step taken"):
        g.mass_noalloc()
    g.step(12)
    a, b = g.mass_noalloc(), g.mass()
    log("  40^3: noalloc %.17g, mass %.17g" % (a, b))
    assert a == b and np.isfinite(a)
    g.step(5)
    assert g.mass_noalloc() == g.mass()
    g.close()
    # a second, smaller case: the shared scratch keeps the first case's
    # values, which the launches overwrite
    h = hip.LatticeCase(2, storage=hip.STORE_F64, device=0)
    cases.cavity2d(h, 24, re=100.0)
    h.step(7)
    a, b = h.mass_noalloc(), h.mass()
    log("  24^2: noalloc %.17g, mass %.17g" % (a, b))
    assert a == b
    h.close()


@pytest.mark.gpu
def test_noalloc_mass_matches_mass_in_the_split_step_form(monkeypatch):
    """Cases of 2^18 nodes and more step (and measure their mass) through
    the bulk and boundary kernels: forced here on a 48^3 FP16 cavity, then a
    smaller 36^3 FP64 one, the two entry points agree bit for bit."""
    monkeypatch.setenv("MPLBHIP_FUSED", "0")
    for n, st in ((48, hip.STORE_F16), (36, hip.STORE_F64)):
        log("no-alloc mass, split step form: 3D cavity %d^3 storage %d" % (n, st))
        g = hip.LatticeCase(3, storage=st, device=0)
        cases.cavity3d(g, n, re=300.0, noise=True)
        g.step(9)
        a, b = g.mass_noalloc(), g.mass()
        log("  %d^3: noalloc %.17g, mass %.17g" % (n, a, b))
        assert a == b and np.isfinite(a)
        g.close()
