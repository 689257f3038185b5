"""mplb_amd: MI355X-native lattice update for MPLB (HIP, gfx950).

The compute path lives in the C-ABI library built from mplb_amd/csrc
(libmplbhip.so); mplb_amd.hip binds it.  There is no CPU fallback: using the
lattice API without the built library raises.
"""
from . import enums, lattice  # noqa: F401
