"""dompap_amd -- MI355X-native engine for dompap's per-step hot path.

``from dompap_amd import Simulation`` is a drop-in for ``from dompap import
Simulation``: neighbour-list build, pair forces / energy / virial /
Laplacian and the G-JF Langevin / leap-frog update run as hand-written HIP
kernels for gfx950 (libdompap_hip.so, C-ABI in include/dompap_hip.h); any
SymPy pair potential is compiled to device code with hiprtc.
"""
from .Simulation import Simulation  # noqa: F401
from .batch import ReplicaBatch  # noqa: F401

__version__ = "0.1.0"
