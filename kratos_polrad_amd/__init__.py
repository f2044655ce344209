"""MI355X-native photon-packet Monte Carlo transport engine.

Drop-in for the photon module of kratos_polrad
(src/modules/particle/radiation/): hand-written gfx950 HIP kernels behind the
C ABI of include/kpolrad.h, with the host mesh layer (block tree, face
neighbour table, geometry) in C++.  This package is the thin Python view of
that ABI used by the tests and the benchmark.
"""
from ._abi import KPError, lib  # noqa: F401
from .engine import BandSpec, Engine, RunSpec  # noqa: F401
from .mesh import Mesh  # noqa: F401
