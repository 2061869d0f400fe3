"""MI355X force engine for cuco's PM+P3M gravity path.

The engine itself is ``libcuco_engine.so`` (hand-written HIP for gfx950
behind the C ABI in ``include/cuco_engine.h``); this package holds its
ctypes binding (:mod:`cuco_amd.engine`), the synthetic workload generators
(:mod:`cuco_amd.ics`) and the build recipe (:mod:`cuco_amd.build`).
"""
from . import ics  # noqa: F401
