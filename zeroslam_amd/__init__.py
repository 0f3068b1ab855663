"""zeroslam_amd — MI355X bundle adjustment for ZeroSLAM-style factor graphs.

The compute path is ``libzsba.so`` (HIP kernels for gfx950 behind the C ABI in
``include/zsba.h``); this package holds its build, a ctypes binding, the Python
mirror of the reference's factor-graph API and the synthetic workloads.
"""
__all__ = ["build", "problem", "factor_graph", "engine"]
