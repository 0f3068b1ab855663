"""zeroslam_amd — MI355X bundle adjustment for ZeroSLAM-style factor graphs.

The compute path is ``libzsba.so`` (HIP kernels for gfx950 behind the C ABI in
``include/zsba.h``); this package holds its build (``build``), the ctypes
binding (``engine``), the synthetic workloads (``problem``) and a
``torch.autograd`` wrapper of the adjoint (``autograd``).  The reference's
factor-graph classes are mirrored in C++ by ``include/zsba_factor_graph.hpp``.
"""
__all__ = ["build", "problem", "engine"]
