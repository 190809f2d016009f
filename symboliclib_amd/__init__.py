"""MI355X execution backend for sparsity-specific generated expression kernels.

The C++ library (csrc/, built into lib/libsymlib.so) keeps the reference's
Sym::ComputeUnit<double>(Device(UseHIP(), ThreadsPerBlock(n)), AS, BS) surface; this package
holds its Python bindings (unit.Unit), the benchmark workload generators (workloads) and the
build helper (build.build_native).
"""
from . import workloads  # noqa: F401

__all__ = ["workloads", "unit", "build"]
