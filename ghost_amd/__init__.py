"""ghost_amd: GHOST's pseudospectral hot path on AMD MI355X (gfx950).

The compute lives in ``libghx.so`` (hand-written HIP kernels behind the C ABI
of include/ghx.h); this package is the Python host mirror of GHOST's
``fftp`` and pseudospec interfaces, used by the tests and the benchmark.
"""
from ._lib import GhxError, LIB_PATH, declared_symbols, load  # noqa: F401

__all__ = ["GhxError", "LIB_PATH", "declared_symbols", "load"]
