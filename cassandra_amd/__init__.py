"""cassandra_amd: the pattern-evaluator fit of dominichofer/Cassandra on
MI355X (gfx950) -- HIP kernels behind a C-ABI (include/cassandra_amd.h).

Importing the package does not touch the GPU; the HIP library is loaded on the
first call (cassandra_amd.capi.lib()) and its absence is an error.
"""
from . import capi, names  # noqa: F401

__all__ = ["capi", "names"]
