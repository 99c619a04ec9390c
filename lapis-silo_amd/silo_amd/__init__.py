"""MI355X (gfx950) engine for SILO/RhyDB's filter, Mutations and Aggregated query path.

The compute runs in `csrc/libsilo_amd.so` (HIP kernels behind the C ABI of include/rdb.h); this
package holds the Python binding, the workload generator binding and the ingest description.
"""
from . import datagen, table  # noqa: F401

__all__ = ["datagen", "table", "engine", "workloads"]
