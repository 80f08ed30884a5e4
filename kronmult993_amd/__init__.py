"""kronmult993_amd -- MI355X-native batched Kronecker mat-vec (`kronmult_batched`).

The product is libkronmult_hip.so (hand-written HIP for gfx950 behind the C ABI of include/kronmult_hip.h,
with the C++ drop-in template in include/kronmult_hip.hpp). This package is its Python host side:
  * `api`       -- the reference's `kronmult_batched` interface over the C ABI (ctypes);
  * `workloads` -- seeded synthetic batches (the five benchmark workloads, the reference's test cases);
  * `device`    -- a batch placed in device memory with its pointer arrays (torch owns the memory);
  * `gen`       -- the counter-based random generator.
"""
from . import gen, workloads  # noqa: F401

__all__ = ["gen", "workloads", "api", "device"]
