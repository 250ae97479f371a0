"""oxav1 — Python plumbing for the gfx950 AV1 post-filter chain.

The product is the C-ABI library liboxav1_pf.so (HIP kernels for gfx950);
this package only binds it with ctypes for tests, benchmarks and tools:
  abi      struct layouts and function signatures (checked against the C side)
  frames   seeded synthetic frames and drop-in argument bundles
  dropin   Python-level calls of the five drop-in entry points
  engine   the device-resident chain engine and HBM buffers (library allocator)
"""
from . import abi  # noqa: F401
from .frames import Frame, planes_struct, workload_seeds  # noqa: F401
