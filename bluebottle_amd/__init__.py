"""bluebottle_amd -- MI355X-native per-timestep hot path of Bluebottle.

The product is the C-ABI library ``lib/libbluebottle_hip.so`` (HIP kernels
for gfx950 behind the reference's ``cuda_*`` entry points).  This package
holds its Python-side plumbing: ctypes mirrors of the ABI structs
(``abi``), library loading and global-state access (``runtime``), case
set-up and stepping (``sim``), and the benchmark configurations
(``configs``).
"""
from . import abi, runtime  # noqa: F401

__all__ = ["abi", "runtime", "sim", "configs"]
