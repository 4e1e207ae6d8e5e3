"""gmpt_amd: MI355X-native GPU Merkle Patricia Trie batch insert + root hash.

The product is libgmpt_amd.so (HIP kernels for gfx950 behind the C-ABI in
include/gmpt_amd.h and include/libgmpt.h); `gmpt` is its ctypes binding and
`workloads` builds batches in the reference's packed layout.
"""
from . import workloads  # noqa: F401
from . import gmpt  # noqa: F401
from .gmpt import MPT, DeviceBatch, GmptError, Server, build_root_noalloc, keccak256_batch  # noqa: F401
