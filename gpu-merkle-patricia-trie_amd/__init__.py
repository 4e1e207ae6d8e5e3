"""gmpt_amd: MI355X-native GPU Merkle Patricia Trie batch insert + root hash."""
from . import workloads  # noqa: F401
