"""MI355X-native force-model graph-embedding hot path (DistEmbed's
EmbeddingAlgo::algo_force2_vec_ns) behind a C ABI: HIP kernels for gfx950,
C++ host code, and this thin Python view used by tests and benchmarks."""
from ._lib import FeError, load  # noqa: F401

__all__ = ["FeError", "load"]
