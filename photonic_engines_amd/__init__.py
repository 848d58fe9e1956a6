"""MI355X-native SLOS strong-simulation engine: drop-in for FockSLOSKernel.probs(U, loss=...) of
Quandela/photonic-engines, computed by hand-written gfx950 HIP kernels behind a C ABI
(include/slos_mi355x.h)."""
from .masks import Mask, graded_state_space_size, is_valid, state_space_size

__version__ = "0.1.0"

__all__ = ["Mask", "graded_state_space_size", "is_valid", "state_space_size", "__version__"]
