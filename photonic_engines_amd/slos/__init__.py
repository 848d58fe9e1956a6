"""SLOS_full strong simulation (Heurtel, Mansfield, Senellart, Valiron, arXiv:2206.10549) on MI355X.

  basis   lex-ranked layers (rank / unrank instead of stored states), get_layers
  kernel  FockSLOSKernel: the layer recurrence as gfx950 HIP kernels, fused readout
  autograd hand-written adjoint of probs(U) with every-other-layer checkpointing
  haar    seeded Haar-random unitaries (the test and bench input generator)
  batched UnevenBatch: kernels of different shapes in one launch per layer, autotuned, hipGraph
"""
from .basis import FockLayers, RankedLayers, clear_layers_cache, get_layers
from .haar import haar_unitary
from .kernel import FockSLOSKernel
from .batched import UnevenBatch

__all__ = ["FockLayers", "FockSLOSKernel", "RankedLayers", "clear_layers_cache", "get_layers", "haar_unitary",
           "UnevenBatch"]
