"""SLOS_full strong simulation (Heurtel, Mansfield, Senellart, Valiron, arXiv:2206.10549) on MI355X.

  basis   lex-ranked layers (rank / unrank instead of stored states), get_layers
  kernel  FockSLOSKernel: the layer recurrence as gfx950 HIP kernels, fused readout
  autograd hand-written adjoint of probs(U) with every-other-layer checkpointing
  haar    seeded Haar-random unitaries (the test and bench input generator)
  batched UnevenBatch: kernels of different shapes in one launch per layer, autotuned, hipGraph
  species partial distinguishability (x-model): probs(indist=x, order=b)
  sampling strong-then-sample on the device: sample_from_probs, draw_indices
"""
from .basis import FockLayers, RankedLayers, clear_layers_cache, get_layers
from .haar import haar_unitary
from .kernel import FockSLOSKernel, GroupedSLOSKernel, SuperpositionSLOSKernel
from .batched import UnevenBatch
from .sampling import sample_from_probs

__all__ = ["FockLayers", "FockSLOSKernel", "RankedLayers", "clear_layers_cache", "get_layers", "haar_unitary",
           "UnevenBatch", "sample_from_probs",
           "GroupedSLOSKernel", "SuperpositionSLOSKernel"]
