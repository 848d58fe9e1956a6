"""Strong-simulation backend face of the SLOS engine (photonic_engines/adapters/perceval/backends.py:
`_TorchStrongBase` :54-180 and `SLOSTorchBackend` :185-244).

`StrongBackend` holds the backend's logic with plain Python types: states are occupation tuples,
distributions are dicts. It mirrors the reference backend's methods (set_circuit, set_input_state,
prob_amplitude, probability, all_prob, prob_distribution, samples, evolve), errors and
normalisation (<t|U|s> = per / sqrt(prod s_i! prod t_j!)). The output-state index is a rank lookup
(`RankedLayers.index_of`): no final_states dict of the whole basis is built, which at
n = 11, m = 32 would hold 4.3e9 entries (the reference's `_idx`, backends.py:105-111).

`SLOSTorchBackend` is the Perceval plugin: the same logic behind Perceval's
AStrongSimulationBackend. It is defined only when `perceval` can be imported. Its Perceval types
(BasicState, StateVector, BSDistribution, allstate_iterator) are used exactly as in the reference.
"""
from __future__ import annotations

import math
from collections.abc import Sequence

import numpy as np
import torch

from ..imperfections import spec


def _unitary(circuit) -> np.ndarray:
    """A circuit with compute_unitary() (a Perceval circuit) or a matrix."""
    u = circuit.compute_unitary() if hasattr(circuit, "compute_unitary") else circuit
    return np.asarray(u, dtype=np.complex128)


class StrongBackend:
    """Strong simulation of one Fock input or a coherent superposition of Fock inputs (one kernel per
    photon number), with optional uniform loss, x-model partial distinguishability, or a species
    partition (photons of one species interfere; `species=(groups, singles)` as input columns)."""

    name = "SLOSTorch"

    def __init__(self, device: str = "cuda", dtype: torch.dtype = torch.complex64, loss: float | None = None,
                 indist=None, order: int | None = None, mask: str = "fock"):
        self.device, self.dtype, self.mask = device, dtype, mask
        self.loss = None if loss is None or float(loss) >= 1.0 else float(loss)
        self.indist = spec(indist)
        self.order = order
        self._U: torch.Tensor | None = None
        self._m: int | None = None
        self._kernels: dict[int, object] = {}
        self._species = None
        self._amps: dict | None = None
        self._probs: dict | None = None
        self._lossy = None

    # ---- configuration -------------------------------------------------------------------------------
    def set_circuit(self, circuit) -> None:
        u = _unitary(circuit)
        self._m = u.shape[0]
        self._U = torch.as_tensor(u).to(self.dtype)
        self._invalidate()

    def set_input_state(self, state, species=None) -> None:
        """`state`: an occupation sequence (one Fock input), or (n, m, occupations, amplitudes) as an
        object with those attributes (a coherent superposition, one kernel per photon number)."""
        from ..slos import FockSLOSKernel, SuperpositionSLOSKernel
        self._kernels, self._species = {}, None
        if hasattr(state, "occupations") and hasattr(state, "amplitudes"):
            if self.indist is not None:
                raise NotImplementedError("partial distinguishability of a coherent superposition: no per-photon identity")
            occs = [list(map(int, o)) for o in state.occupations]
            amps = np.asarray(state.amplitudes, dtype=np.complex128)
            amps = amps / math.sqrt(float(np.sum(np.abs(amps) ** 2)))
            blocks: dict = {}
            for o, a in zip(occs, amps):
                blocks.setdefault(sum(o), ([], []))
                blocks[sum(o)][0].append(o)
                blocks[sum(o)][1].append(a)
            for n, (oc, am) in blocks.items():
                self._kernels[n] = SuperpositionSLOSKernel(n, len(oc[0]), oc, mask=self.mask, device=self.device,
                                                          dtype=self.dtype, weights=am)
        else:
            occ = [int(c) for c in state]
            if species is not None and self.indist is not None:
                raise NotImplementedError("annotated input and indist= together: annotate, or use the x-model, not both")
            self._species = species
            self._kernels[sum(occ)] = FockSLOSKernel(sum(occ), len(occ), occ, mask=self.mask, device=self.device,
                                                     dtype=self.dtype)
        self._invalidate()

    def _invalidate(self) -> None:
        self._amps = self._probs = self._lossy = None

    # ---- computation -----------------------------------------------------------------------------------
    def _ensure(self) -> None:
        if self._probs is not None:
            return
        if self._U is None or not self._kernels:
            raise RuntimeError("circuit and input state must be set")
        self._amps, self._probs = {}, {}
        with torch.no_grad():
            for n, k in self._kernels.items():
                if self._species is not None:
                    from ..slos.species import species_probs
                    Ub, _ = k._prepare(self._U)
                    self._probs[n] = species_probs(k, Ub, *self._species)[0]
                    self._amps = None
                elif self.indist is not None:
                    self._probs[n] = k.probs(self._U, indist=self.indist, order=self.order)
                    self._amps = None
                else:
                    # probabilities from the fused readout; amplitudes only when asked for (_amplitudes)
                    self._probs[n] = k.probs(self._U)
        if self.loss is not None:
            from ..imperfections import thin_distribution
            self._lossy = {n: thin_distribution(k.layers, self._probs[n], self.loss)
                           for n, k in self._kernels.items()}

    def _amplitudes(self, n: int) -> torch.Tensor:
        """Raw amplitudes times the input norm of block n ([S] of the backend's dtype), computed on first use."""
        if n not in self._amps:
            self._amps[n] = self._kernels[n].amplitudes(self._U)
        return self._amps[n]

    @staticmethod
    def _out_norm(t) -> float:
        """sqrt(prod t_j!): Perceval's normalisation of a raw amplitude of output t."""
        return math.sqrt(math.prod(math.factorial(int(c)) for c in t))

    def _idx(self, t):
        """(photon-number block, lex index in that block's final basis): a rank lookup."""
        t = tuple(int(c) for c in t)
        n = sum(t)
        k = self._kernels.get(n)
        if k is None:
            return n, None
        return n, k.layers.index_of(t)

    # ---- the backend API ---------------------------------------------------------------------------------
    def prob_amplitude(self, output_state) -> complex:
        self._ensure()
        if self.loss is not None:
            raise NotImplementedError("amplitudes are undefined with loss (mixed state)")
        if self.indist is not None or self._species is not None:
            raise NotImplementedError("amplitudes are undefined with partial distinguishability (mixed state)")
        n, i = self._idx(output_state)
        return 0j if i is None else complex(self._amplitudes(n)[i]) * self._out_norm(output_state)

    def probability(self, output_state) -> float:
        self._ensure()
        t = tuple(int(c) for c in output_state)
        if self.loss is not None:
            tot = 0.0
            for n, ld in self._lossy.items():
                k = sum(t)
                if k <= n:
                    i = self._kernels[n].layers.index_of(t)
                    tot += float(ld.probs[k][i])
            return tot
        n, i = self._idx(t)
        return 0.0 if i is None else float(self._probs[n][i])

    def all_prob(self, input_state=None) -> list[float]:
        """Probabilities of every output state of the input's photon number, in the kernel's basis
        order (`layers.final_states`)."""
        if input_state is not None:
            self.set_input_state(input_state)
        self._ensure()
        if self.loss is not None:
            raise NotImplementedError("all_prob runs over the n-photon states only; with loss use prob_distribution()")
        return [float(v) for p in self._probs.values() for v in p.cpu().tolist()]

    def prob_distribution(self) -> dict:
        self._ensure()
        out: dict = {}
        if self._lossy is not None:
            for ld in self._lossy.values():
                for t, p in ld.as_dict().items():
                    out[t] = out.get(t, 0.0) + p
            return out
        for n, k in self._kernels.items():
            p = self._probs[n].cpu().tolist()
            for t, v in zip(k.layers.final_states, p):
                out[t] = out.get(t, 0.0) + v
        return out

    def samples(self, count: int, generator: torch.Generator | None = None) -> list[tuple]:
        """Occupation tuples drawn from the computed distribution (lossy when `loss` is set)."""
        self._ensure()
        from ..slos.sampling import sample_from_probs
        if self._lossy is None and len(self._kernels) == 1:
            (n, k), = self._kernels.items()
            rows = sample_from_probs(k.layers, self._probs[n], int(count), generator=generator)
            return [tuple(r) for r in rows.cpu().tolist()]
        dist = self.prob_distribution()
        states = list(dist)
        probs = torch.tensor([dist[s] for s in states], dtype=torch.float64)
        idx = torch.multinomial(probs / probs.sum(), int(count), replacement=True, generator=generator)
        return [states[i] for i in idx.tolist()]

    def evolve(self) -> dict:
        """The output state as {occupation tuple: amplitude} (pure outputs only)."""
        self._ensure()
        if self.loss is not None:
            raise NotImplementedError("evolve needs a pure output (no loss)")
        if self.indist is not None or self._species is not None:
            raise NotImplementedError("evolve needs a pure output (no partial distinguishability)")
        out: dict = {}
        for n, k in self._kernels.items():
            for t, a in zip(k.layers.final_states, self._amplitudes(n).cpu().tolist()):
                out[t] = complex(a) * self._out_norm(t)
        return out


try:  # the Perceval plugin, when Perceval is installed
    import perceval as _pcvl  # noqa: F401
    from perceval.backends import AStrongSimulationBackend as _ASSB
    from perceval.utils import BasicState as _BS
    from perceval.utils import BSDistribution as _BSD
    from perceval.utils import StateVector as _SV
    from perceval.utils import allstate_iterator as _all

    class SLOSTorchBackend(_ASSB):
        """AStrongSimulationBackend on the MI355X SLOS engine (backends.py:185-244)."""

        name = "SLOSTorch"

        def __init__(self, device: str = "cuda", dtype: torch.dtype = torch.complex64, loss=None, indist=None,
                     order: int | None = None):
            super().__init__()
            self._core = StrongBackend(device=device, dtype=dtype, loss=loss, indist=indist, order=order)

        def set_circuit(self, circuit):
            super().set_circuit(circuit)
            self._core.set_circuit(np.asarray(self._umat))

        def set_mask(self, *a, **k):
            raise NotImplementedError("masks: use the engine's mask= argument directly")

        def set_input_state(self, input_state):
            super().set_input_state(input_state)
            if isinstance(input_state, _SV):
                class _Sup:
                    pass
                sup = _Sup()
                keys = list(input_state.keys())
                sup.occupations = [list(s) for s in keys]
                sup.amplitudes = [complex(input_state[s]) for s in keys]
                self._core.set_input_state(sup)
            else:
                self._core.set_input_state([int(input_state[i]) for i in range(input_state.m)])

        def prob_amplitude(self, output_state) -> complex:
            return self._core.prob_amplitude(list(output_state))

        def probability(self, output_state) -> float:
            return self._core.probability(list(output_state))

        def all_prob(self, input_state=None):
            if input_state is not None:
                self.set_input_state(input_state)
            self._core._ensure()
            return [self._core.probability(list(s)) for s in _all(self._input_state)]

        def prob_distribution(self):
            bsd = _BSD()
            for t, p in self._core.prob_distribution().items():
                bsd.add(_BS(list(t)), p)
            return bsd

        def evolve(self):
            res = _SV()
            for t, a in self._core.evolve().items():
                res += complex(a) * _BS(list(t))
            res.normalize()
            return res
except ImportError:
    SLOSTorchBackend = None
