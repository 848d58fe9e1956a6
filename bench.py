"""Headline benchmark: SLOS strong simulation, n=11 photons in m=32 modes, complex64, one GPU.

    python bench.py [--steps K] [--warmup W] [--no-cpu-baseline] [--no-batched] [--workload c5|c2]

A step is one FockSLOSKernel.probs(U) call (11 photon layers, the last one fused with the
|amp|^2 * prod t! readout, float32 output) on one seeded Haar unitary already resident in HBM: the
drop-in op itself, timed with perf_counter around `steps` calls between two device
synchronisations. `value` = output probabilities per second (S_11 / step time). The dominant
kernel is the last layer: after the timed steps the same launch (same inputs) runs 5 times
bracketed by HIP events on the stream it runs on, and `roofline.achieved` = its algorithmic bytes
(read layer 10 once: 8 S_10, write the probabilities: 4 S_11) / its mean duration. Accuracy of
the timed output: max|dp| and max relative error over 10,000 seeded outputs against FP64 Ryser
permanents computed on the GPU, and the FP64 total mass.
`cpu_baseline`: the oracle's C++/OpenMP pull-form restatement of the same workload on this host's
cores, every layer complete (no extrapolation). `cpu_reference_algorithm`: the reference's own
stored-edge algorithm restated in torch on the CPU at n=8, m=24 (its edges do not fit at n=11,
m=32), beside the GPU engine on the same input.
`variants`: the compensated-accumulation forward of the same workload; the lossy output
(probs(U, loss=0.9): the forward with the top binomial weight in its readout and the 11 descent
steps, with the k=11 descent's time and roofline and the FP64 mass of every photon-number block
against the binomial law); the gradient step of n=6, m=16 (forward, L = -mean log p over 1,024
outputs drawn from the exact distribution, dL/dU through the hand-written adjoint) and the same
at n=8, m=24 with the time and roofline of its top (block-structured) adjoint layer; and the
uneven batch of 512 unitaries ((4,12), (6,16), (8,20)) as a single-U loop, per-shape batches, one
group launch per layer, and the autotuned group forward replayed as a hipGraph; and the components
built on the path (`next_rows`): device sampling of 10^6 patterns from the n=11, m=32
distribution, the x-model and a 4-component coherent superposition at n=8, m=24, grouped
probabilities, and rank lookups of the strong-simulation backend at n=11, m=32; and
(`published_shapes`) probs(U) on every shape the reference publishes a time for (its torch kernel
on an H100, its C engine on 48 CPU threads), timed the way its product bench times it, with that
time and its source file beside ours.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SLOS probs/s + % HBM roofline (n=11,m=32,c64) at max|Δp|≤1e-6 vs FP64"
PEAK_HBM_GBS = 8000.0
WORKLOADS = {
    "c5": dict(n=11, m=32, seed=11032),
    "c2": dict(n=8, m=24, seed=8024),
}


def bytes_alg(n: int, m: int) -> int:
    """Compulsory HBM bytes of one forward: every layer read once and written once (complex64),
    the last written as float32: sum_{k=1..n} 8 S_{k-1} + sum_{k=1..n-1} 8 S_k + 4 S_n."""
    S = [math.comb(m + k - 1, k) for k in range(n + 1)]
    return sum(8 * S[k - 1] for k in range(1, n + 1)) + sum(8 * S[k] for k in range(1, n)) + 4 * S[n]


def load_traffic(tag: str):
    """HBM bytes per launch of the last-layer kernel, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    and scaled by a plain-copy calibration; used only if the kernel sources are the ones measured."""
    path = os.path.join(ROOT, "profiles", f"layer_block_{tag}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh)
    src = b"".join(open(os.path.join(ROOT, "photonic_engines_amd", "csrc", p), "rb").read()
                   for p in ("layer_block.hpp", "block_index.hpp"))
    return rec if hashlib.sha256(src).hexdigest() == rec.get("source_sha256") else None


UNEVEN = [(4, 12, 171), (6, 16, 171), (8, 20, 170)]  # (n, m, unitaries): 512 in all, photons in modes 0..n-1


def bench_uneven(reps: int = 5) -> dict:
    """512 seeded Haar unitaries, 1/3 each of (n, m) = (4,12), (6,16), (8,20), complex64, lossless:
    throughput of the plain single-U path (one probs(U) call per unitary), of one batched call per
    shape, of one group launch per layer (untuned), and of the autotuned group forward replayed as
    one hipGraph. All four produce bit-identical outputs; accuracy is stated once, over 32 seeded
    outputs of every instance against FP64 Ryser permanents, plus every instance's FP64 mass."""
    from photonic_engines_amd.slos import FockSLOSKernel, UnevenBatch, haar_unitary
    dev = torch.device("cuda", 0)
    ks = [FockSLOSKernel(n, m, [1] * n + [0] * (m - n), device=str(dev)) for n, m, _ in UNEVEN]
    Us = [torch.stack([haar_unitary(m, 100_000 * (g + 1) + i) for i in range(c)]).to(torch.complex64).to(dev)
          for g, (n, m, c) in enumerate(UNEVEN)]
    total = sum(k.output_size() * c for k, (_, _, c) in zip(ks, UNEVEN))
    # compulsory HBM bytes of the whole workload: every layer of every instance read once and written
    # once (complex64), the last written as float32
    alg = sum(bytes_alg(n, m) * c for n, m, c in UNEVEN)

    def rate(t: float) -> dict:
        return {"ms": t * 1e3, "probs_per_s": total / t, "alg_GBps": alg / t / 1e9, "frac": alg / t / 1e9 / PEAK_HBM_GBS}

    def clock(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    res = {"workload": "512 Haar unitaries: 171 x (4,12), 171 x (6,16), 170 x (8,20), photons in modes 0..n-1, "
                       "complex64, float32 probabilities", "outputs": total, "alg_bytes": alg}
    t = clock(lambda: [k.probs(U[i]) for k, U in zip(ks, Us) for i in range(U.shape[0])])
    res["single_u_loop"] = rate(t)
    t = clock(lambda: [k.probs(U) for k, U in zip(ks, Us)])
    res["batched_per_shape"] = rate(t)
    batch = UnevenBatch(ks)
    t = clock(lambda: batch.probs(Us))
    res["group_untuned"] = rate(t)
    cfg = batch.tune(Us, reps=2)
    batch.capture(Us)
    t = clock(lambda: batch.replay())
    res["group_tuned_graph"] = dict(rate(t), tuned=[f"L{c['layer']}:{c['kernel']}/{c['per_wave']}/{c['grid']}"
                                                    for c in cfg])
    outs = batch.replay()
    torch.cuda.synchronize()
    rng = np.random.default_rng(512)
    worst, mass_dev = 0.0, 0.0
    for k, U, P in zip(ks, Us, outs):
        mass_dev = max(mass_dev, float((k.total_mass(P) - 1).abs().max()))
        U128 = U.cpu().to(torch.complex128)
        for i in range(U.shape[0]):
            idx = torch.as_tensor(np.sort(rng.choice(P.shape[1], 32, replace=False)).astype(np.int64), device=dev)
            ref = k.spot_probs(U128[i], idx).cpu().numpy()
            worst = max(worst, float(np.abs(P[i, idx].double().cpu().numpy() - ref).max()))
    res["accuracy"] = {"max_abs_dp": worst, "spot_outputs_per_instance": 32, "oracle": "FP64 Ryser on the GPU",
                       "max_abs_mass_minus_1": mass_dev}
    res["speedup_graph_vs_single_u"] = res["single_u_loop"]["ms"] / res["group_tuned_graph"]["ms"]
    return res


def bench_lossy(kern, U: torch.Tensor, eta: float = 0.9, steps: int = 3) -> dict:
    """probs(U, loss=eta) end to end, and the top descent step alone (HIP events on its stream)."""
    from photonic_engines_amd.imperfections import binomial_pmf
    n, plan, S = kern.n, kern._plan, kern._plan.sizes
    ld = kern.probs(U, loss=eta)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        del ld
        ld = kern.probs(U, loss=eta)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    pmf = binomial_pmf(n, eta).tolist()
    masses = [float(kern.total_mass(q)) for q in ld.probs]
    top = ld.probs[n].unsqueeze(0)
    del ld
    stream = torch.cuda.current_stream()
    ms = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        q = plan.loss_descend(n, top)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
        del q
    d_ms = float(np.median(ms))
    d_bytes = 4 * S[n] + 4 * S[n - 1]
    outputs = int(sum(S[: n + 1]))
    return {"eta": eta, "ms_per_step": dt * 1e3, "outputs": outputs, "probs_per_s": outputs / dt,
            "descent_top": {"kernel": f"k_descend_block (layer {n} -> {n - 1})", "kernel_ms": d_ms,
                            "alg_bytes": d_bytes, "achieved_GBps": d_bytes / (d_ms * 1e-3) / 1e9,
                            "frac": d_bytes / (d_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
            "max_abs_block_mass_minus_binomial": max(abs(a - b) for a, b in zip(masses, pmf))}


def bench_gradient(steps: int = 20) -> dict:
    """n=6, m=16: one step = probs(U) with U requiring a gradient, L = -mean log p[idx] over 1,024
    indices drawn (seeded) from the exact distribution, L.backward() (adjoint kernels)."""
    from photonic_engines_amd.slos import FockSLOSKernel, haar_unitary
    n, m = 6, 16
    k = FockSLOSKernel(n, m, [1] * n + [0] * (m - n), device="cuda")
    U0 = haar_unitary(m, 6016).to(torch.complex64).cuda()
    with torch.no_grad():
        p = k.probs(U0)
    idx = torch.multinomial(p.double(), 1024, replacement=True, generator=torch.Generator(device="cuda").manual_seed(4))

    def step():
        U = U0.clone().requires_grad_(True)
        loss = -torch.log(k.probs(U)[idx]).mean()
        loss.backward()
        return U.grad

    for _ in range(3):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"workload": "n=6 m=16, first 6 modes, L = -mean log p over 1,024 sampled outputs", "states": int(p.numel()),
            "ms_per_step": dt * 1e3, "grad_nonzero_columns": int((g.abs().sum(0) > 0).sum())}


def bench_adjoint(steps: int = 5) -> dict:
    """n=8, m=24 (7.9M outputs): one gradient step (probs(U) with U requiring a gradient, L = -mean
    log p over 1,024 outputs drawn from the exact distribution, L.backward() through the adjoint
    kernels), and the top adjoint layer (layer 8 -> 7, the block-structured adjoint) alone,
    bracketed by HIP events on its stream. Its algorithmic bytes: read g_8 and psi_7, write g_7:
    8 S_8 + 16 S_7."""
    from photonic_engines_amd import _native as nat
    from photonic_engines_amd.slos import FockSLOSKernel, haar_unitary
    n, m = 8, 24
    k = FockSLOSKernel(n, m, [1] * n + [0] * (m - n), device="cuda")
    U0 = haar_unitary(m, 8024).to(torch.complex64).cuda()
    with torch.no_grad():
        p = k.probs(U0)
    idx = torch.multinomial(p.double(), 1024, replacement=True, generator=torch.Generator(device="cuda").manual_seed(8))

    def step():
        U = U0.clone().requires_grad_(True)
        loss = -torch.log(k.probs(U)[idx]).mean()
        loss.backward()
        return U.grad

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    plan, modes = k._plan, k.input_modes
    S = plan.sizes
    U = U0.unsqueeze(0).contiguous()
    psi = torch.ones((1, 1), dtype=torch.complex64, device="cuda")
    for kk in range(1, n):
        psi = plan.layer_step(U, kk, modes[kk - 1], psi, nat.OUT_RAW)
    gk = torch.randn((1, S[n]), dtype=torch.complex64, device="cuda")
    gU = torch.zeros_like(U)
    stream = torch.cuda.current_stream()
    plan.adjoint_layer(U, n, modes[n - 1], psi, gk, gU)
    ms = []
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        plan.adjoint_layer(U, n, modes[n - 1], psi, gk, gU)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    a_ms = float(np.median(ms))
    a_bytes = 8 * S[n] + 16 * S[n - 1]
    return {"workload": "n=8 m=24, first 8 modes, L = -mean log p over 1,024 sampled outputs", "states": int(S[n]),
            "ms_per_step": dt * 1e3, "grad_nonzero_columns": int((g.abs().sum(0) > 0).sum()),
            "adjoint_top": {"kernel": "k_adjoint_block (layer 8 -> 7, with dL/dU partials and their FP64 sum)",
                            "kernel_ms": a_ms, "alg_bytes": a_bytes,
                            "achieved_GBps": a_bytes / (a_ms * 1e-3) / 1e9,
                            "frac": a_bytes / (a_ms * 1e-3) / 1e9 / PEAK_HBM_GBS}}


def bench_next_rows(kern, U: torch.Tensor) -> dict:
    """Timings of the components built on the path (sampling, x-model, superposition, grouped
    probabilities, backend lookups); their parity is in the GPU tests."""
    from photonic_engines_amd.adapters.perceval import StrongBackend
    from photonic_engines_amd.slos import FockSLOSKernel, SuperpositionSLOSKernel, haar_unitary
    from photonic_engines_amd.slos.sampling import draw_indices, unrank_rows

    def clock(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, r

    out = {}
    n, S = kern.n, kern._plan.sizes[kern.n]
    p = kern.probs(U)
    shots = 1_000_000
    g = torch.Generator().manual_seed(17)
    dt_draw, idx = clock(lambda: draw_indices(p, shots, generator=g))
    dt_rows, _ = clock(lambda: unrank_rows(kern.layers, n, idx))
    out["sampling_n11_m32"] = {"shots": shots, "draw_ms": dt_draw * 1e3, "unrank_ms": dt_rows * 1e3,
                               "draw_read_GBps": 4 * S / dt_draw / 1e9, "frac": 4 * S / dt_draw / 1e9 / PEAK_HBM_GBS}
    del p, idx
    torch.cuda.empty_cache()
    n2, m2 = 8, 24
    U2 = haar_unitary(m2, 8024).to(torch.complex64).cuda()
    k2 = FockSLOSKernel(n2, m2, [1] * n2 + [0] * (m2 - n2), device="cuda")
    S2 = k2.output_size()
    dt_plain, _ = clock(lambda: k2.probs(U2))
    dt_x, px = clock(lambda: k2.probs(U2, indist=0.9))
    dt_x2, _ = clock(lambda: k2.probs(U2, indist=0.9, order=2))
    out["xmodel_n8_m24"] = {"x": 0.9, "lossless_ms": dt_plain * 1e3, "all_subsets_ms": dt_x * 1e3, "order2_ms": dt_x2 * 1e3,
                            "mass_fp64": float(k2.total_mass(px))}
    rng = np.random.default_rng(5)
    occs = []
    for _ in range(4):
        o = np.zeros(m2, dtype=int)
        np.add.at(o, rng.choice(m2, size=n2, replace=False), 1)
        occs.append(o.tolist())
    ks = SuperpositionSLOSKernel(n2, m2, occs, device="cuda")
    dt_s, ps = clock(lambda: ks.probs(U2))
    out["superposition_n8_m24_K4"] = {"ms": dt_s * 1e3, "probs_per_s": S2 / dt_s, "mass_fp64": float(k2.total_mass(ps))}
    occ2 = torch.as_tensor(k2.layers.occupations(n2))
    groups = occ2[:, : m2 // 2].sum(1).to(torch.int64)          # photons in the first half of the modes
    kg = k2.group_matrix(groups)
    dt_g, _ = clock(lambda: kg.probs(U2))
    out["grouped_n8_m24"] = {"groups": kg.output_size(), "ms": dt_g * 1e3}
    # dtype=torch.complex128: every layer in complex128, float64 probabilities (FP64 kernels)
    k3 = FockSLOSKernel(n2, m2, [1] * n2 + [0] * (m2 - n2), device="cuda", dtype=torch.complex128)
    U3 = haar_unitary(m2, 8024).to(torch.complex128).cuda()
    dt_z, pz = clock(lambda: k3.probs(U3))
    out["complex128_n8_m24"] = {"ms": dt_z * 1e3, "probs_per_s": S2 / dt_z, "mass_minus_1": float(pz.sum()) - 1.0}
    del pz
    torch.cuda.empty_cache()
    be = StrongBackend(device="cuda")
    be.set_circuit(U.cpu().numpy())
    be.set_input_state([1] * n + [0] * (kern.m - n))
    t0 = time.perf_counter()
    be.probability([1] * n + [0] * (kern.m - n))
    t_first = time.perf_counter() - t0
    rows = unrank_rows(kern.layers, n, torch.as_tensor(rng.integers(0, S, 200, dtype=np.int64))).tolist()
    t0 = time.perf_counter()
    for r in rows:
        be.probability(r)
    dt_l = (time.perf_counter() - t0) / len(rows)
    out["backend_n11_m32"] = {"first_probability_ms": t_first * 1e3, "lookup_us": dt_l * 1e6}
    del be
    torch.cuda.empty_cache()
    return out


def compare_every_output(p: torch.Tensor, p64: torch.Tensor) -> dict:
    """max|dp|, TVD and the max relative error over the outputs above 1e-12 of a float32
    distribution against a float64 one of the same size, on the GPU in chunks."""
    mx = tvd = rel = 0.0
    chunk = 1 << 28
    for i in range(0, p64.numel(), chunk):
        a, b = p[i:i + chunk].double(), p64[i:i + chunk]
        d = (a - b).abs()
        mx = max(mx, float(d.max()))
        tvd += float(d.sum())
        big = b > 1e-12
        if bool(big.any()):
            rel = max(rel, float((d[big] / b[big]).max()))
        del a, b, d, big
    return {"outputs": int(p64.numel()), "max_abs_dp": mx, "tvd": 0.5 * tvd, "max_rel_above_1e-12": rel}


def fp64_forward(n: int, m: int, occ, U128: torch.Tensor, dev: str):
    """The engine's complex128 forward (FP64 throughout, pinned to the oracle and to Ryser
    permanents in the GPU tests) of U128: (float64 probabilities, seconds, FP64 mass - 1)."""
    from photonic_engines_amd.slos import FockSLOSKernel
    k = FockSLOSKernel(n, m, occ, device=dev, dtype=torch.complex128)
    Uz = U128.to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p64 = k.probs(Uz)
    torch.cuda.synchronize()
    return p64, time.perf_counter() - t0, float(k.total_mass(p64)) - 1.0


def full_fp64_accuracy(p: torch.Tensor, n: int, m: int, occ, U128: torch.Tensor, dev: str) -> dict:
    """Every output of the complex64 step `p` against the engine's complex128 forward of the exact
    U (the complex64 step sees U rounded to complex64): max|dp|, TVD, the max relative error over
    outputs above 1e-12, and the complex128 forward's own time and FP64 mass."""
    p64, dt, mass = fp64_forward(n, m, occ, U128, dev)
    res = compare_every_output(p, p64)
    res.update({"oracle": "the engine's complex128 forward of the same U (FP64)", "complex128_ms": dt * 1e3,
                "complex128_mass_minus_1": mass})
    del p64
    torch.cuda.empty_cache()
    return res


def bench_gradient_frontier(steps: int = 2) -> dict:
    """n=11, m=32 (4.28e9 outputs): one gradient step of L = -mean log p over 1,024 outputs drawn
    from the exact distribution (probs(U) under autograd, L.backward() through the adjoint
    kernels: every other layer kept, the readout adjoint pulled from layer 10 at the sampled
    outputs); time and peak device memory. Skipped when the device has less than 100 GB free."""
    from photonic_engines_amd.slos import FockSLOSKernel, haar_unitary
    from photonic_engines_amd.slos.sampling import draw_indices
    dev = torch.device("cuda", 0)
    torch.cuda.empty_cache()
    free = torch.cuda.mem_get_info(dev)[0]
    if free < 100e9:
        return {"skipped": f"needs about 100 GB free, {free / 1e9:.0f} GB are"}
    n, m = 11, 32
    k = FockSLOSKernel(n, m, [1] * n + [0] * (m - n), device="cuda")
    U0 = haar_unitary(m, 1132).to(torch.complex64).cuda()
    with torch.no_grad():
        p = k.probs(U0)
        idx = draw_indices(p, 1024, torch.Generator(device="cuda").manual_seed(8))
    del p
    torch.cuda.empty_cache()

    def step():
        U = U0.clone().requires_grad_(True)
        loss = -torch.log(k.probs(U)[idx]).mean()
        loss.backward()
        return U.grad

    g = step()
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        g = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    res = {"workload": "n=11 m=32, first 11 modes, L = -mean log p over 1,024 sampled outputs",
           "states": int(k.output_size()), "ms_per_step": dt * 1e3,
           "peak_GB": torch.cuda.max_memory_allocated(dev) / 1e9,
           "grad_nonzero_columns": int((g.abs().sum(0) > 0).sum()), "grad_finite": bool(torch.isfinite(g).all())}
    del g, k
    torch.cuda.empty_cache()
    return res


def cpu_stored_edge(threads: int, reps: int = 3) -> dict:
    """The reference's own algorithm on the CPU, where its edge tables fit: the torch-CPU
    restatement of the stored-edge mode-slab forward (oracle/stored_edge.py) at n=8, m=24 (7.9M
    outputs, 63M edges), setup timed apart, then one warm-up and the median of `reps` probs(U)
    calls; beside it the GPU engine's probs(U) on the same U. At n=11, m=32 its edges alone would
    take 377 GB, so there it is not run."""
    from oracle.stored_edge import StoredEdgeSLOS
    from photonic_engines_amd.slos import FockSLOSKernel, haar_unitary
    n, m = 8, 24
    occ = [1] * n + [0] * (m - n)
    U128 = haar_unitary(m, 8024)
    torch.set_num_threads(threads)
    ref = StoredEdgeSLOS(n, m, occ)
    Uc = U128.to(torch.complex64)
    ref.probs(Uc)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        pc = ref.probs(Uc)
        ts.append(time.perf_counter() - t0)
    t_cpu = float(np.median(ts))
    kg = FockSLOSKernel(n, m, occ, device="cuda")
    Ug = Uc.cuda()
    kg.probs(Ug)
    torch.cuda.synchronize()
    gs = []
    for _ in range(10):
        t0 = time.perf_counter()
        pg = kg.probs(Ug)
        torch.cuda.synchronize()
        gs.append(time.perf_counter() - t0)
    t_gpu = float(np.median(gs))
    S = pc.numel()
    return {"workload": "n=8 m=24 (7,888,725 outputs), input [1]*8+[0]*16, seed 8024, complex64",
            "kind": "torch-CPU restatement of the reference's stored-edge mode-slab forward",
            "cores": threads, "setup_s": ref.setup_s, "cpu_ms": t_cpu * 1e3, "cpu_probs_per_s": S / t_cpu,
            "gpu_ms": t_gpu * 1e3, "gpu_probs_per_s": S / t_gpu,
            "max_abs_dp_gpu_vs_cpu": float((pg.cpu() - pc).abs().max())}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# The reference's own published times for this computation: (tag, n, m, input occupation, dtype
# label, reference ms, hardware, source). The H100 rows are FockSLOSKernel.probs(U) (`ms` column,
# median of the timed calls after one warm-up); the CPU rows are its C engine on one 48-thread
# rank (`total_s`, basis build included).
H100 = "1x NVIDIA H100 PCIe, FockSLOSKernel.probs"
CPU48 = "48-core CPU node, the reference's C engine slos_dist.c v2, 1 rank x 48 threads"
PUBLISHED_SHAPES = [
    ("n11_m22", 11, 22, [1] * 11 + [0] * 11, "complex64", 98.908, H100,
     "results/keep/sbs_20260919_product_slos_r2_n11.csv:2"),
    ("n4_m8", 4, 8, [1] * 4 + [0] * 4, "complex64", 1.99, H100, "results/keep/sbs_20260919_product_slos_r2.csv:2"),
    ("n6_m12", 6, 12, [1] * 6 + [0] * 6, "complex64", 3.739, H100, "results/keep/sbs_20260919_product_slos_r2.csv:4"),
    ("n8_m16", 8, 16, [1] * 8 + [0] * 8, "complex64", 5.139, H100, "results/keep/sbs_20260919_product_slos_r2.csv:6"),
    ("n10_m20", 10, 20, [1] * 10 + [0] * 10, "complex64", 19.357, H100,
     "results/keep/sbs_20260919_product_slos_r2.csv:8"),
    ("n4_m12", 4, 12, [1] * 4 + [0] * 8, "unrecorded", 1.965, H100,
     "results/keep/sbs_20260919_product_slos_r3_c128.csv:2"),
    ("n6_m18", 6, 18, [1] * 6 + [0] * 12, "unrecorded", 4.226, H100,
     "results/keep/sbs_20260919_product_slos_r3_c128.csv:4"),
    ("n8_m24", 8, 24, [1] * 8 + [0] * 16, "unrecorded", 12.335, H100,
     "results/keep/sbs_20260919_product_slos_r3_c128.csv:6"),
    ("n6_m12_alternating", 6, 12, [1, 0] * 6, "complex64", 0.29, H100,
     "results/keep/bench_emitter_slos_h100_20260909_modeslab.csv:2"),
    ("n8_m16_alternating", 8, 16, [1, 0] * 8, "complex64", 0.40, H100,
     "results/keep/bench_emitter_slos_h100_20260909_modeslab.csv:3"),
    ("n10_m20_alternating", 10, 20, [1, 0] * 10, "complex64", 10.57, H100,
     "results/keep/bench_emitter_slos_h100_20260909_modeslab.csv:4"),
    ("n11_m22_c_engine", 11, 22, [1] * 11 + [0] * 11, "complex128", 1918.0, CPU48,
     "results/keep/cpu48_scaling_v2.csv:8"),
    ("n12_m24_c_engine", 12, 24, [1] * 12 + [0] * 12, "complex128", 11284.0, CPU48,
     "results/keep/cpu48_scaling_v2.csv:14"),
    ("n13_m26_c_engine", 13, 26, [1] * 13 + [0] * 13, "complex128", 73087.0, CPU48,
     "results/keep/cpu48_scaling_v2_n13.csv:2"),
]


def bench_published_shapes(repeats: int = 7) -> dict:
    """FockSLOSKernel.probs(U) on every shape the reference publishes a time for, timed as the
    reference's product bench times it (bench/bench_product_slos.py:45-58: one warm-up call, then
    the median of `repeats` calls, each closed by a device synchronisation). The U is this build's
    seeded Haar unitary, seed 100 n + m. Rows whose reference dtype is not complex64 (complex128,
    or not recorded) are timed in complex128 as well."""
    from photonic_engines_amd.slos import FockSLOSKernel, haar_unitary

    def median_ms(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return sorted(ts)[len(ts) // 2]

    out = {}
    for tag, n, m, occ, dt_label, ref_ms, hw, src in PUBLISHED_SHAPES:
        U128 = haar_unitary(m, 100 * n + m)
        rec = {"n": n, "m": m, "input": occ, "reference_ms": ref_ms, "reference_dtype": dt_label,
               "reference_hardware": hw, "reference_source": src}
        dtypes = [torch.complex64] if dt_label == "complex64" else [torch.complex64, torch.complex128]
        for cd in dtypes:
            k = FockSLOSKernel(n, m, occ, device="cuda", dtype=cd)
            Ud = U128.to(cd).cuda()
            ms = median_ms(lambda: k.probs(Ud))
            name = "complex64" if cd == torch.complex64 else "complex128"
            rec[name] = {"ms": ms, "probs_per_s": k.output_size() / (ms * 1e-3), "speedup_vs_reference": ref_ms / ms}
            rec["states"] = k.output_size()
            del k, Ud
            torch.cuda.empty_cache()
        out[tag] = rec
    return out


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-last", type=int, default=0,
                    help="layer-n states in the CPU sample (0: the whole layer, no extrapolation)")
    ap.add_argument("--no-cpu-reference-algorithm", action="store_true",
                    help="skip the torch-CPU stored-edge restatement at n=8, m=24")
    ap.add_argument("--no-compensated", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 Ryser spot checks (profiling passes)")
    ap.add_argument("--no-batched", action="store_true", help="skip the uneven-batch measurement")
    ap.add_argument("--no-lossy", action="store_true", help="skip the lossy-output measurement")
    ap.add_argument("--no-gradient", action="store_true", help="skip the gradient-step measurement")
    ap.add_argument("--no-next-rows", action="store_true",
                    help="skip the sampling / x-model / superposition / grouped / backend timings")
    ap.add_argument("--no-published-shapes", action="store_true",
                    help="skip probs(U) on the shapes the reference publishes times for")
    ap.add_argument("--copy-calibration-gb", type=float, default=0.0,
                    help="also run one plain device copy of this many GB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the SLOS engine runs on a single MI355X in this build "
              "(multi-GPU sharding is designed, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from photonic_engines_amd import _native as nat
    from photonic_engines_amd.slos import FockSLOSKernel, haar_unitary

    wl = WORKLOADS[args.workload]
    n, m = wl["n"], wl["m"]
    occ = [1] * n + [0] * (m - n)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    kern = FockSLOSKernel(n, m, occ, device=str(dev))
    plan, modes = kern._plan, kern.input_modes
    U128 = haar_unitary(m, wl["seed"])
    U2 = U128.to(torch.complex64).to(dev).contiguous()        # [m, m], resident in HBM
    U = U2.unsqueeze(0)
    S = plan.sizes
    stream = torch.cuda.current_stream(dev)

    def timed(k):
        """`args.steps` calls of k.probs(U) (the drop-in op itself), after `args.warmup` untimed."""
        for _ in range(args.warmup):
            p = k.probs(U2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            p = k.probs(U2)
        torch.cuda.synchronize()
        return p, (time.perf_counter() - t0) / args.steps

    def last_layer_ms(flags: int, reps: int = 5) -> float:
        """The last-layer launch alone (the same kernel probs(U) ends with, on the same inputs):
        layer n-1 computed once, then `reps` launches of layer n bracketed by HIP events on the
        stream they run on; the mean duration. Compensated sums carry hi/lo layers, as the
        compensated forward does (the last layer then reads the parents' low-order parts too)."""
        if flags & nat.FLAG_COMPENSATED:
            flags |= nat.FLAG_HILO
        psi = torch.ones((1, 1), dtype=torch.complex64, device=dev)
        if flags & nat.FLAG_HILO:
            psi = torch.stack([psi, torch.zeros_like(psi)])
        for k in range(1, n):
            psi = plan.layer_step(U, k, modes[k - 1], psi, nat.OUT_RAW, flags=flags)
        q = plan.layer_step(U, n, modes[n - 1], psi, nat.OUT_PROBS, scale=1.0, flags=flags)
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            q = plan.layer_step(U, n, modes[n - 1], psi, nat.OUT_PROBS, scale=1.0, flags=flags, out=q)
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        del psi, q
        return float(np.mean(ms))

    def per_layer_ms(reps: int = 3) -> list:
        """Each layer launch of the forward alone (the kernels probs(U) runs, on the same inputs),
        from the layer below: the mean of `reps` launches bracketed by HIP events on their stream,
        with the layer's algorithmic bytes (8 B per parent read, 8 B per complex64 child or 4 B per
        probability written) over that time."""
        rows = []
        psi = torch.ones((1, 1), dtype=torch.complex64, device=dev)
        for k in range(1, n + 1):
            kind = nat.OUT_PROBS if k == n else nat.OUT_RAW
            q = plan.layer_step(U, k, modes[k - 1], psi, kind, scale=1.0)
            ms = []
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                q = plan.layer_step(U, k, modes[k - 1], psi, kind, scale=1.0, out=q)
                e1.record(stream)
                torch.cuda.synchronize()
                ms.append(e0.elapsed_time(e1))
            t = float(np.mean(ms))
            b = 8 * S[k - 1] + (4 if k == n else 8) * S[k]
            rows.append({"layer": k, "ms": t, "alg_bytes": int(b), "alg_GBps": b / (t * 1e-3) / 1e9,
                         "frac": b / (t * 1e-3) / 1e9 / PEAK_HBM_GBS})
            psi = q if k < n else None
            del q
        torch.cuda.empty_cache()
        return rows

    if args.copy_calibration_gb > 0:
        nbytes = int(args.copy_calibration_gb * 1e9) // 4 * 4
        src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    p, dt = timed(kern)
    S_n = S[n]
    # accuracy of this output: 10,000 seeded outputs against FP64 Ryser on the GPU, FP64 mass
    rng = np.random.default_rng(wl["seed"])
    idx = np.sort(rng.choice(S_n, size=min(10_000, S_n), replace=False)).astype(np.int64)
    idx_t = torch.as_tensor(idx, device=dev)
    max_abs = max_rel = mass = full = None
    if not args.no_accuracy:
        ref = kern.spot_probs(U128, idx_t).cpu().numpy()
        got = p[idx_t].double().cpu().numpy()
        max_abs = float(np.abs(got - ref).max())
        max_rel = float((np.abs(got - ref) / np.maximum(ref, 1e-300)).max())
        mass = float(kern.total_mass(p))
        full = full_fp64_accuracy(p, n, m, occ, U128, str(dev))
    del p
    torch.cuda.empty_cache()
    last_ms = last_layer_ms(0)
    layers = per_layer_ms()

    last_bytes = 8 * S[n - 1] + 4 * S[n]
    achieved = last_bytes / (last_ms * 1e-3) / 1e9
    traffic = load_traffic(args.workload)
    out = {
        "metric": METRIC,
        "value": S_n / dt,
        "unit": "probs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"SLOS fock n={n} m={m}, input [1]*{n}+[0]*{m - n}, one seeded Haar U "
                               f"(seed {wl['seed']}), lossless, float32 probabilities; a step is one "
                               f"FockSLOSKernel.probs(U) call",
                   "n": n, "m": m, "states": S_n, "global_batch": 1, "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "kernel": (f"k_layer_group (layer {n}, group order)" if 7 <= n <= 11 and math.comb(m + 4, 5) <= 4 << 20
                                else f"k_layer_block (layer {n})"), "kernel_ms": last_ms,
                     "alg_bytes_per_launch": last_bytes},
        "whole_forward": {"alg_bytes": bytes_alg(n, m), "alg_GBps": bytes_alg(n, m) / dt / 1e9,
                          "alg_frac": bytes_alg(n, m) / dt / 1e9 / PEAK_HBM_GBS, "layers": layers},
        "accuracy": {"max_abs_dp": max_abs, "max_rel": max_rel, "spot_outputs": int(idx.size),
                     "oracle": "FP64 Ryser on the GPU", "mass_fp64": mass, "every_output": full},
    }
    if not args.no_compensated and not args.no_accuracy:
        kc = FockSLOSKernel(n, m, occ, device=str(dev), compensated=True)
        pc, dtc = timed(kc)
        gc = pc[idx_t].double().cpu().numpy()
        del pc
        torch.cuda.empty_cache()
        last_c = last_layer_ms(nat.FLAG_COMPENSATED)
        out["variants"] = {"compensated": {"ms_per_step": dtc * 1e3, "value": S_n / dtc, "kernel_ms": last_c,
                                           "max_abs_dp": float(np.abs(gc - ref).max()),
                                           "max_rel": float((np.abs(gc - ref) / np.maximum(ref, 1e-300)).max())}}
        # the arithmetic alone: both arms against FP64 arithmetic on the complex64 U they are given
        # (the complex128 forward of that U upcast), every output; the FP64 forward's time beside
        p64, dt64, _ = fp64_forward(n, m, occ, U[0].to(torch.complex128), str(dev))
        pc = kc.probs(U2)
        cmp_c = compare_every_output(pc, p64)
        del pc
        pp = kern.probs(U2)
        cmp_p = compare_every_output(pp, p64)
        del pp, p64
        torch.cuda.empty_cache()
        out["variants"]["arithmetic_vs_fp64_on_the_same_complex64_u"] = {
            "compensated": cmp_c, "plain": cmp_p, "fp64_ms_per_step": dt64 * 1e3,
            "compensated_ms_per_step": dtc * 1e3, "plain_ms_per_step": dt * 1e3,
            "compensated_speedup_vs_fp64": dt64 / dtc}
    torch.cuda.empty_cache()
    if not args.no_lossy:
        out.setdefault("variants", {})["lossy"] = bench_lossy(kern, U[0])
        torch.cuda.empty_cache()
    if not args.no_gradient:
        out.setdefault("variants", {})["gradient_n6_m16"] = bench_gradient()
        out["variants"]["gradient_n8_m24"] = bench_adjoint()
        out["variants"]["gradient_n11_m32"] = bench_gradient_frontier()
        torch.cuda.empty_cache()
    if not args.no_next_rows:
        out.setdefault("variants", {})["next_rows"] = bench_next_rows(kern, U[0])
        torch.cuda.empty_cache()
    if not args.no_batched:
        out.setdefault("variants", {})["batched_uneven"] = bench_uneven()
        torch.cuda.empty_cache()
    if not args.no_published_shapes:
        out.setdefault("variants", {})["published_shapes"] = bench_published_shapes()
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        from oracle import pull
        threads = len(os.sched_getaffinity(0))
        last = S_n if args.cpu_last <= 0 else min(args.cpu_last, S_n)
        t_pre, t_last, done, _ = pull.forward_sample(n, m, occ, U128.numpy(), last, prec=64, threads=threads)
        full = t_pre + t_last * S_n / done
        how = (f"layers 1..{n} complete ({t_pre + t_last:.2f} s)" if done == S_n else
               f"layers 1..{n - 1} complete ({t_pre:.2f} s) + first {done} of {S_n} layer-{n} states "
               f"({t_last:.2f} s), extrapolated to {full:.1f} s")
        out["cpu_baseline"] = {"value": S_n / full, "unit": "probs/s", "cores": threads, "kind": "port",
                               "sample": f"oracle C++/OpenMP pull form of the same workload, complex64, {how}; "
                                         f"CPU {cpu_model()}"}
        if not args.no_cpu_reference_algorithm:
            out["cpu_reference_algorithm"] = cpu_stored_edge(threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
