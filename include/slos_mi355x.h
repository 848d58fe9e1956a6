/* slos_mi355x — C ABI of the MI355X (gfx950) SLOS strong-simulation engine.
 *
 * The reference exposes this path as the Python class FockSLOSKernel
 * (photonic_engines/slos/kernel.py:83-293) over a cached basis (get_layers, basis.py:275-292).
 * Its only native precedent is the cffi ABI of the distributed CPU engine
 * (slos/csrc/slos_dist.c:100-136 bound in slos/distributed.py:44-114: opaque handle,
 * create / run / rank / unrank / get / destroy, caller-owned buffers). This header is the
 * drop-in boundary a ctypes / cffi binding of FockSLOSKernel binds to; every entry point below
 * names the reference interface it replaces.
 *
 * Conventions
 *  - Every entry returns an int status: SLOS_OK (0) or an SLOS_E* code; slos_last_error() gives
 *    the message of the last failure on the calling thread. Nothing throws across the ABI.
 *  - Large buffers are caller-owned device memory. The library owns only the plan and its small
 *    counting tables (one copy per device, created on first use on that device).
 *  - Device work is stream-ordered on the stream argument (a hipStream_t; NULL = the null
 *    stream), with no hidden host synchronisation. A plan is not used from several threads at
 *    once.
 *  - U is complex64, [batch][m][m] row-major, U[b][j][i] = row j (output mode), column i
 *    (input mode) — the reference's indexing U[:, j, i] (kernel.py:224,242).
 *  - Layer k holds the k-photon states of the mask in the lexicographic order of their sorted
 *    photon positions, bit-exactly the order of FockLayers.states[k] (basis.py:120-148).
 */
#ifndef SLOS_MI355X_H
#define SLOS_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slos_plan slos_plan;
typedef struct slos_group slos_group;
typedef void* slos_stream_t; /* hipStream_t */

/* output masks: photonic_engines/masks.py:18-28 */
enum {
  SLOS_MASK_FOCK = 0,
  SLOS_MASK_UNBUNCHED = 1,
  SLOS_MASK_DUAL_RAIL = 2,
  SLOS_MASK_SPARSE_1IN4 = 3,
  SLOS_MASK_TAIL_EMPTY = 4
};

enum {
  SLOS_OK = 0,
  SLOS_EINVAL = 1,   /* bad argument (ValueError in the reference) */
  SLOS_ENOMEM = 2,   /* workspace too small / allocation failed (OutOfMemoryError) */
  SLOS_EHIP = 3,     /* a HIP runtime call failed */
  SLOS_ERANGE = 4,   /* a size does not fit the supported range */
  SLOS_ENODEV = 5    /* no usable GPU */
};

/* output kinds of slos_forward */
enum {
  SLOS_OUT_PROBS = 0, /* float32 [batch][S]: |amp * scale|^2 * prod_j t_j!   (kernel.py:262-283) */
  SLOS_OUT_AMPS = 1   /* complex64 [batch][S]: raw amplitude * scale          (kernel.py:191-260) */
};

/* flags of slos_forward */
#define SLOS_FLAG_COMPENSATED 1u /* two-sum (compensated) FP32 accumulation of every child sum */
#define SLOS_FLAG_IMMEDIATE 2u   /* amplitudes in the per-edge sqrt normalisation (kernel.py:239-244):
                                    raw amplitude * sqrt(prod_j t_j!) */
#define SLOS_FLAG_SIMPLE 4u      /* force the thread-per-state kernel (ablation / debugging) */
#define SLOS_FLAG_HILO 8u        /* slos_layer_step with SLOS_FLAG_COMPENSATED: the raw layers in
                                    (psi_prev) and out (kind 2) are hi/lo pairs [2][batch][S], the
                                    high parts then the low-order parts; the low parts enter every
                                    child sum and the stored layer keeps its rounding error
                                    (slos_forward with SLOS_FLAG_COMPENSATED does this internally) */
#define SLOS_FLAG_F64 16u        /* slos_forward / slos_forward_noalloc / slos_forward_workspace /
                                    slos_layer_step in complex128: U, the layers and the amplitudes
                                    complex128, the probabilities float64, every sum in FP64 (the
                                    reference's dtype=torch.complex128, kernel.py:92,117-118).
                                    Thread-per-state kernels; not with SLOS_FLAG_COMPENSATED */

/* Version string of the library. */
const char* slos_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* slos_last_error(void);

/* Build a plan for n photons in m modes under `mask`, photon k entering column input_modes[k]
 * (ascending occupation expansion, kernel.py:124). Replaces FockSLOSKernel.__init__
 * (kernel.py:90-167) together with get_layers / FockLayers (basis.py:92-117,275-292): no state
 * or edge table is stored; only the counting tables (a few KB). Errors (kernel.py:96-126):
 * SLOS_EINVAL for an unknown mask, n < 0, m < 1, a mode outside [0, m), or an empty output
 * space; SLOS_ERANGE if a layer size does not fit 64 bits or n > 32. n = 0 is the vacuum: one
 * output state, which slos_forward writes as probability |scale|^2 (amplitude scale). */
int slos_plan_create(int n, int m, int mask, const int32_t* input_modes, slos_plan** out);

/* Free a plan and its device tables. Replaces clear_layers_cache (basis.py:295-301). */
int slos_plan_destroy(slos_plan* plan);

/* Number of states in layer k (0 <= k <= n). Replaces FockLayers.size (basis.py:231-233). */
int slos_layer_size(const slos_plan* plan, int k, uint64_t* size);

/* Host: sorted positions [count][k] (int32) -> lex ranks. Replaces the packed-key searchsorted
 * of FockLayers._edges (basis.py:151-193) as the state -> index map. Positions must describe
 * admissible states; otherwise SLOS_EINVAL. */
int slos_rank(const slos_plan* plan, int k, const int32_t* positions, int64_t count, uint64_t* idx);

/* Host: lex ranks -> sorted positions [count][k]. Replaces FockLayers.states[k]
 * (basis.py:96,107-117) served lazily. SLOS_ERANGE for an index >= S_k. */
int slos_unrank(const slos_plan* plan, int k, const uint64_t* idx, int64_t count, int32_t* positions);

/* Host: prod_j t_j! (float64) of states first .. first+count-1 of layer k. Replaces
 * FockLayers.prob_norms / output_prob_norms (basis.py:254-272). */
int slos_prob_norms(const slos_plan* plan, int k, uint64_t first, int64_t count, double* out);

/* Device bytes of the workspace slos_forward needs for `batch` unitaries when it stops at layer
 * `layer` (layer = n for the full run) with these flags (SLOS_FLAG_COMPENSATED doubles it: the
 * intermediate layers then carry their low-order parts). */
int slos_forward_workspace(const slos_plan* plan, int64_t batch, int layer, unsigned flags, size_t* bytes);

/* The layer recurrence psi_{k+1}[c] = sum_{j distinct in c} U[b, j, i_{k+1}] psi_k[c - e_j]
 * (deferred normalisation), run from the vacuum to layer `layer` with input columns
 * `modes[0..layer-1]` (NULL: the plan's own input modes), written to `out` as SLOS_OUT_PROBS or
 * SLOS_OUT_AMPS. Replaces FockSLOSKernel._amplitudes / amplitudes / probs (kernel.py:191-283)
 * for the lossless, indist=None case; `modes` with layer < n is the subset form used by
 * slos/species.py:94. `scale` multiplies the amplitude (the input norm 1/sqrt(prod s_i!),
 * kernel.py:134-138). U: device complex64 [batch][m][m]; out: device [batch][S_layer].
 * The workspace (>= slos_forward_workspace bytes, device) is caller-owned scratch. */
int slos_forward(const slos_plan* plan, const void* U, int64_t batch, const int32_t* modes,
                 int layer, int out_kind, void* out, double scale, unsigned flags,
                 void* workspace, size_t workspace_bytes, slos_stream_t stream);

/* slos_forward without a workspace argument, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. It calls no hipMalloc / hipFree: the intermediate layers live in a
 * fixed 64 MiB device array of the library, one per GPU, held until the process ends; a forward
 * whose workspace (slos_forward_workspace) exceeds 64 MiB fails with SLOS_ERANGE (n = 8, m = 24
 * needs 21 MB). The plan's tables must already be resident on the current GPU (any earlier
 * slos_forward or slos_layer_step of the plan there), else SLOS_EINVAL. Every call on one GPU
 * shares that array: calls that overlap on one GPU (different streams, or threads) race. Results
 * equal slos_forward's bit for bit (same kernels). */
int slos_forward_noalloc(const slos_plan* plan, const void* U, int64_t batch, const int32_t* modes,
                         int layer, int out_kind, void* out, double scale, unsigned flags,
                         slos_stream_t stream);

/* One layer of the recurrence (layer k from layer k-1 with input column `col`): the building
 * block of a checkpointed forward under autograd (kernel.py:40-80 recomputes layers the same
 * way). psi_prev: device complex64 [batch][S_{k-1}]; out: SLOS_OUT_PROBS / SLOS_OUT_AMPS as in
 * slos_forward, or 2 for the raw complex64 layer. */
int slos_layer_step(const slos_plan* plan, const void* U, int64_t batch, int k, int col, const void* psi_prev,
                    int out_kind, void* out, double scale, unsigned flags, slos_stream_t stream);

/* Adjoint of the probability readout (kernel.py:278-282 under autograd): grad_amps =
 * 2 scale^2 prod_j t_j! grad_p * amps, for raw layer-n amplitudes `amps` (complex64
 * [batch][S_n]) and float32 grad_p [batch][S_n]. */
int slos_readout_adjoint(const slos_plan* plan, int64_t batch, const void* amps, const float* grad_p,
                         double scale, void* grad_amps, slos_stream_t stream);

/* The same adjoint from layer n-1 instead of layer n, for a grad_p that is zero almost everywhere (a
 * loss over sampled outputs): where grad_p is nonzero the state's raw amplitude is pulled from
 * psi_prev (complex64 [batch][S_{n-1}], layer n-1 of U, complex64 [batch][m][m]); grad_amps is zero
 * elsewhere. One pass over grad_p and grad_amps in place of recomputing layer n. No counterpart of
 * its own in the reference (whose autograd forms the dense product). */
int slos_readout_adjoint_pull(const slos_plan* plan, const void* U, int64_t batch, const void* psi_prev,
                              const float* grad_p, double scale, void* grad_amps, slos_stream_t stream);

/* Adjoint of one layer (the backward of _CheckpointedLayer, kernel.py:63-80, written as a pull
 * over children instead of scatter_add_): grad_prev[t] = sum_j conj(U[b,j,col]) grad_k[t + e_j]
 * (skipped if grad_prev is NULL) and grad_U[b][j][col] (+)= sum_t conj(psi_prev[t]) grad_k[t + e_j]
 * (accumulate != 0 adds to grad_U). Deterministic two-stage reduction, no atomics; workspace
 * >= slos_adjoint_workspace bytes. Requires m <= 96 (LDS). */
int slos_adjoint_workspace(const slos_plan* plan, int64_t batch, int k, size_t* bytes);
int slos_adjoint_layer(const slos_plan* plan, const void* U, int64_t batch, int k, int col, const void* psi_prev,
                       const void* grad_k, void* grad_prev, void* grad_U, int accumulate, void* workspace,
                       size_t workspace_bytes, slos_stream_t stream);

/* One step of the uniform photon-removal recurrence of output loss (imperfections.py:254-276):
 * q_{k-1}(t) = scale * (sum_j q_k(t + e_j) (t_j + 1)) / k, for 1 <= k <= n. Pull form (each
 * (k-1)-photon state gathers its children), no atomics; children in ascending mode order. q_k:
 * device float32 [batch][S_k]; q_km1: device float32 [batch][S_{k-1}]. `scale` lets a caller carry
 * the binomial weights (imperfections.py:106-111) down the layers (weight ratio of the two layers);
 * 1 gives the plain recurrence. flags: SLOS_FLAG_SIMPLE forces the thread-per-state kernel
 * (otherwise multiset spaces with k - 1 > 4 use the block kernel; both give bit-identical results). */
int slos_loss_descend(const slos_plan* plan, int k, const float* q_k, float* q_km1, int64_t batch, unsigned flags,
                      double scale,
                      slos_stream_t stream);

/* Probability-level raise of partial distinguishability (species.classical_raise,
 * slos/species.py:37-56): a photon that interferes with nothing enters column col, and
 * q_k(c) = sum_{j distinct in c} w[b][j] q_{k-1}(c - e_j) with w[b][j] = |U[b, j, col]|^2 (device
 * float32 [batch][m]), as a pull over parents instead of index_add over edges. q_km1: [batch][S_{k-1}],
 * q_k: [batch][S_k], device float32. */
int slos_raise(const slos_plan* plan, int k, const float* w, const float* q_km1, float* q_k, int64_t batch,
               slos_stream_t stream);

/* Adjoint of slos_loss_descend (the backward of imperfections._descend under autograd):
 * g_k(c) = sum_{j distinct in c} g_{k-1}(c - e_j) c_j / k. g_km1: [batch][S_{k-1}], g_k:
 * [batch][S_k], device float32. */
int slos_loss_ascend(const slos_plan* plan, int k, const float* g_km1, float* g_k, int64_t batch,
                     slos_stream_t stream);

/* Group sums of float32 probabilities: out[b][g] = sum of probs[b][i] over the states i of group g,
 * in FP64 and in a fixed order (no atomics), rounded to float32. Replaces the scatter_add_ of
 * GroupedSLOSKernel.probs (kernel.py:604-620). order [count]: the state indices sorted by group
 * (stable); chunks [n_chunks][2]: the [begin, end) positions in `order` of each chunk (one wave
 * sums a chunk: 4096 entries balance the work), a group's chunks consecutive; first [n_groups + 1]: group g owns chunks
 * [first[g], first[g + 1]) (an empty group sums to 0). Device pointers. workspace >= batch *
 * n_chunks * 8 bytes. */
int slos_group_sums(const float* probs, int64_t batch, uint64_t count, const int64_t* order, const int64_t* chunks,
                    uint32_t n_chunks, const uint32_t* first, int n_groups, float* out, void* ws, size_t ws_bytes,
                    slos_stream_t stream);

/* FP64 total mass of float32 probabilities: out[b] = sum_s probs[b][s] (device double [batch]).
 * Two-stage deterministic reduction (no atomics). workspace >= slos_mass_workspace bytes. */
int slos_mass_workspace(int64_t batch, uint64_t count, size_t* bytes);
int slos_total_mass(const float* probs, int64_t batch, uint64_t count, double* out,
                    void* workspace, size_t workspace_bytes, slos_stream_t stream);

/* FP64 Ryser spot oracle on the device: out[q] = |per(U[t_q rows][:, input cols])|^2 /
 * (prod s_i! prod t_j!) for the layer-n states with lex indices idx[q] (device uint64 [count]),
 * for unitary `U` (device complex128 [m][m]). Restates slos/reference.py:24-36,88-105 in FP64
 * (Gray-code Ryser), for accuracy reports at sizes where no full FP64 run exists. */
int slos_spot_probs(const slos_plan* plan, const void* U_c128, const uint64_t* idx, int64_t count,
                    double* out, slos_stream_t stream);

/* ---- strong-then-sample --------------------------------------------------------------------
 * Draws of output indices from float32 probabilities on the device (the inverse-CDF branch of
 * draw_indices, imperfections.py:221-237, used by sample_from_probs, slos/emitter.py:215-230, and
 * LossyDistribution.sample, imperfections.py:213-218), without a float64 copy of the [count]
 * vector: FP64 sums of 64-entry chunks (in index order) and 4096-entry blocks (in chunk order), the
 * block CDF (sequential within segments of 1024 blocks, plus the sequential prefix of the segment
 * totals), then per shot u = uniforms[s] * total: the first block whose CDF reaches u, the first
 * chunk and entry whose in-order prefix reaches the remainder (clamped to the block / count end).
 * probs: device float32 [count] (non-negative, unnormalised); uniforms: device float64 [shots] in
 * [0, 1); idx: device uint64 [shots]. count <= 2^32. The workspace holds the block sums, the block
 * CDF and the FP64 chunk sums (count / 64 of them). */
int slos_sample_workspace(uint64_t count, size_t* bytes);
int slos_sample(const float* probs, uint64_t count, const double* uniforms, int64_t shots, uint64_t* idx,
                void* workspace, size_t workspace_bytes, slos_stream_t stream);

/* Device unrank of layer k: occ (device int32 [count][m]) is the occupation row of each lex index
 * (device uint64 [count]); an index >= S_k gives a row of -1. The device form of slos_unrank for
 * sampled indices (occupations as in sample_from_probs, slos/emitter.py:228-229). */
int slos_unrank_device(const slos_plan* plan, int k, const uint64_t* idx, int64_t count, int32_t* occ,
                       slos_stream_t stream);

/* ---- uneven batches ------------------------------------------------------------------------
 * A group runs many instances of up to 8 plans of different (n, m, input, mask) together: one
 * launch per photon layer K covers every instance of every plan with n >= K (batched U, as
 * kernel.py:208-215 batches one shape, widened to several shapes). The reference has no
 * counterpart for mixed shapes; each plan's output equals slos_forward's (SLOS_OUT_PROBS) for the
 * same U, bit for bit. Per plan g: U[g] complex64 [batch[g]][m][m] (device), out[g] float32
 * [batch[g]][S_n] (device), scale[g] the input norm 1/sqrt(prod s_i!). Plans must outlive the group. */
int slos_group_create(int count, const slos_plan* const* plans, slos_group** out);
int slos_group_destroy(slos_group* group);
int slos_group_workspace(const slos_group* group, const int64_t* batch, unsigned flags, size_t* bytes);
int slos_group_forward(slos_group* group, const void* const* U, const int64_t* batch, float* const* out,
                       const double* scale, unsigned flags, void* workspace, size_t workspace_bytes,
                       slos_stream_t stream);

/* Autotuner: times candidate launches of every layer on these buffers (block or thread-per-state
 * kernel, children per wave / states per thread, persistent grid size), `reps` timed launches
 * each, and keeps the fastest per layer for later forwards and captures. Synchronises `stream`.
 * Leaves the full forward result in `out`. */
int slos_group_tune(slos_group* group, const void* const* U, const int64_t* batch, float* const* out,
                    const double* scale, unsigned flags, void* workspace, size_t workspace_bytes, int reps,
                    slos_stream_t stream);
/* The launch chosen for layer K: block (1 block kernel, 0 thread-per-state, 2 one launch per plan
 * through the single-plan kernels, -1 default: one launch per plan where a single plan reaches
 * layer K, else the block kernel where it fits), per_wave, grid (0: one work group per item). */
int slos_group_config(const slos_group* group, int K, int* block, uint64_t* per_wave, uint32_t* grid);

/* Capture one forward over these buffers as a hipGraph (on an internal stream; nothing is
 * allocated or memset inside the graph), then replay it on any stream. The buffers must stay
 * allocated while the graph is replayed; new inputs are written into the same U buffers. */
int slos_group_capture(slos_group* group, const void* const* U, const int64_t* batch, float* const* out,
                       const double* scale, unsigned flags, void* workspace, size_t workspace_bytes);
int slos_group_replay(slos_group* group, slos_stream_t stream);

/* complex128 / float64 twins (FockSLOSKernel(dtype=torch.complex128), kernel.py:92,117-118) of
 * the entry points with typed float pointers or no flags argument; same arguments and results
 * as their complex64 forms, in FP64, thread-per-state:
 *   slos_loss_descend_f64   slos_loss_descend (imperfections.py:254-276) on float64 q
 *   slos_loss_ascend_f64    slos_loss_ascend on float64 gradients
 *   slos_readout_adjoint_f64  slos_readout_adjoint: complex128 amps / grad_amps, float64 grad_p
 *   slos_adjoint_workspace_f64, slos_adjoint_layer_f64  slos_adjoint_workspace / slos_adjoint_layer
 *     (kernel.py:63-80): complex128 U, psi_prev, grad_k, grad_prev and grad_U
 *   slos_group_sums_f64     slos_group_sums (kernel.py:604-620) on float64 probabilities and sums */
int slos_loss_descend_f64(const slos_plan* plan, int k, const double* q_k, double* q_km1, int64_t batch, double scale,
                          slos_stream_t stream);
int slos_loss_ascend_f64(const slos_plan* plan, int k, const double* g_km1, double* g_k, int64_t batch,
                         slos_stream_t stream);
int slos_readout_adjoint_f64(const slos_plan* plan, int64_t batch, const void* amps, const double* grad_p,
                             double scale, void* grad_amps, slos_stream_t stream);
int slos_adjoint_workspace_f64(const slos_plan* plan, int64_t batch, int k, size_t* bytes);
int slos_group_sums_f64(const double* probs, int64_t batch, uint64_t count, const int64_t* order,
                        const int64_t* chunks, uint32_t n_chunks, const uint32_t* first, int n_groups, double* out,
                        void* workspace, size_t workspace_bytes, slos_stream_t stream);
int slos_adjoint_layer_f64(const slos_plan* plan, const void* U, int64_t batch, int k, int col, const void* psi_prev,
                           const void* grad_k, void* grad_prev, void* grad_U, int accumulate, void* workspace,
                           size_t workspace_bytes, slos_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SLOS_MI355X_H */
