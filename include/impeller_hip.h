/*
 * impeller_hip.h -- C ABI of the MI355X-native BN254 NTT / MSM / fused-gate
 * library (libimpeller_hip.so, built for gfx950).
 *
 * The first block keeps the reference's exported symbols, argument order and
 * meaning, with cudaError_t / cudaStream_t replaced by hipError_t /
 * hipStream_t, so the reference's Rust wrappers (core/src/ntt.rs,
 * core/src/msm.rs, core/src/fused_kernels.rs) can bind it through libloading
 * unchanged apart from those two type names.  Every entry point returns
 * hipSuccess (0) or the first HIP error / hipErrorInvalidValue for bad
 * arguments, the reference's C-side CUDA_CHECK convention
 * (core/src/common/error/src/check.cuh:1-10).
 *
 * Data layout (same as the reference): a field element is 8 little-endian u32
 * limbs in Montgomery form (R = 2^256), canonical; a G1 affine point is x || y
 * (16 u32), identity (0, 0).
 */
#ifndef IMPELLER_HIP_H
#define IMPELLER_HIP_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* core/src/common/iter/src/iter.cuh:12-26 and core/src/poly_ptr.rs:4-21:
 * logical element i lives at physical index (offset + i) % whole_len, where a
 * non-zero rotate (only on an unsliced view) means offset = len - rotate. */
typedef struct PolyPtr {
    uint32_t *ptr;
    size_t len;
    size_t rotate;
    size_t offset;
    size_t whole_len;
} PolyPtr;

typedef struct ConstPolyPtr {
    const uint32_t *ptr;
    size_t len;
    size_t rotate;
    size_t offset;
    size_t whole_len;
} ConstPolyPtr;

/* core/src/common/slice/poly_slice.h:3-13, common/src/typ.rs:13-40:
 * to_origin(i) = (begin + i) % len */
typedef struct PolySlice {
    unsigned long long begin, slice_len, len;
} PolySlice;

/* ===================== reference-compatible entry points ================== */

/* Replaces core/src/ntt/src/ntt.h:6 (ssip_ntt).  In-place forward NTT of x
 * (natural order in and out), bit-identical to halo2 best_fft with the same
 * omega, for any omega (core/tests/ntt.rs:63,151-153).  twiddle[i] =
 * omega^i for i < 2^(log_len-1), device memory.  Asynchronous on stream;
 * allocates nothing (the passes run in place, as in the reference, whose
 * runtime plans all device memory).  The two thread-count hints are accepted
 * and ignored: the pass split comes from the library's in-place plan. */
hipError_t ssip_ntt(PolyPtr x, const unsigned int *twiddle, unsigned int log_len,
                    hipStream_t stream, const unsigned int max_threads_stage1_log,
                    const unsigned int max_threads_stage2_log);

/* Replaces core/src/ntt/src/ntt.h:9 (ssip_precompute).  Writes unit^i,
 * i < len, into the HOST buffer twiddle (callers pass len = n/2,
 * core/src/ntt.rs:224-244).  Synchronous. */
hipError_t ssip_precompute(unsigned int *twiddle, unsigned int len, const unsigned int *unit);

/* Replaces core/src/ntt/src/ntt.h:12 (recompute_ntt).  Same transform as
 * ssip_ntt without an n/2 twiddle table: every pass work-group derives its
 * twiddles into LDS from pq_d[i] = omega^(i 2^(log_len - pq_deg)) (i <
 * 2^(pq_deg-1), used for a pass of at most pq_deg levels; may be NULL) and
 * omegas_d[k] = omega^(2^k) (k < 32), both device memory as written by
 * gen_pq_omegas.  Asynchronous; allocates nothing; device memory in use is
 * x, pq and omegas only (recompute_ntt.cuh:8-20,304-333). */
hipError_t recompute_ntt(PolyPtr x, const unsigned int *pq_d, unsigned int pq_deg,
                         const unsigned int *omegas_d, unsigned int log_len,
                         hipStream_t stream, const unsigned int max_threads_stage1_log,
                         const unsigned int max_threads_stage2_log);

/* Replaces core/src/ntt/src/ntt.h:15 (gen_pq_omegas), computed on the host
 * into host memory (core/src/ntt/src/precompute.cuh:89-112): omegas[i] =
 * unit^(2^i) for i < 32 (always written), pq[0] = 1 and pq[i] = (unit^(len
 * >> pq_deg))^i for 0 < i < 2^(pq_deg-1).  NULL pointers or pq_deg > 31
 * leave both outputs untouched. */
void gen_pq_omegas(unsigned int *pq, unsigned int *omegas, unsigned int pq_deg,
                   unsigned int len, unsigned int *unit);

/* Replaces core/src/ntt/src/ntt.h:17 (distribute_powers):
 * x[i] *= powers[poly_slice->to_origin(i) % power_num] for logical i < x.len. */
hipError_t distribute_powers(PolyPtr poly, const unsigned int *powers,
                             unsigned long long power_num, hipStream_t stream,
                             const PolySlice *poly_slice);

/* Replaces core/src/msm/src/msm.h:4-9 (msm).  Synchronous; h_points and
 * h_scaler_batch are HOST arrays (points: n_precompute tables of len affine
 * points; scalars: batches arrays of len Fr); h_result[i] receives the affine
 * result of batch i.  With buffers == NULL, buffer_sizes[card] receives the
 * device temp bytes each card needs and nothing else happens.  As in the
 * reference (msm_impl.cuh:373-580,703-756), the points are processed in
 * `parts` ranges through `stage_points` device buffers and the scalars
 * through `stage_scalers` buffers, copied on a second stream while the
 * previous range computes, `batch_per_run` batches at a time (0 counts as
 * 1), so the device buffer can be far smaller than the point set.  One card
 * (cards[0]) does all the work; the others' sizes are 0.  This symbol uses
 * the library's default window configuration (window 16, no precompute
 * tables); per-configuration symbols are exported as msm_cfg below. */
hipError_t msm(void *const *buffers, unsigned long *buffer_sizes, unsigned long long len,
               unsigned int batch_per_run, unsigned int parts, unsigned int stage_scalers,
               unsigned int stage_points, unsigned int num_cards, const unsigned int *cards,
               unsigned int const *const *h_points, unsigned int batches,
               unsigned int const *const *h_scaler_batch, unsigned int *const *h_result);

/* Replaces core/src/msm/src/msm.h:11 (msm_precompute) for the default
 * configuration: h_points[0] holds len affine points; h_points[j] (j <
 * n_precompute) receives 2^(j * n_windows * window) * P_i.  Synchronous. */
hipError_t msm_precompute(unsigned long long len, unsigned int *const *h_points,
                          unsigned int max_cards);

/* The reference bakes (window, target_window) into each .so name
 * (libmsm_bn254_<s>_<alpha>_<debug>.so, core/src/msm.rs:101-112); these take
 * them as arguments so thin per-configuration shims can forward to them. */
hipError_t msm_cfg(unsigned int window, unsigned int target_window, void *const *buffers,
                   unsigned long *buffer_sizes, unsigned long long len,
                   unsigned int batch_per_run, unsigned int parts,
                   unsigned int stage_scalers, unsigned int stage_points,
                   unsigned int num_cards, const unsigned int *cards,
                   unsigned int const *const *h_points, unsigned int batches,
                   unsigned int const *const *h_scaler_batch, unsigned int *const *h_result);
hipError_t msm_precompute_cfg(unsigned int window, unsigned int target_window,
                              unsigned long long len, unsigned int *const *h_points,
                              unsigned int max_cards);

/* Replaces a generated fused kernel with the signature emitted by
 * core/src/fused_kernels/kernel_gen.rs:309-330, for the Lagrange-form gate
 * out = a*b*c - d + q*e*(1 - q) over Fr.  vars (device array of 6
 * ConstPolyPtr) = {a, b, c, d, q, e}; mut_vars (device array of 1 PolyPtr)
 * = {out}, read by the kernel itself; len elements; slice_info (host) only
 * places Coef-form constants, of which this gate has none; local_buffer
 * unused (no spills).  Asynchronous, kernel launch only (capturable). */
hipError_t fused_gate_abc_d_qe(const ConstPolyPtr *vars, const PolyPtr *mut_vars,
                               unsigned long long len, const PolySlice *slice_info,
                               void *local_buffer, hipStream_t stream);

/* ------- polynomial library: core/src/poly/src/poly.h:6-30 (libpoly.so) -------
 * Every entry is asynchronous on stream (kernel launches only, so they can
 * be captured into a HIP graph) and allocates nothing.  Element-wise views
 * follow PolyPtr / ConstPolyPtr above.  The four with a temporary buffer
 * keep the reference's convention: with temp == NULL the needed bytes go to
 * *size and nothing runs; with a buffer (16-byte aligned), size may be NULL
 * (the runtime passes NULL, core/src/poly.rs:395-410) or hold the buffer's
 * bytes, which are then checked. */

/* poly.h:6-10 (poly_basic.cu:4-54): r[i] = a[i] op b[i] for i < min(len);
 * for min(len) <= i < max(len), the longer operand alone (negated when it
 * is the subtrahend of poly_sub; poly_mul copies it).  r.len >= max(len). */
hipError_t poly_add(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream);
hipError_t poly_sub(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream);
hipError_t poly_mul(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream);

/* poly.h:12-16 (poly_basic.cuh:58-120): one everywhere; one where
 * slice_info->to_origin(i) == 0 and zero elsewhere (slice_info NULL: the
 * whole array, i.e. index 0); zero everywhere. */
hipError_t poly_one_lagrange(PolyPtr target, hipStream_t stream);
hipError_t poly_one_coef(PolyPtr target, hipStream_t stream, const PolySlice *slice_info);
hipError_t poly_zero(PolyPtr target, hipStream_t stream);

/* poly.h:18 (poly_eval.cuh): *res = sum_i poly[i] * x^i; x and res are one
 * device element each; 0 for an empty poly. */
hipError_t poly_eval(void *temp_buf, unsigned long *temp_buf_size, ConstPolyPtr poly,
                     unsigned int *res, const unsigned int *x, hipStream_t stream);

/* poly.h:20 (kate_division.cuh:48-88): q = (p(X) - p(b)) / (X - b) for p of
 * 2^log_p coefficients (p.len == q.len == 2^log_p): q[i] = sum_{j > i}
 * p[j] b^(j-i-1), q[2^log_p - 1] = 0.  b: one device element.  The size
 * query reads only log_p. */
hipError_t kate_division(void *temp_buf, unsigned long *temp_buf_size, unsigned int log_p,
                         ConstPolyPtr p, const unsigned int *b, PolyPtr q, hipStream_t stream);

/* poly.h:22 (scan_mul.cuh): target[i] := target[0] * ... * target[i], in place. */
hipError_t scan_mul(void *temp_buffer, unsigned long *buffer_size, PolyPtr target,
                    hipStream_t stream);

/* poly.h:24 (batched_invert.cuh:50-101): every non-zero element replaced by
 * its inverse, zeros kept, in place. */
hipError_t batched_invert(void *temp_buffer, unsigned long *buffer_size, PolyPtr poly,
                          hipStream_t stream);

/* poly.h:26-28 (poly_basic.cu:72-80): *target := target^-1 (0 stays 0);
 * *target := target^exp.  One device element, one thread. */
hipError_t inv_scalar(unsigned int *target, hipStream_t stream);
hipError_t scalar_pow(unsigned int *target, unsigned long long exp, hipStream_t stream);

/* poly.h:30 (permute.cuh): the lookup permutation of the first `usable`
 * rows: res_input = those inputs sorted ascending by canonical value;
 * res_table = at each run start of res_input that value (one copy taken
 * from the table's first `usable` rows), at every other row the table's
 * leftover values in ascending order.  input.len == table.len >= usable,
 * res_input.len == res_table.len == usable.  An input value absent from
 * the table keeps its own value at its run start (the reference asserts). */
hipError_t permute(void *temp_buffer, unsigned long *buffer_size, unsigned long usable,
                   ConstPolyPtr input, ConstPolyPtr table, PolyPtr res_input, PolyPtr res_table,
                   hipStream_t stream);

/* ======================= device-resident extensions ====================== */

/* The same gate over contiguous device arrays, without the view round trip
 * (kernel launch only, so it can be captured): inputs is a HOST array of six
 * device pointers {a, b, c, d, q, e}. */
hipError_t impeller_gate(const uint32_t *const *inputs, uint32_t *out, unsigned long long len,
                         hipStream_t stream);
/* No reference analogue: entry points for callers that keep everything in
 * HBM (the batched / graph-captured proving path and the benchmarks). */

/* Forward (inverse = 0) or inverse (inverse = 1, result scaled by n^-1) NTT
 * of a contiguous device array, twiddle = omega^i table (i < n/2) of the
 * omega actually used (omega^-1 for the inverse).  coset_shift (nullable,
 * device, 1 element) multiplies x[i] by shift^i before a forward transform
 * or by shift^-i after an inverse one (shift passed as its inverse then).
 * plan_passes (n_passes entries, or NULL for the library default):
 * {fuse, lds_log, d0, d1, ...}: fuse = DIT levels done inside the
 * bit-reversal kernel (0; 4 when log_len >= 8, 16 x 16 tiles; 5 when
 * log_len >= 10, 32 x 32 tiles), or 1: no bit-reversal kernel, the first
 * pass reads x bit-reversed and writes a second 2^log_len-element buffer
 * (at least two passes); lds_log = log2 of the elements one work-group
 * stages in LDS (10 or 11), d_i = levels of each following pass; fused
 * levels + sum(d_i) == log_len.  Scratch (coset / inverse factor tables,
 * the second buffer) comes from the device's stream-ordered pool, whose
 * settings are left alone; without room for the second buffer the same
 * passes run in place (same result). */
hipError_t impeller_ntt(uint32_t *x, const uint32_t *twiddle, unsigned int log_len,
                        int inverse, const uint32_t *coset_shift,
                        const unsigned int *plan_passes, unsigned int n_passes,
                        hipStream_t stream);

/* impeller_ntt with caller-owned scratch (no reference analogue): allocates
 * nothing.  workspace_size gives the bytes for a plan (0 for an invalid
 * plan): the factor tables, plus the second buffer when the plan fuses the
 * bit reversal; a workspace of at least the factor part but without room
 * for the second buffer runs the same passes in place. */
size_t impeller_ntt_workspace_size(unsigned int log_len, const unsigned int *plan_passes,
                                   unsigned int n_passes);
hipError_t impeller_ntt_ws(uint32_t *x, const uint32_t *twiddle, unsigned int log_len,
                           int inverse, const uint32_t *coset_shift,
                           const unsigned int *plan_passes, unsigned int n_passes,
                           void *workspace, size_t workspace_bytes, hipStream_t stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): impeller_ntt with the same arguments and results, whose
 * scratch (the coset / inverse factor tables and n^-1, 1.5 MiB + 256 B) is a
 * static device array that each GPU the library runs on holds until the
 * process ends; it calls no hipMalloc / hipFree.  Plans that fuse the bit
 * reversal run the same passes in place (there is no second buffer).
 * log_len above 31 is rejected (hipErrorInvalidValue).  Calls on one GPU
 * share that GPU's array: calls that overlap on one GPU (different streams)
 * race on it and must be serialised by the caller; each call rewrites the
 * parts it reads before reading them. */
hipError_t impeller_ntt_noalloc(uint32_t *x, const uint32_t *twiddle, unsigned int log_len,
                                int inverse, const uint32_t *coset_shift,
                                const unsigned int *plan_passes, unsigned int n_passes,
                                hipStream_t stream);

/* Batched NTT (no reference analogue): count independent transforms run as
 * one launch per stage (bit reversal, then each LDS pass), work-groups
 * striding over the concatenated tiles of all transforms.  Per item: data
 * (2^log_n elements, in place) receives the result; src, when non-NULL, is
 * the 2^log_src-element input, zero-extended to 2^log_n (coset extension);
 * twiddle/inverse/coset_shift/plan as for impeller_ntt.  prepare fills the
 * device buffer dev (at least impeller_ntt_batch_bytes(count) bytes) and the
 * host descriptor *batch, synchronously; run only launches kernels (it can be
 * captured into a HIP graph); max_blocks (0 = no cap) bounds the grid of
 * each stage launch.  Items whose plan (given or default) fuses the bit
 * reversal into the first pass need a second 2^log_n-element buffer: a dev
 * buffer of impeller_ntt_batch_bytes_items(items, count) bytes holds them
 * all; items without room run the same passes in place after the
 * bit-reversal stage (same result). */
typedef struct ImpNttItem {
    uint32_t *data;
    const uint32_t *src;
    unsigned int log_n, log_src;
    const uint32_t *twiddle;
    int inverse;
    const uint32_t *coset_shift;
    unsigned int plan[10];
    unsigned int plan_len;
} ImpNttItem;

typedef struct ImpNttBatch {
    void *dev;
    unsigned int count, nstages;
    unsigned int total[9];
    unsigned int lds[9];
} ImpNttBatch;

size_t impeller_ntt_batch_bytes(unsigned int count);
size_t impeller_ntt_batch_bytes_items(const ImpNttItem *items, unsigned int count);
hipError_t impeller_ntt_batch_prepare(const ImpNttItem *items, unsigned int count, void *dev,
                                      size_t bytes, ImpNttBatch *batch, hipStream_t stream);
hipError_t impeller_ntt_batch_run(const ImpNttBatch *batch, unsigned int max_blocks,
                                  hipStream_t stream);

/* Stream capture into a HIP graph (no reference analogue): begin/end wrap
 * hipStreamBeginCapture / hipStreamEndCapture + hipGraphInstantiate; the
 * executable graph is returned as an opaque handle. */
hipError_t impeller_graph_begin(hipStream_t stream);
hipError_t impeller_graph_end(hipStream_t stream, void **graph_exec);
hipError_t impeller_graph_launch(void *graph_exec, hipStream_t stream);
hipError_t impeller_graph_destroy(void *graph_exec);

/* Fills twiddle[i] = omega^i (i < len) on the device.  omega is host memory. */
hipError_t impeller_twiddles(uint32_t *twiddle_d, unsigned long long len,
                             const uint32_t *omega_host, hipStream_t stream);

/* Device MSM over device-resident affine bases / Fr scalars.  window is the
 * bucket width c (8..20); the result (affine, 16 u32) is written to device
 * memory.  workspace: impeller_msm_workspace_size(len, window) bytes. */
size_t impeller_msm_workspace_size(unsigned long long len, unsigned int window);
hipError_t impeller_msm(const uint32_t *bases_d, const uint32_t *scalars_d,
                        unsigned long long len, unsigned int window, uint32_t *result_d,
                        void *workspace, size_t workspace_bytes, hipStream_t stream);

/* Configurable device MSM.  tables_d holds impeller_msm_tables(window,
 * target_windows) tables of len affine points: table t = 2^(t * n_windows *
 * window) * P with n_windows = min(target_windows, floor(254/window) + 1)
 * (target_windows = 0: one bucket set per window, a single table); the
 * reference's precompute layout (core/src/msm/src/msm_impl.cuh:56-67).
 * With more than one table, impeller_msm_precompute fills tables 1.. from
 * table 0 (affine, Montgomery R = 2^256) in place and leaves every table,
 * table 0 included, in the table format: each coordinate in Montgomery form
 * with R' = 2^261, canonical, as a 256-bit integer in 8 u32 (the form the
 * bucket kernels read); call it once per table buffer.  A single table
 * (target_windows 0, or at least floor(254/window) + 1) is the plain bases,
 * left as given and converted by each MSM call into its workspace.
 * table_stride is the number of points between consecutive tables (0 = len,
 * larger when a shorter MSM uses a prefix of a shared table set).  task is
 * the number of bucket entries one thread accumulates before the per-bucket
 * partial-sum tree takes over. */
size_t impeller_msm_cfg_workspace_size(unsigned long long len, unsigned int window,
                                       unsigned int target_windows, unsigned int task);
unsigned int impeller_msm_tables(unsigned int window, unsigned int target_windows);
hipError_t impeller_msm_precompute(uint32_t *tables_d, unsigned long long len,
                                   unsigned int window, unsigned int target_windows,
                                   hipStream_t stream);
hipError_t impeller_msm_cfg(const uint32_t *tables_d, const uint32_t *scalars_d,
                            unsigned long long len, unsigned long long table_stride,
                            unsigned int window,
                            unsigned int target_windows, unsigned int task, uint32_t *result_d,
                            void *workspace, size_t workspace_bytes, hipStream_t stream);

/* Batched MSM (no reference analogue): count independent MSMs, possibly of
 * different sizes and configurations and sharing one table set (table_stride
 * = points between consecutive tables of that set, 0 = len), run as one
 * launch per pipeline stage over the concatenated work of all of them, so
 * the latency-bound tail stages of all MSMs overlap.  prepare is
 * synchronous and carves every MSM's workspace out of dev
 * (impeller_msm_batch_bytes bytes); run only launches kernels and can be
 * captured into a HIP graph.  Each result (affine) goes to item.result. */
typedef struct ImpMsmItem {
    const uint32_t *tables;
    const uint32_t *scalars;
    unsigned long long len;
    unsigned long long table_stride;
    unsigned int window, target_windows, task;
    uint32_t *result;
} ImpMsmItem;

typedef struct ImpMsmBatch {
    void *dev;
    unsigned int count, nstages;
    unsigned int kind[128];
    unsigned int arg[128];
    unsigned int total[128];
} ImpMsmBatch;

size_t impeller_msm_batch_bytes(const ImpMsmItem *items, unsigned int count);
hipError_t impeller_msm_batch_prepare(const ImpMsmItem *items, unsigned int count, void *dev,
                                      size_t bytes, ImpMsmBatch *batch, hipStream_t stream);
hipError_t impeller_msm_batch_run(const ImpMsmBatch *batch, unsigned int max_blocks,
                                  hipStream_t stream);

/* Seeded synthetic inputs on the device (the CPU checker has the same
 * definition): kind 0 uniform Fr, 1 selector-like (90% zero, else one),
 * 2 small (< 2^16).  Bases: P_i = k_i * G with k_i = uniform Fr (seed). */
hipError_t impeller_gen_scalars(uint32_t *out_d, unsigned long long n, int kind,
                                unsigned long long seed, unsigned long long start,
                                hipStream_t stream);
hipError_t impeller_gen_bases(uint32_t *out_d, unsigned long long n, unsigned long long seed,
                              unsigned long long start, hipStream_t stream);

/* Device field helpers used by tests of the arithmetic itself (no reference
 * analogue): op 0 mul, 1 add, 2 sub, 3 neg(a), 4 from_number(a),
 * 5 to_number(a), 6 pow(a, b as integer), 7 inv(a), and through the
 * 9 x 29-bit form (R-form in and out): 8 mul, 9 a - b, 10 a - 3b, 11 a^2;
 * field 0 = Fr, 1 = Fq; n elements. */
hipError_t impeller_field_op(int field, int op, uint32_t *r_d, const uint32_t *a_d,
                             const uint32_t *b_d, unsigned long long n, hipStream_t stream);
/* G1 helpers: op 0 madd(xyzz, aff), 1 add(xyzz, xyzz), 2 dbl(xyzz),
 * 3 to_affine(xyzz) -> aff, 4 scalar_mul(aff, integer) -> aff, the
 * 29-bit forms 5 madd, 6 add, 7 dbl, 8 -aff through the table format, and
 * the 29-bit group-of-four-lanes forms 9 add, 10 dbl; n items. */
hipError_t impeller_g1_op(int op, uint32_t *r_d, const uint32_t *a_d, const uint32_t *b_d,
                          unsigned long long n, hipStream_t stream);
/* Montgomery-multiply throughput probe (no reference analogue): blocks x 256
 * threads, each running two independent chains of iters products (form 0:
 * the 8 x 32-bit Fr product; form 1: the 9 x 29-bit Fq product used by the
 * MSM bucket kernels); out_d receives blocks * 256 words.  Products per
 * launch = blocks * 256 * 2 * iters. */
hipError_t impeller_mont_mul_rate(int form, unsigned int blocks, unsigned int iters,
                                  uint32_t *out_d, hipStream_t stream);

/* FP32 cost model of the autotuner (no reference analogue): a 3-layer MLP
 * F -> H -> H -> 1 with ReLU, parameters W1[H][F] b1[H] W2[H][H] b2[H]
 * W3[H] b3 in one float buffer of impeller_mlp_param_count(F, H) floats,
 * trained with MSE and Adam.  All buffers are device memory; workspace holds
 * impeller_mlp_workspace_floats(F, H, B) floats.  loss_sum (nullable)
 * receives the sum of squared errors of the batch. */
unsigned int impeller_mlp_param_count(unsigned int F, unsigned int H);
size_t impeller_mlp_workspace_floats(unsigned int F, unsigned int H, unsigned int B);
hipError_t impeller_mlp_forward(const float *params, const float *X, unsigned int B,
                                unsigned int F, unsigned int H, float *Y, float *workspace,
                                hipStream_t stream);
hipError_t impeller_mlp_grad(const float *params, const float *X, const float *T,
                             unsigned int B, unsigned int F, unsigned int H, float *grads,
                             float *loss_sum, float *workspace, hipStream_t stream);
hipError_t impeller_mlp_train_step(float *params, float *m, float *v, const float *X,
                                   const float *T, unsigned int B, unsigned int F,
                                   unsigned int H, float lr, unsigned int step, float *loss_sum,
                                   float *workspace, hipStream_t stream);

/* Per-kernel HIP-event timing (off by default).  Kernel ids: 0 NTT
 * bit-reversal, 1 NTT DIT pass, 2 MSM digits, 3 MSM bucket bounds + task
 * scan, 5 MSM bucket accumulation, 6 MSM bucket reduction, 7 MSM window
 * sums, 8 MSM final combine, 9 fused gate, 10 MSM radix sort, 11 MSM
 * partial-sum tree.  read() synchronises pending events. */
void impeller_timing_enable(int on);
void impeller_timing_reset(void);
int impeller_timing_read(int kernel_id, double *total_ms, unsigned long long *count);

/* ---------------------------------------------------------------------
 * Device-runtime layer (csrc/runtime_api.hip): what the reference's Rust
 * runtime takes from its zkpoly_cuda_api crate (the .rs files of
 * cuda_api/src), over the
 * HIP runtime.  A hip_api crate binds these in place of generated CUDA
 * bindings.  Errors are returned (hipErrorInvalidValue for a broken
 * contract the reference panics on), never raised.
 * ------------------------------------------------------------------- */
typedef struct ImpRtEvent ImpRtEvent;
typedef struct ImpRtStatic ImpRtStatic;
typedef struct ImpRtPages ImpRtPages;

/* CudaStream (cuda_api/src/stream.rs:137-257): create, destroy, sync,
 * stream-ordered alloc / free, async copies (kind 1 h2d, 2 d2h, 3 d2d). */
hipError_t impeller_rt_stream_create(int device, hipStream_t *out);
hipError_t impeller_rt_stream_destroy(hipStream_t s);
hipError_t impeller_rt_stream_sync(hipStream_t s);
hipError_t impeller_rt_stream_alloc(hipStream_t s, size_t bytes, void **out);
hipError_t impeller_rt_stream_free(hipStream_t s, void *p);
hipError_t impeller_rt_memcpy_async(int device, void *dst, const void *src, size_t bytes,
                                    int kind, hipStream_t s);

/* CudaEvent / CudaEventRaw (stream.rs:1-135).  host_ordered = 1: sync and
 * stream_wait_event first block until the event has been recorded (by any
 * host thread); reset forgets the record.  elapsed: ms from a to b. */
hipError_t impeller_rt_event_create(int device, int host_ordered, ImpRtEvent **out);
hipError_t impeller_rt_event_record(ImpRtEvent *e, hipStream_t s);
hipError_t impeller_rt_event_sync(ImpRtEvent *e);
hipError_t impeller_rt_stream_wait_event(hipStream_t s, ImpRtEvent *e);
hipError_t impeller_rt_event_reset(ImpRtEvent *e);
hipError_t impeller_rt_event_elapsed(ImpRtEvent *a, ImpRtEvent *b, float *ms);
hipEvent_t impeller_rt_event_raw(ImpRtEvent *e);
hipError_t impeller_rt_event_destroy(ImpRtEvent *e);

/* alloc_pinned / free_pinned (cuda_api/src/mem.rs:18-46). */
hipError_t impeller_rt_alloc_pinned(size_t bytes, void **out);
hipError_t impeller_rt_free_pinned(void *p);

/* StaticAllocator (cuda_api/src/mem/static_allocator.rs): one device
 * allocation of max_size bytes; allocate hands out [offset, offset + bytes)
 * chosen by the caller, refusing a range that meets a live one when
 * check_overlap is set. */
hipError_t impeller_rt_static_create(int device, size_t max_size, int check_overlap,
                                     ImpRtStatic **out);
hipError_t impeller_rt_static_allocate(ImpRtStatic *a, size_t offset, size_t bytes, void **out);
hipError_t impeller_rt_static_free(ImpRtStatic *a, void *p);
void *impeller_rt_static_base(ImpRtStatic *a);
hipError_t impeller_rt_static_destroy(ImpRtStatic *a);

/* PageAllocator (cuda_api/src/mem/page_allocator.rs) over HIP virtual
 * memory: page_num device pages of page_size bytes (a multiple of
 * impeller_rt_page_granularity); allocate maps page_ids, in order, at the
 * start of a va_size reservation the first time vaid is seen and returns
 * the same address for the same vaid afterwards; with memory_check a page
 * belongs to one live address at a time, a later call for a vaid must name
 * the pages mapped there, and destroy refuses while pages are assigned.  Device pages only (the reference's host-NUMA pages are not
 * offered). */
hipError_t impeller_rt_page_granularity(int device, size_t *out);
hipError_t impeller_rt_page_create(int device, size_t page_size, size_t page_num,
                                   int memory_check, ImpRtPages **out);
hipError_t impeller_rt_page_allocate(ImpRtPages *a, size_t va_size, const size_t *page_ids,
                                     size_t n_pages, size_t vaid, void **out);
hipError_t impeller_rt_page_deallocate(ImpRtPages *a, void *p, size_t vaid);
size_t impeller_rt_page_size(ImpRtPages *a);
hipError_t impeller_rt_page_destroy(ImpRtPages *a);

/* get_num_sms / sm_version (cuda_api/src/device_info.rs): compute units and
 * the gfx target name ("gfx950"). */
hipError_t impeller_rt_num_cus(int device, int *out);
hipError_t impeller_rt_arch_name(int device, char *buf, size_t len);

/* In place of cuFile (cuda_api/src/file.rs): bytes [file_offset,
 * file_offset + bytes) of a file descriptor to / from device memory through
 * a caller's pinned staging buffer (the read overlaps file reads with
 * copies through its two halves); both return with the data in place. */
hipError_t impeller_rt_file_read_to_device(int fd, uint64_t file_offset, size_t bytes, void *dst,
                                           void *staging, size_t staging_bytes, hipStream_t s);
hipError_t impeller_rt_file_write_from_device(int fd, uint64_t file_offset, size_t bytes,
                                              const void *src, void *staging,
                                              size_t staging_bytes, hipStream_t s);

/* Library version string. */
const char *impeller_version(void);

#ifdef __cplusplus
}
#endif
#endif
