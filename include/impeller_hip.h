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
 * omega^i for i < 2^(log_len-1), device memory.  Asynchronous on stream.
 * The two thread-count hints are accepted and ignored: the pass split comes
 * from the library's planner/autotuner. */
hipError_t ssip_ntt(PolyPtr x, const unsigned int *twiddle, unsigned int log_len,
                    hipStream_t stream, const unsigned int max_threads_stage1_log,
                    const unsigned int max_threads_stage2_log);

/* Replaces core/src/ntt/src/ntt.h:9 (ssip_precompute).  Writes unit^i,
 * i < len, into the HOST buffer twiddle (callers pass len = n/2,
 * core/src/ntt.rs:224-244).  Synchronous. */
hipError_t ssip_precompute(unsigned int *twiddle, unsigned int len, const unsigned int *unit);

/* Replaces core/src/ntt/src/ntt.h:12 (recompute_ntt).  Same transform as
 * ssip_ntt with twiddles derived from omegas_d[i] = omega^(2^i) (i < 32);
 * pq_d / pq_deg follow gen_pq_omegas.  Needs no n/2 twiddle table. */
hipError_t recompute_ntt(PolyPtr x, const unsigned int *pq_d, unsigned int pq_deg,
                         const unsigned int *omegas_d, unsigned int log_len,
                         hipStream_t stream, const unsigned int max_threads_stage1_log,
                         const unsigned int max_threads_stage2_log);

/* Replaces core/src/ntt/src/ntt.h:15 (gen_pq_omegas), host memory:
 * pq[i] = (unit^(len >> pq_deg))^i for i < 2^(pq_deg-1); omegas[i] =
 * unit^(2^i) for i < 32 (core/src/ntt/src/precompute.cuh:89-112). */
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
 * device temp bytes each card needs and nothing else happens.  This symbol
 * uses the library's default window configuration (window 16, no precompute
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
 * core/src/fused_kernels/kernel_gen.rs:309-330, for the gate
 * out = a*b*c - d + q*e*(1 - q) over Fr.  vars (device array of 6
 * ConstPolyPtr) = {a, b, c, d, q, e}; mut_vars (device array of 1 PolyPtr)
 * = {out}; len elements; local_buffer unused (no spills).  Asynchronous. */
hipError_t fused_gate_abc_d_qe(const ConstPolyPtr *vars, const PolyPtr *mut_vars,
                               unsigned long long len, const PolySlice *slice_info,
                               void *local_buffer, hipStream_t stream);

/* ======================= device-resident extensions ====================== */
/* No reference analogue: entry points for callers that keep everything in
 * HBM (the batched / graph-captured proving path and the benchmarks). */

/* Forward (inverse = 0) or inverse (inverse = 1, result scaled by n^-1) NTT
 * of a contiguous device array, twiddle = omega^i table (i < n/2) of the
 * omega actually used (omega^-1 for the inverse).  coset_shift (nullable,
 * device, 1 element) multiplies x[i] by shift^i before a forward transform
 * or by shift^-i after an inverse one (shift passed as its inverse then).
 * plan_passes: pass split (level counts, summing to log_len) or NULL for
 * the library default. */
hipError_t impeller_ntt(uint32_t *x, const uint32_t *twiddle, unsigned int log_len,
                        int inverse, const uint32_t *coset_shift,
                        const unsigned int *plan_passes, unsigned int n_passes,
                        hipStream_t stream);

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

/* Seeded synthetic inputs on the device (the CPU checker has the same
 * definition): kind 0 uniform Fr, 1 selector-like (90% zero, else one),
 * 2 small (< 2^16).  Bases: P_i = k_i * G with k_i = uniform Fr (seed). */
hipError_t impeller_gen_scalars(uint32_t *out_d, unsigned long long n, int kind,
                                unsigned long long seed, unsigned long long start,
                                hipStream_t stream);
hipError_t impeller_gen_bases(uint32_t *out_d, unsigned long long n, unsigned long long seed,
                              unsigned long long start, hipStream_t stream);

/* Device field helpers used by tests of the arithmetic itself:
 * op 0 mul, 1 add, 2 sub, 3 neg(a), 4 from_number(a), 5 to_number(a),
 * 6 pow(a, b as integer), 7 inv(a); field 0 = Fr, 1 = Fq; n elements. */
hipError_t impeller_field_op(int field, int op, uint32_t *r_d, const uint32_t *a_d,
                             const uint32_t *b_d, unsigned long long n, hipStream_t stream);
/* G1 helpers: op 0 madd(xyzz, aff), 1 add(xyzz, xyzz), 2 dbl(xyzz),
 * 3 to_affine(xyzz) -> aff, 4 scalar_mul(aff, integer) -> aff; n items. */
hipError_t impeller_g1_op(int op, uint32_t *r_d, const uint32_t *a_d, const uint32_t *b_d,
                          unsigned long long n, hipStream_t stream);

/* Per-kernel HIP-event timing (off by default).  Kernel ids: 0 NTT
 * bit-reversal, 1 NTT DIT pass, 2 MSM digits, 3 MSM scan, 4 MSM scatter,
 * 5 MSM bucket accumulation, 6 MSM bucket reduction, 7 MSM window sums,
 * 8 MSM final combine, 9 fused gate.  read() synchronises pending events. */
void impeller_timing_enable(int on);
void impeller_timing_reset(void);
int impeller_timing_read(int kernel_id, double *total_ms, unsigned long long *count);

/* Library version string. */
const char *impeller_version(void);

#ifdef __cplusplus
}
#endif
#endif
