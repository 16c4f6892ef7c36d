/*
 * zkgpu.h - C ABI of libzkgpu_hip.so, the MI355X (gfx950) BN254 prover core.
 *
 * Conventions
 * -----------
 * - Every entry returns 0 on success and a non-zero status on failure; the
 *   message of the calling thread's last failure is zk_last_error().
 * - Pointers named like device buffers are device addresses (from zk_alloc or
 *   any HIP allocation); "host" pointers are ordinary host memory.
 * - Field elements are 4 little-endian uint64 limbs (32 bytes). Rows called
 *   "mont" hold a*2^256 mod m (Montgomery form, R = 2^256), the layout of the
 *   reference's (n, 4) uint64 arrays (src/zkgpu/field256_gpu.py:12-20).
 *   Rows called "plain" hold the canonical value itself.
 * - Jacobian G1 points are 12 limbs (X, Y, Z), affine points 8 limbs (x, y),
 *   coordinates in Montgomery form over the base field p; the identity is
 *   (1, 1, 0) (src/zkgpu/curve_gpu.py:15-23).
 * - "bcast" arguments are 1 when that operand is a single row broadcast
 *   against the other (the reference's step-0 launch, field256_gpu.py:471).
 * - stream is a hipStream_t (NULL = the null stream); work is enqueued
 *   asynchronously unless the entry says it synchronises.
 * - field ids: ZK_FIELD_FR (BN254 scalar field r) and ZK_FIELD_FQ (BN254
 *   base field p) are built in with compile-time constants; other odd moduli
 *   below 2^255 get an id from zk_field_register.
 */
#ifndef ZKGPU_H
#define ZKGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZK_FIELD_FR 0
#define ZK_FIELD_FQ 1

/* ---- lifecycle, memory, streams (replaces src/zkgpu/gpu_backend.py:79-159:
 *      CuPy detection, NVRTC compile and CuPy's allocator/streams) ---- */
int zk_init(int device);
const char* zk_last_error(void);
int zk_device_count(int* count);
int zk_device_info(char* name, int name_len, int* compute_units, int* clock_khz);
int zk_mem_info(uint64_t* free_bytes, uint64_t* total_bytes);
int zk_alloc(void** ptr, uint64_t bytes, void* stream);
int zk_free(void* ptr);
int zk_memcpy_h2d(void* dst, const void* host_src, uint64_t bytes, void* stream);  /* synchronous for host_src */
int zk_memcpy_d2h(void* host_dst, const void* src, uint64_t bytes, void* stream);  /* synchronises stream */
int zk_memcpy_d2d(void* dst, const void* src, uint64_t bytes, void* stream);
int zk_memset(void* dst, int value, uint64_t bytes, void* stream);
/* Page-locked host memory for uploads that run at PCIe speed (the public int-taking calls stage
 * converted rows in it): zk_host_alloc returns a block of at least `bytes`, zk_host_free takes it
 * back into a small per-process cache (blocks are pinned once and reused).
 * zk_memcpy_h2d_async copies from such a block without waiting: the block must stay unchanged
 * until the stream has passed the copy. No counterpart in the reference (CuPy's pageable copies). */
int zk_host_alloc(void** ptr, uint64_t bytes);
int zk_host_free(void* ptr);
int zk_memcpy_h2d_async(void* dst, const void* pinned_src, uint64_t bytes, void* stream);
/* A plain copy kernel (each byte read and written once; 16-byte aligned buffers and sizes): the
 * known-size reference that hardware byte counters are scaled by. No counterpart in the reference. */
int zk_copy(void* dst, const void* src, uint64_t bytes, void* stream);
int zk_stream_create(void** stream);
int zk_stream_destroy(void* stream);
int zk_stream_sync(void* stream);
int zk_device_sync(void);
int zk_event_create(void** event);
int zk_event_destroy(void* event);
int zk_event_record(void* event, void* stream);
int zk_event_elapsed_ms(void* start, void* end, float* ms);
/* *done = 1 once every piece of work before the event's record has finished, else 0; never blocks. */
int zk_event_query(void* event, int* done);
int zk_event_sync(void* event);
int zk_stream_wait_event(void* stream, void* event);
/* HIP graphs over a created stream: everything enqueued between begin and end (e.g. plan
 * launches, which allocate nothing) is recorded once and replayed by zk_graph_launch (no
 * counterpart in the reference, which launches from Python every time). */
int zk_graph_begin(void* stream);
int zk_graph_end(void* stream, void** graph_exec);
int zk_graph_launch(void* graph_exec, void* stream);
int zk_graph_destroy(void* graph_exec);
/* Opt-in per-kernel timing with HIP events on each launch's stream (no counterpart in the
 * reference, whose benchmarks time whole calls, benchmarks/_harness.py:67-112):
 * zk_prof_read sums the recorded durations of the kernels named name ("msm_accumulate",
 * "ntt_pass", ...) and optionally forgets all records. */
int zk_prof_enable(int on);
int zk_prof_read(const char* name, double* total_ms, int64_t* count, int reset);
/* Measured throughput of v_mad_u64_u32 (32x32+64-bit multiply-add, the instruction every
 * field product is built from) on the current GPU, in 10^12 per second: the VALU roofline
 * of the field, NTT and MSM kernels (no counterpart in the reference). */
int zk_valu_peak(double* mads_per_second_t);
/* Measured throughput of whole 9-limb Montgomery products of the Fq field (independent chains on
 * every lane, no memory traffic), in 10^9 products per second: the ceiling the field arithmetic
 * itself sets for the NTT and MSM kernels, with its instructions beside the multiply-adds (no
 * counterpart in the reference). */
int zk_product_peak(double* gproducts_per_s);
/* Probe kernel: replaces the CuPy probe (gpu_backend.py:94-99). */
int zk_selftest(void);

/* ---- field arithmetic (replaces the kernels zk256_mul / zk256_add / zk256_sub,
 *      src/zkgpu/field256_gpu.py:172-191, and Field256's Montgomery constants,
 *      field256_gpu.py:277-288) ---- */
/* Registers an odd modulus in [3, 2^255) (plain limbs) and returns its id. */
int zk_field_register(const uint64_t modulus[4], int* field_id);
/* mont[i] = (plain[i] mod m) * 2^256 mod m, for any plain[i] < 2^256 (Field256.from_ints, :297-310). */
int zk_field_to_mont(int field_id, const uint64_t* plain, uint64_t* mont, int64_t n, void* stream);
/* plain[i] = mont[i] * 2^-256 mod m (Field256.to_ints, :312-319). */
int zk_field_from_mont(int field_id, const uint64_t* mont, uint64_t* plain, int64_t n, void* stream);
/* out[i] = a[i] * b[i] * 2^-256, a[i] + b[i], a[i] - b[i] mod m (Field256.mul/add/sub, :325-341). */
int zk_field_mul(int field_id, const uint64_t* a, int a_bcast, const uint64_t* b, int b_bcast, uint64_t* out, int64_t n, void* stream);
int zk_field_add(int field_id, const uint64_t* a, int a_bcast, const uint64_t* b, int b_bcast, uint64_t* out, int64_t n, void* stream);
int zk_field_sub(int field_id, const uint64_t* a, int a_bcast, const uint64_t* b, int b_bcast, uint64_t* out, int64_t n, void* stream);

/* ---- NTT (replaces zk256_bit_reverse + zk256_ntt_stage and Field256.ntt / _ntt_tables,
 *      src/zkgpu/field256_gpu.py:198-234,347-387,492-508; and the coset transforms of
 *      Radix2Domain, src/zkgpu/domain.py:112-135) ---- */
/* out = DFT_n(in) with omega (plain) a primitive 2^log_n-th root, natural order in and out.
 * inverse != 0: in is transformed with omega and scaled by n^-1 (pass omega^-1 for an inverse NTT).
 * coset (plain, or NULL): forward multiplies in[j] by coset^j first; inverse multiplies out[j]
 * by coset^-j afterwards. in and out are mont rows; in may equal out. Tables are cached per
 * (field, log_n, omega, coset). */
int zk_ntt(int field_id, const uint64_t* in, uint64_t* out, int log_n, int inverse,
           const uint64_t omega[4], const uint64_t* coset, void* stream);
/* Frees the cached tables that no plan holds (plans, and graphs recorded from them, keep theirs
 * until destroyed); the cache also evicts such tables, least recently used first, beyond 4 GiB.
 * Replaces _ntt_tables.cache_clear() (field256_gpu.py:492). zk_ntt_cache_stats reports the
 * cached tables, their device bytes and how many are held by plans. */
int zk_ntt_cache_clear(void);
int zk_ntt_cache_stats(uint64_t* tables, uint64_t* bytes, uint64_t* pinned);
/* Radix-2 stages fused per LDS-staged pass for transforms of size 2^log_n (1..10; 0 restores
 * the default of 8), as picked by the autotuner; plans created afterwards use it. No counterpart
 * in the reference, whose NTT runs one kernel launch per stage (src/zkgpu/ntt.py:149-203). */
int zk_ntt_set_fusion(int log_n, int max_stages);
int zk_ntt_get_fusion(int log_n, int* max_stages);
/* Groth16 quotient over Fr (demo/prover_simulation.py:142-150 with coset shift g):
 * h = coset_intt((coset_ntt(intt a) * coset_ntt(intt b) - coset_ntt(intt c)) * (g^n - 1)^-1).
 * a, b, c, h: n mont rows each (h may alias none of them). */
int zk_groth16_quotient(const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t* h,
                        int log_n, const uint64_t shift[4], void* stream);
/* The quotient above for count jobs of mixed sizes as one launch sequence: every job's three
 * inverse NTTs (with their coset post-scale) in one launch per pass index, then every forward
 * NTT, then every final inverse NTT with the fused (A B - C) / Z load. Job j reads a[j], b[j],
 * c[j] (2^log_n[j] mont rows each) and writes h[j]; shifts holds count plain 4-limb coset
 * shifts. The plan holds the coset evaluations and pass scratch; launching allocates nothing.
 * No counterpart in the reference, which runs one job's transforms one call at a time
 * (demo/prover_simulation.py:142-150). */
int zk_groth16_plan_create(int count, const uint64_t* const* a, const uint64_t* const* b,
                           const uint64_t* const* c, uint64_t* const* h, const int32_t* log_n,
                           const uint64_t* shifts, void** plan, void* stream);
int zk_groth16_plan_launch(void* plan, void* stream);
int zk_groth16_plan_destroy(void* plan);
/* Device bytes zk_groth16_plan_create holds for these sizes (for admission by HBM budget). */
int zk_groth16_workspace_bytes(int count, const int32_t* log_n, uint64_t* bytes);
/* Several Fr NTTs of different sizes as one launch per pass index (every transform's pass p
 * in the same launch). Per job j: in[j], out[j] (device addresses, in may equal out), log_n[j]
 * in 1..28, inverse[j] (the inverse includes the n^-1 scale); roots from the generator 5, as
 * Field256.ntt (field256_gpu.py:347-387). A plan allocates all scratch and uploads the work
 * tables; launching it allocates nothing (capturable into a HIP graph). */
int zk_ntt_plan_create(int count, const uint64_t* const* in, uint64_t* const* out, const int32_t* log_n,
                       const int32_t* inverse, void** plan, void* stream);
int zk_ntt_plan_launch(void* plan, void* stream);
int zk_ntt_plan_destroy(void* plan);
/* Plan, launch and release in one call. */
int zk_ntt_batch(int count, const uint64_t* const* in, uint64_t* const* out, const int32_t* log_n,
                 const int32_t* inverse, void* stream);

/* ---- G1 point kernels (replaces zk_g1_add / zk_g1_add_affine / zk_g1_double,
 *      src/zkgpu/curve_gpu.py:236-274; same EFD formulas, same (X, Y, Z)) ---- */
int zk_g1_add(const uint64_t* a, int a_bcast, const uint64_t* b, int b_bcast, uint64_t* out, int64_t n, void* stream);
int zk_g1_add_affine(const uint64_t* a, int a_bcast, const uint64_t* b_affine, int b_bcast, uint64_t* out, int64_t n, void* stream);
int zk_g1_double(const uint64_t* a, uint64_t* out, int64_t n, void* stream);
/* Replaces the host-side Point(x, y) validation in msm_pippenger_gpu (msm_gpu.py:441-442):
 * *first_bad (host) = smallest i whose affine mont point is off y^2 = x^3 + 3, or -1. Synchronises. */
int zk_g1_on_curve(const uint64_t* points_affine, int64_t n, int64_t* first_bad, void* stream);
/* The same check without waiting: the first bad index, or n when every point is on the curve, is
 * written to the device word *first_bad_dev in stream order (so a caller can go on converting
 * other inputs and read the verdict later). */
int zk_g1_on_curve_async(const uint64_t* points_affine, int64_t n, uint64_t* first_bad_dev, void* stream);
/* out_affine[i] = k[i] * G for plain scalars k[i] < 2^256 (all-zero rows for the identity);
 * the SRS generator behind the device-resident basis (no counterpart: the reference builds
 * its bases with Python scalar multiplications, benchmarks/bench_msm.py:40-44). */
int zk_g1_fixed_base(const uint64_t* k, int64_t n, uint64_t* out_affine, void* stream);

/* ---- Pippenger MSM (replaces msm_pippenger_gpu's device work and the CPU window
 *      combination, src/zkgpu/msm_gpu.py:336-377,447-449 and msm.py:156-169) ---- */
/* out_jac (device, 12 limbs) = sum_i s_i * P_i as a Jacobian mont point, for n affine mont points
 * and plain scalars s_i < 2^256 (reduced mod r on the device). window: bits per signed digit
 * (0 = automatic); chunk_len: bucket entries per thread in the accumulation (0 = automatic). */
int zk_msm_g1(const uint64_t* points_affine, const uint64_t* scalars, int64_t n, int window,
              int chunk_len, uint64_t* out_jac, void* stream);
/* zk_msm_g1 that also returns in *entries the number of (bucket, point) entries its bucket
 * accumulation processed: the nonzero signed digits of the scalars' GLV halves, the unit of work
 * a measurement prices the accumulation by. Synchronises stream. No counterpart in the reference
 * (measurement only). */
int zk_msm_g1_entries(const uint64_t* points_affine, const uint64_t* scalars, int64_t n, int window,
                      int chunk_len, uint64_t* out_jac, uint64_t* entries, void* stream);
/* zk_msm_g1 with no allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels and Jacobian result as zk_msm_g1; every device buffer of the
 * call (point records, digits, sort, buckets, run summaries, reduction) lives in one file-scope
 * device array of zk_msm_noalloc_max_bytes bytes (64 MiB) that each GPU holds until the process
 * ends, so only MSMs whose workspace fits are accepted (2^14 points with the automatic window and
 * chunk; larger ones fail with an error naming the bytes needed). Calls on one GPU share that array
 * and keep the last call's contents: calls that overlap on one GPU (two streams, two threads) race. */
int zk_msm_g1_noalloc(const uint64_t* points_affine, const uint64_t* scalars, int64_t n, int window,
                      int chunk_len, uint64_t* out_jac, void* stream);
int zk_msm_noalloc_max_bytes(uint64_t* bytes);
/* The signed-digit window zk_msm_g1 picks for n points when window = 0. */
int zk_msm_default_window(int64_t n, int* window);
/* Whether plans created afterwards split each scalar k = k1 + k2 lambda (|k1|, |k2| < 2^127)
 * over P_i and phi(P_i) = (beta x_i, y_i) (on by default; off gives plain 254-bit digits).
 * No counterpart in the reference, whose Pippenger recodes the full scalar (msm.py:112-207). */
int zk_msm_set_glv(int on);
int zk_msm_get_glv(int* on);
/* The bucket accumulation kernel plans created afterwards use (picked by the autotuner):
 * bit 0 = the low-latency product schedule, bit 1 = XYZZ bucket coordinates (madd-2008-s);
 * 0 = Jacobian madd-2007-bl with serial column products. No counterpart in the reference. */
int zk_msm_set_accumulate(int variant);
int zk_msm_get_accumulate(int* variant);
/* log2 of the bucket-reduction segment length L of plans created afterwards (picked by the
 * autotuner): each thread forms the running sums of L consecutive buckets, the segments' sums
 * then meet in log2(buckets / L) bit rows; 0 = automatic (L = 4 for windows of 14 bits and
 * more, else L = 2). The reference's segments are a
 * fixed length and their sums are scaled one by one (msm_gpu.py:196-223). */
int zk_msm_set_segment(int log_seg);
int zk_msm_get_segment(int* log_seg);
/* A device-resident basis: n affine mont points converted once into the kernels' internal
 * layout and kept in HBM (the shared SRS of batched and served MSMs; no counterpart in the
 * reference, which converts its points on every call, msm_gpu.py:348). */
int zk_srs_create(const uint64_t* points_affine, int64_t n, void** srs, void* stream);
int zk_srs_destroy(void* srs);
/* Adds a window table of width c (4..20) to an SRS: 2^(c w) P_i for every window w of a 254-bit
 * scalar, W = floor(254 / c) + 1 windows, W x 64 bytes per point of HBM. MSM plans over the SRS
 * then send every window's digits to one shared set of 2^(c-1) buckets (one bucket reduction,
 * no window combination), each job using the table with the least work for its size, or the
 * table of its requested window. Fails while plans over the SRS exist. No counterpart in the
 * reference, which has no resident basis. */
int zk_srs_precompute(void* srs, int window, void* stream);
/* A batch of count MSMs over prefixes of one SRS: job j sums n[j] <= SRS size points with the
 * plain scalars at scalars[j] (device address, n[j] rows) and writes its Jacobian mont result to
 * out_jac + 12 * j. window[j] may be 0 (automatic) or window may be NULL; chunk_len 0 = automatic.
 * Creating a plan allocates every buffer the batch needs; launching it enqueues a fixed kernel
 * sequence that allocates nothing, so it can be captured into a HIP graph. The plan reads the
 * scalar buffers at launch time. Replaces the per-call structure of msm_pippenger_gpu
 * (msm_gpu.py:336-377) with one sort, one accumulation and one reduction for all jobs. */
int zk_msm_plan_create(void* srs, int count, const uint64_t* const* scalars, const int64_t* n,
                       const int32_t* window, int chunk_len, uint64_t* out_jac, void** plan, void* stream);
int zk_msm_plan_launch(void* plan, void* stream);
int zk_msm_plan_destroy(void* plan);
/* Device bytes zk_msm_plan_create holds for these sizes, windows and chunk_len over srs (for
 * admission by HBM budget; the same window-table choice as the plan). srs may be NULL: plans of
 * zk_msm_g1 over raw points. */
int zk_msm_workspace_bytes(void* srs, int count, const int64_t* n, const int32_t* window, int chunk_len, uint64_t* bytes);
/* Device bytes a created plan holds. */
int zk_msm_plan_bytes(void* plan, uint64_t* bytes);
/* Plan, launch and release in one call. */
int zk_msm_g1_batch(void* srs, int count, const uint64_t* const* scalars, const int64_t* n,
                    const int32_t* window, uint64_t* out_jac, void* stream);

/* ---- 64-bit prime fields: int64 elements modulo 2 <= mod < 2^63 ----
 * Operands are device arrays of any integer dtype (codes below), widened to int64 as NumPy's
 * astype(int64) does (uint64 keeps its bit pattern) and floor-reduced into [0, mod) on load.
 * Broadcasting is by strides: shape has ndim <= 8 dimensions (the output's, C order), strides
 * are in elements per dimension and 0 where an operand is broadcast. Outputs are contiguous
 * int64 in [0, mod). */
#define ZK_I8 0
#define ZK_I16 1
#define ZK_I32 2
#define ZK_I64 3
#define ZK_U8 4
#define ZK_U16 5
#define ZK_U32 6
#define ZK_U64 7
/* out = a * b (op 0) or a + b (op 1) mod mod. Replaces the zk_modmul / zk_modadd elementwise
 * kernels (src/zkgpu/field_gpu.py:73-101) behind gpu_batch_mul (:147) and gpu_batch_add (:181). */
int zk_f64_binary(int op, const void* a, int a_dtype, const int64_t* a_strides, const void* b, int b_dtype,
                  const int64_t* b_strides, int ndim, const int64_t* shape, int64_t* out, uint64_t mod,
                  void* stream);
/* out = base^e mod mod for e >= 0 given as exp_len little-endian 64-bit words (e = 0 gives 1).
 * Replaces the launch-per-bit square-and-multiply of _gpu_pow_nonnegative
 * (src/zkgpu/field_gpu.py:135-144) behind gpu_batch_pow (:199). */
int zk_f64_pow(const void* base, int dtype, const int64_t* strides, int ndim, const int64_t* shape,
               const uint64_t* exp_words, int exp_len, int64_t* out, uint64_t mod, void* stream);
/* *result = 1 if some element of base is 0 mod mod, else 0; synchronises stream. Replaces the
 * invertibility check of gpu_batch_pow for negative exponents (src/zkgpu/field_gpu.py:221-223). */
int zk_f64_any_zero(const void* base, int dtype, const int64_t* strides, int ndim, const int64_t* shape,
                    uint64_t mod, int* result, void* stream);
/* The NTT X_k = sum_j in_j omega^(jk) of 2^log_n values (log_n <= 30) modulo an odd prime
 * mod < 2^63, omega a primitive 2^log_n-th root of unity; invert uses omega^-1 and scales by
 * n^-1. in (contiguous, any integer dtype, reduced on load) is left unchanged; out (contiguous
 * int64, natural order) must be another buffer. Moduli below 2^31 run on 32-bit arithmetic.
 * Twiddle tables are cached per (mod, omega, log_n, invert) with the other NTT tables. Replaces ntt_gpu's per-stage CuPy butterflies,
 * bit-reversal gather and cached twiddle table (src/zkgpu/ntt_gpu.py:48-155). */
int zk_ntt64(const void* in, int dtype, int64_t* out, int log_n, uint64_t mod, uint64_t omega, int invert,
             void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ZKGPU_H */
