/*
 * pintt_gpu.h - C ABI of the MI355X (gfx950) exact big-integer product engine.
 *
 * Magnitudes only: little-endian arrays of 64-bit limbs (GMP's mp_limb_t
 * layout on x86-64, as returned by mpz_limbs_read), least significant first.
 * Signs, zero operands and the small-operand threshold are the caller's
 * business (the C++ HipNttMultiplier does them, like the reference wrapper at
 * src/arithmetic/int_ntt_multiplier.cpp:115-159).  No exceptions cross this
 * boundary; every entry point returns a status code (0 = success).
 *
 * Products are exact: the engine computes the cyclic convolution of 32-bit
 * coefficients modulo three primes (below 2^29 for N <= 2^22, below 2^31
 * above; transform length N = 2^k, 2^10 <= N <= 2^25 with
 * N >= 2*an + 2*bn - 1), reconstructs every coefficient with the Chinese
 * remainder theorem and propagates carries on the device.  The result is
 * bit-identical to mpz_mul.
 *
 * Thread safety: every entry point may be called from any thread; calls on one
 * device are serialised inside the library (one product at a time).
 */
#ifndef PINTT_GPU_H
#define PINTT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PINTT_OK = 0,
    PINTT_ERR_NO_DEVICE = 1,  /* no usable gfx950 device / init failed            */
    PINTT_ERR_TOO_LARGE = 2,  /* transform would exceed 2^25 (an + bn > 2^24)      */
    PINTT_ERR_NO_MEMORY = 3,  /* device allocation failed                          */
    PINTT_ERR_HIP = 4,        /* a HIP runtime call or kernel launch failed        */
    PINTT_ERR_ARGS = 5        /* null pointer, zero length, aliasing               */
};

/* Select and initialise a device (default: device 0 on first use).
 * Replaces the engine constructors gpu::IntNttEngine(int device_id)
 * (src/gpu/int_ntt.h:48) and gpu::NttEngine (src/gpu/ntt_engine.h). */
int gpu_ntt_init(int device);

/* Release all device memory, tables and streams. */
void gpu_ntt_shutdown(void);

/* out[0 .. an+bn) = a[0 .. an) * b[0 .. bn), host pointers, synchronous.
 * an, bn >= 1; out must not overlap a or b.  a == b with an == bn runs the
 * squaring path (one forward transform).
 * Replaces gpu::IntNttEngine::multiply(const uint32_t*, size_t, const uint32_t*,
 * size_t, uint32_t*, size_t, uint64_t) (src/gpu/int_ntt.h:66-69,
 * src/gpu/int_ntt.cu:301-380) and gpu::NttEngine::multiply
 * (src/gpu/ntt_engine.h:58-60, src/gpu/ntt_engine.cu:109-158), which took
 * base-2^15 / 2^12 digit vectors built on the host. */
int gpu_ntt_mul(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, uint64_t* out);

/* out[0 .. 2an) = a^2.  Same contract as gpu_ntt_mul(a, an, a, an, out).
 * Replaces Multiplier::square's default route multiply(r, a, a)
 * (src/arithmetic/multiplier.h:36-38) on the device side. */
int gpu_ntt_sqr(const uint64_t* a, size_t an, uint64_t* out);

/* Device-resident variant: d_a, d_b, d_out are device pointers on the current
 * device, `stream` a hipStream_t (NULL = the library's own stream).  The work
 * is enqueued and the call returns without waiting; products issued on any
 * stream are ordered one after another inside the library.  d_a == d_b with
 * an == bn squares.  No reference counterpart: the reference copies host digit
 * vectors on every call (src/gpu/int_ntt.cu:329-330,348-349). */
int gpu_ntt_mul_device(const uint64_t* d_a, size_t an, const uint64_t* d_b, size_t bn,
                       uint64_t* d_out, void* stream);

/* Same product, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  It calls no hipMalloc, hipFree or other
 * allocating routine: its residues, carry maps and twiddle tables live in one
 * static device array of 72.1 MiB per device (set up by gpu_ntt_init or the
 * first call of any entry point), which limits it to an + bn <=
 * PINTT_NOALLOC_MAX_LIMBS (transforms of at most 2^PINTT_NOALLOC_MAX_LOG2N
 * points); larger products return PINTT_ERR_TOO_LARGE.  Tile plans are the
 * loaded or tuned ones (never tuned here).  The work is enqueued on `stream`
 * (NULL = library stream) and not ordered with any other call: every call on
 * a device shares that device's scratch, so calls that overlap in time on one
 * device (other streams, other threads) race and must be ordered by the
 * caller; calls on one stream are. */
#define PINTT_NOALLOC_MAX_LOG2N 20
#define PINTT_NOALLOC_MAX_LIMBS (1u << (PINTT_NOALLOC_MAX_LOG2N - 1))
int gpu_ntt_mul_device_noalloc(const uint64_t* d_a, size_t an, const uint64_t* d_b, size_t bn,
                               uint64_t* d_out, void* stream);

/* Batches: count independent products out[i] = a[i] * b[i] in one call
 * (a[i] == b[i] with an[i] == bn[i] squares).  The engine buckets them by
 * transform length and runs each bucket as one launch per pass, so a level of
 * binary splitting costs four launches per distinct length instead of four
 * per product.  Host-pointer form is synchronous.  Replaces the per-product
 * calls of BinarySplitting::merge_parallel (src/engine/binary_splitting.cpp:
 * 161-293), which each paid a full engine round trip. */
int gpu_ntt_mul_batch(size_t count, const uint64_t* const* a, const size_t* an, const uint64_t* const* b,
                      const size_t* bn, uint64_t* const* out);
/* Device-pointer form: enqueued on `stream` (NULL = library stream). */
int gpu_ntt_mul_batch_device(size_t count, const uint64_t* const* d_a, const size_t* an,
                             const uint64_t* const* d_b, const size_t* bn, uint64_t* const* d_out,
                             void* stream);

/* A batch captured once into a HIP graph and replayed: the device pointers
 * and sizes are fixed at creation (the graph owns its scratch arena, carry
 * state and descriptors); gpu_ntt_graph_launch enqueues one replay on
 * `stream` (NULL = library stream), ordered with all other products.  Creating
 * a graph may tune tile plans first (it waits on the device). */
int gpu_ntt_graph_create(size_t count, const uint64_t* const* d_a, const size_t* an, const uint64_t* const* d_b,
                         const size_t* bn, uint64_t* const* d_out, void** graph);
int gpu_ntt_graph_launch(void* graph, void* stream);
size_t gpu_ntt_graph_nodes(void* graph);
void gpu_ntt_graph_destroy(void* graph);

/* Device buffers and batched copies, for callers that keep operands on the
 * device between products (binary splitting's upper tree levels).  No
 * reference counterpart: the reference copies every operand per call
 * (src/gpu/int_ntt.cu:321-350).  Upload and download go through pinned
 * staging, gathered and scattered by host threads, and return when done. */
int gpu_ntt_dev_alloc(size_t limbs, uint64_t** d_ptr);
int gpu_ntt_dev_free(uint64_t* d_ptr);  /* waits for the library stream first */
int gpu_ntt_upload(size_t count, uint64_t* const* d_dst, const uint64_t* const* h_src, const size_t* limbs);
int gpu_ntt_download(size_t count, uint64_t* const* h_dst, const uint64_t* const* d_src, const size_t* limbs);
int gpu_ntt_dev_copy(uint64_t* d_dst, const uint64_t* d_src, size_t limbs);  /* enqueued */

/* out[i] = x[i] + y[i] (max(xn, yn) + 1 limbs) or, when subtract[i], x[i] -
 * y[i] (max(xn, yn) limbs, requires x >= y; lengths may count zero top limbs):
 * magnitudes, device pointers, outputs distinct from operands.  The signed
 * R = Qr*Rl + Pl*Rr of binary-splitting merges (src/engine/binary_splitting.cpp:199,229,258,280)
 * on device-resident values.  Waits for completion; PINTT_ERR_ARGS if a
 * subtraction found x < y (the outputs are then undefined). */
int gpu_ntt_add_batch_device(size_t count, const uint64_t* const* d_x, const size_t* xn, const uint64_t* const* d_y,
                             const size_t* yn, const int* subtract, uint64_t* const* d_out);

/* Same batch enqueued without waiting (ordered with products on the library
 * stream): a subtraction with x < y sets the device word *d_negative to 1
 * (it is never cleared here; the caller zeroes and reads it) and writes
 * x - y + 2^(64 max(xn, yn)) to that output, as gpu_ntt_add_batch_device_flags
 * does.  Lets a caller enqueue a whole stretch of binary-splitting
 * levels and check once.  No reference counterpart. */
int gpu_ntt_add_batch_device_async(size_t count, const uint64_t* const* d_x, const size_t* xn,
                                   const uint64_t* const* d_y, const size_t* yn, const int* subtract,
                                   uint64_t* const* d_out, uint32_t* d_negative);

/* Same, but a subtraction with x < y is no error: negative[i] is set to 1
 * and out[i] holds x - y + 2^(64 w), w = max(xn[i], yn[i]); 0 otherwise.
 * Lets a caller try "x - y if not negative" on a whole batch (the remainder
 * corrections of a device-resident decimal conversion).  No reference
 * counterpart. */
int gpu_ntt_add_batch_device_flags(size_t count, const uint64_t* const* d_x, const size_t* xn,
                                   const uint64_t* const* d_y, const size_t* yn, const int* subtract,
                                   uint64_t* const* d_out, int* negative);

/* Decimal digits of device numbers: leaf i (xn[i] limbs at d_x[i], value
 * below 10^width[i], width[i] <= PINTT_DECIMAL_LEAF_MAX_DIGITS, xn[i] <=
 * PINTT_DECIMAL_LEAF_MAX_LIMBS) is written as exactly width[i] ASCII digits,
 * left-padded with '0', at d_out + offset[i] (inside d_out's out_len bytes).
 * A value not below 10^width[i] loses its top digits.  Waits for completion.
 * The leaves of the device-resident decimal conversion: the reference
 * converts with mpz_get_str on the host (src/io/base_converter.cpp). */
#define PINTT_DECIMAL_LEAF_MAX_DIGITS 4096
#define PINTT_DECIMAL_LEAF_MAX_LIMBS 214
int gpu_ntt_decimal_leaves_device(size_t count, const uint64_t* const* d_x, const size_t* xn, const size_t* width,
                                  const size_t* offset, uint8_t* d_out, size_t out_len);

/* The bottom of the Chudnovsky binary-splitting tree on the device: subtree i
 * covers the terms [a + 64 i, a + 64 (i + 1)) (terms at or past b count as
 * the identity P = Q = 1, R = 0), and d_out + 3 * PINTT_BS_LEAF_LIMBS * i
 * receives the magnitudes of its P, Q and R in slots of PINTT_BS_LEAF_LIMBS
 * limbs; lengths[3 i + 0..2] their exact limb counts (limbs past them are
 * unspecified), signs[2 i + 0..1]
 * the signs of P and R (+1 or -1).  Requires a < b <= 2^24.  Waits.  Replaces
 * the host subtrees of the level-synchronous tree (the reference's recursion,
 * src/engine/binary_splitting.cpp:137-160, at the bottom of the tree). */
#define PINTT_BS_LEAF_TERMS 64
#define PINTT_BS_LEAF_LIMBS 255
int gpu_ntt_bs_leaves(uint64_t a, uint64_t b, size_t count, uint64_t* d_out, uint32_t* lengths, int32_t* signs);

/* Wait for all enqueued products (PINTT_ERR_HIP if the stream reports an
 * error). */
int gpu_ntt_synchronize(void);

/* Very unbalanced operands: the longer one may be cut into up to 4 chunks that
 * share one transform of the shorter one.  mode 1 (default): when the tuned
 * timings say that is cheaper; 0 (or PINTT_SPLIT=0): never; 2..4: that many
 * chunks whenever the shapes allow it (for tests).  Products are identical
 * either way. */
void gpu_ntt_set_split(int mode);

/* Graphs made by gpu_ntt_graph_create after this call capture each
 * transform-length bucket (and the schoolbook products) as an independent
 * branch with its own working memory, so a replay may run them at the same
 * time (enable != 0, the default; the graph then holds the sum of the
 * buckets' working memory rather than the largest), or capture every launch
 * in one chain (0).  Products are identical either way.  No reference counterpart (the
 * reference has no batch launch; its caller is
 * src/engine/binary_splitting.cpp:161-293). */
void gpu_ntt_set_graph_branches(int enable);

/* log2 of the transform length an unsplit an-limb by bn-limb product uses,
 * or -1 when it exceeds 2^25. */
int gpu_ntt_transform_log2(size_t an, size_t bn);

/* How gpu_ntt_mul would run an an-limb by bn-limb product (not a square):
 * transform length 2^*log2n, the longer operand cut into *chunks pieces of
 * *chunk_limbs limbs (may be null) sharing one transform of the other.  May
 * tune tile plans first (waits on the device).  No reference counterpart: the
 * reference sizes one transform per call (src/gpu/int_ntt.cu:305-314). */
int gpu_ntt_product_plan(size_t an, size_t bn, int* log2n, int* chunks, size_t* chunk_limbs);

/* Chunk counts priced by the caller (the learned dispatcher).  For an an x bn
 * product (an >= bn, not a square) the function writes predicted costs of
 * cutting the longer operand into m = 1 .. 4 chunks to cost[m - 1] and returns
 * nonzero, or returns 0 to leave the choice to the engine's tuned per-length
 * costs; the engine takes the cheapest cut its rules allow.  It is called
 * with the engine lock held while a batch is planned and must not call into
 * this library.  NULL restores the engine's own choice.  Forced splits
 * (gpu_ntt_set_split with 0 or 2..4) take precedence.  Replaces the
 * reference's one transform length per product (src/gpu/int_ntt.cu:303-314). */
typedef int (*pintt_chunk_cost_fn)(size_t an, size_t bn, double* cost, void* ctx);
void gpu_ntt_set_chunk_cost(pintt_chunk_cost_fn fn, void* ctx);

/* Whether an an x bn product (an >= bn) can be cut into `chunks` pieces
 * (1 = not cut): returns 1, the limbs of each chunk of the longer operand
 * (*chunk_limbs, the last chunk may be shorter) and the transform length
 * 2^*log2n it then uses, else 0.  Either pointer may be null.  Pure
 * arithmetic: needs no device. */
int gpu_ntt_chunk_plan(size_t an, size_t bn, int chunks, size_t* chunk_limbs, int* log2n);

/* Largest an + bn accepted. */
size_t gpu_ntt_max_limbs(void);

/* Message for the last failure on the calling thread ("" if none). */
const char* gpu_ntt_last_error(void);

/* Tile plans.  A transform of length 2^n splits into R x C (R = 2^log2r
 * column length, C = 2^(n - log2r) row length) with 2^log2w columns per
 * column tile and 2^log2rows rows per row tile.  By default the library times
 * every compiled plan the first time a length is used and keeps the fastest
 * (set PINTT_AUTOTUNE=0 for the fixed heuristic).  Tuning waits on the device,
 * so call gpu_ntt_autotune beforehand when products will be enqueued inside a
 * stream capture.  No reference counterpart: the reference had one fixed
 * radix-2 layout (src/gpu/int_ntt.cu:192-220). */
int gpu_ntt_autotune(int log2n_min, int log2n_max);
/* The plan of a length and its tuned time *ms: milliseconds per product of
 * two N/4-limb operands (timed as a batch of 2^(22-n) products below 2^22
 * points, where one product is launch-bound), 0 if the length was not
 * tuned. */
int gpu_ntt_get_plan(int log2n, int* log2r, int* log2w, int* log2rows, float* ms);
/* Force a plan (log2r <= 0 selects the heuristic); PINTT_ERR_ARGS if that
 * tile shape is not compiled in. */
int gpu_ntt_set_plan(int log2n, int log2r, int log2w, int log2rows);
/* Writes up to max_plans (log2r, log2w, log2rows) triples; returns the count. */
int gpu_ntt_list_plans(int log2n, int* out, int max_plans);
/* Batch plans: a length bucket of a batch with 2^24 points of forward
 * transforms or more (a tree level's hundreds of products of one length) is
 * throughput-bound and takes its length's batch plan when one is set; other
 * buckets, single products among them, take the plan above.  log2r <= 0
 * clears it.  PINTT_ERR_ARGS if the tile shape is not compiled in (set) or no
 * batch plan is set (get).  Plan files hold both tables. */
int gpu_ntt_set_batch_plan(int log2n, int log2r, int log2w, int log2rows);
int gpu_ntt_get_batch_plan(int log2n, int* log2r, int* log2w, int* log2rows);
/* Tile plans (and their tuned times per product) to / from a text file, so
 * that a process can skip tuning lengths another run already tuned. */
int gpu_ntt_save_plans(const char* path);
int gpu_ntt_load_plans(const char* path);

/* Per-pass timing with HIP events on the product's stream.  When enabled,
 * every product records events around its passes; gpu_ntt_pass_times writes
 * the last product's times in milliseconds (waiting for it) in the order
 * forward-columns, row-convolution, inverse-columns, CRT+carry, and returns
 * the number of passes written. */
void gpu_ntt_set_timing(int enable);
int gpu_ntt_pass_times(double* ms, int max);

#ifdef __cplusplus
}
#endif
#endif
