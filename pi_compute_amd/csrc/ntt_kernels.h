// ntt_kernels.h - device side of the three-prime NTT product (gfx950).
//
// Transform layout (length N = R * C, four-step without transposes):
//   forward  = log2(R) Cooley-Tukey layers on every column (stride C), then
//              log2(C) layers on every row; the result sits in bit-reversed
//              order, which the pointwise product does not care about;
//   inverse  = the same layers undone in reverse order (Gentleman-Sande), so
//              natural order comes back without any bit-reversal pass.
// Layer s of the length-N transform splits block b (of size N / 2^s) with the
// twiddle Z[2^s + b] = w_{2^(s+1)}^bitrev_s(b).  That table does not depend on
// N, so one table of 2^25 (w, w') pairs per prime serves every length.  The
// first log2(R) layers only pair elements at distances that are multiples of C,
// i.e. inside one column; the rest stay inside one row (block = row index).
//
// Each workgroup owns a tile in LDS; a thread keeps 16 elements in registers
// and runs up to four layers on them between LDS exchanges (radix-16 steps).
//
// Replaces: bit_reverse_kernel, ntt_butterfly_kernel (one launch per layer),
// pointwise_mod_mul_kernel, scale_kernel (src/gpu/int_ntt.cu:57-104,192-220)
// and the host CRT + carry loop (src/gpu/int_ntt.cu:226-245,352-377).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "modarith.h"

namespace pintt {

struct Twiddles {
    const uint2* fwd[kPrimes];  // (w, floor(w 2^32 / p)) in Cooley-Tukey order
    const uint2* inv[kPrimes];  // inverses, same indexing
    const uint32_t* row[kPrimes];  // row factor tables (lo, then hi pairs; ntt_device.h)
};

// Constants for the CRT at one transform length N.  The residues coming out of
// the inverse transform equal N * x * 2^-32 (Montgomery pointwise product), so
// the first multiplier of each prime also carries 2^32 / N.  The wide set uses
// Garner's form (k0 .. c21); the narrow set the explicit sum
// x = sum_i y_i M_i - k M, M_i = M / p_i, y_i = r_i 2^32 N^-1 M_i^-1 mod p_i
// (e), k from a single-precision estimate of sum_i y_i / p_i (inv), with the
// M_i split into 32-bit halves (mlo, mhi) and -M mod 2^96 as three words (negm).
struct CrtConst {
    uint2 k0;        // 2^32 N^-1                  mod p0
    uint2 k1;        // 2^32 N^-1 p0^-1            mod p1
    uint2 c1;        // p0^-1                      mod p1
    uint2 k2;        // 2^32 N^-1 (p0 p1)^-1       mod p2
    uint2 c20;       // (p0 p1)^-1                 mod p2
    uint2 c21;       // p1^-1                      mod p2
    uint2 e[kPrimes];       // 2^32 N^-1 M_i^-1 mod p_i, with its Shoup quotient
    uint32_t mlo[kPrimes];  // M_i mod 2^32
    uint32_t mhi[kPrimes];  // M_i div 2^32
    float inv[kPrimes];     // 1 / p_i rounded to single precision
    uint32_t negm[3];       // 2^96 - M, least significant word first
};

constexpr int kElemsPerThread = 16;
constexpr int kMinLog2N = 10;
constexpr int kCarryWordsPerThread = 8;
// A carry workgroup of T threads sums 8 T words: the kCarryWindowWords before
// its block, whose sums almost always fix the block's carry-in, then the
// carry_block_words(T) words of the block, which it writes.  T is
// carry_threads() (64, 128 or 256; ntt_carry.hip).
constexpr int kCarryWindowWords = 32;
constexpr int kCarryMinThreads = 64;
__host__ __device__ constexpr uint32_t carry_block_words(int threads) {
    return (uint32_t)(threads * kCarryWordsPerThread - kCarryWindowWords);
}
int carry_threads();
constexpr int kMaxChunks = 4;  // pieces the longer operand of one product may be cut into

// Work descriptors.  One launch of a pass processes a bucket of products that
// share the transform length N; each product is a group: its longer operand
// cut into m chunks (m = 1 unless the operands are very unbalanced), all
// multiplied by the one transform of the other operand, and the chunk
// convolutions summed at their offsets by the carry pass.  Residues live in
// arena slots of 3 planes x N words.

struct FwdJob {             // forward column pass of one operand chunk
    const uint32_t* src;    // 32-bit coefficients, little endian
    uint32_t ncoef;         // coefficients present; the rest of the slot is zero
    uint32_t* dst;          // arena slot
};

// Row pass of one chunk product a_i * b (or a_i^2), or of two, a_i * b and
// a_(i+1) * b, sharing b's row layers and the twiddles.  Launches are split by
// kind: squares, products whose b still needs its row layers (the first chunk
// of a group, or its first two chunks as a pair; b's transformed rows are
// written back when more chunks follow), and products whose b slot already
// holds transformed rows.
enum RowKind { kRowSquare = 0, kRowTransformB = 1, kRowReuseB = 2, kRowPair = 3 };
struct RowJob {
    uint32_t* a;       // chunk slot, overwritten by the chunk product
    uint32_t* b;       // slot of the shared operand (unused when squaring)
    uint32_t* a2;      // kRowPair: the second chunk slot
    uint32_t write_b;  // kRowTransformB, kRowPair: store b's transformed rows back
};

struct InvJob {             // inverse column pass of one chunk product, in place
    uint32_t* buf;
};

struct CarryJob {           // CRT + carry of one group
    const uint32_t* res[kMaxChunks];  // chunk product slots
    uint32_t off[kMaxChunks];         // coefficient offset of each chunk product (multiple of 32)
    uint32_t len[kMaxChunks];         // length of each chunk product (<= N)
    uint32_t m;
    uint32_t blk0;                    // first carry block of the group in this launch
    uint32_t* out;                    // product words
    uint64_t nwords;
    uint32_t* maps;                   // one carry map per carry block of the group; the groups of a
                                      // launch have theirs one after another (maps - blk0 is the same)
};

// Row/column split and tile shapes for a transform of length 2^n.
struct Plan {
    int log2n, log2r, log2c;
    int log2w;     // columns per workgroup in the column passes
    int log2rows;  // rows per workgroup in the row pass
};

__host__ inline Plan make_plan(int log2n) {
    Plan pl{};
    pl.log2n = log2n;
    if (log2n <= 24) {
        pl.log2r = log2n / 2 < 10 ? log2n / 2 : 10;
    } else {
        pl.log2r = 11;
    }
    pl.log2c = log2n - pl.log2r;
    if (pl.log2r == 11) {
        pl.log2w = 3;
    } else {
        int w = 12 - pl.log2r;  // 4096 / R elements wide
        if (w < 4) w = 4;
        if (w > pl.log2c) w = pl.log2c;
        pl.log2w = w;
    }
    int rows = 12 - pl.log2c;
    if (rows < 0) rows = 0;
    if (rows > pl.log2r) rows = pl.log2r;
    pl.log2rows = rows;
    return pl;
}

// Tile shapes compiled into the library per prime set (ntt_cols.hip / ntt_rows.hip).
bool col_case_available(int set, int log2r, int log2w);
bool row_case_available(int set, int log2c, int log2rows);
// Pair row kernels hold three tiles in LDS: rows of at most 2^13 words per workgroup.
bool pair_case_available(int set, int log2c, int log2rows);
// Every row kernel of a plan scales its rows and stores its products below 2p.
bool rows_scaled(const Plan& pl);

// A plan with an explicit split, checked against the tile shapes compiled for
// the prime set of its length.
__host__ inline bool plan_valid(const Plan& pl) {
    const int set = set_for(pl.log2n);
    return pl.log2r + pl.log2c == pl.log2n && pl.log2w <= pl.log2c && pl.log2rows <= pl.log2r &&
           col_case_available(set, pl.log2r, pl.log2w) && row_case_available(set, pl.log2c, pl.log2rows);
}

// A launch's job list: up to kInlineJobs jobs travel inside the kernel
// arguments (no descriptor upload, so a single product is only kernels on the
// stream); longer lists are read from device memory.
constexpr int kInlineJobs = 4;
template <class T>
struct JobList {
    const T* dev;   // device array (count jobs) when ninl == 0
    uint32_t ninl;  // jobs held in `inl`
    T inl[kInlineJobs];
};
template <class T>
inline JobList<T> device_jobs(const T* dev) {
    JobList<T> l{};
    l.dev = dev;
    return l;
}
template <class T>
inline JobList<T> inline_jobs(const T* host, uint32_t n) {
    JobList<T> l{};
    l.ninl = n;
    for (uint32_t i = 0; i < n; ++i) l.inl[i] = host[i];
    return l;
}

// Job z of a launch (z uniform across the workgroup).  A statement expression
// over the kernel argument itself: passing the list by reference to a function
// would make the compiler copy the kernel argument block to scratch.
#define PINTT_JOB(T, list, z)                                                        \
    ({                                                                               \
        const uint32_t z_ = __builtin_amdgcn_readfirstlane(z);                      \
        T r_;                                                                        \
        if ((list).ninl == 0) {                                                      \
            r_ = (list).dev[z_];                                                     \
        } else {                                                                     \
            r_ = (list).inl[0];                                                      \
            _Pragma("unroll") for (int i_ = 1; i_ < kInlineJobs; ++i_) if (z_ == (uint32_t)i_) r_ = (list).inl[i_]; \
        }                                                                            \
        r_;                                                                          \
    })

// A small product of a batch (schoolbook.hip): out (an + bn limbs) = a * b,
// an, bn <= kSchoolMaxLimbs; a == b squares.
constexpr int kSchoolThreads = 256;
constexpr int kSchoolMaxLimbs = 512;
struct SchoolJob {
    const uint64_t* a;
    const uint64_t* b;
    uint64_t* out;
    uint32_t an, bn;
};

// Launch entry points (ntt_cols.hip, ntt_rows.hip, ntt_carry.hip,
// schoolbook.hip); counts are host values.
// Tables for one prime of a set: the first 2^log2count entries (log2count <=
// kSetMaxLog2N[set]) of the forward and inverse tables, which are the whole
// tables of every transform of at most 2^log2count points, and the full row
// factor table.  pow_scratch: device room for kPowScratch words (powers of the
// root, used on the stream).
constexpr int kPowLoBits = 12;
constexpr uint32_t kPowLo = 1u << kPowLoBits;
constexpr uint32_t kPowScratch = kPowLo + (1u << (kMaxLog2N - kPowLoBits)) + 1;
// Words of the row factor table of one prime of a set.
__host__ __device__ constexpr uint32_t row_table_words(int set) {
    return kPowLo + 2u * (1u << (kSetMaxLog2N[set] - kPowLoBits));
}
hipError_t launch_build_twiddles(uint2* fwd, uint2* inv, uint32_t* row, uint32_t* pow_scratch, int set, int prime,
                                 int log2count, hipStream_t s);
hipError_t launch_forward_columns(const Plan& pl, const JobList<FwdJob>& jobs, uint32_t njobs, const Twiddles& tw,
                                  hipStream_t s);
hipError_t launch_row_convolution(const Plan& pl, const JobList<RowJob>& jobs, uint32_t njobs, RowKind kind,
                                  const Twiddles& tw, hipStream_t s);
hipError_t launch_inverse_columns(const Plan& pl, const JobList<InvJob>& jobs, uint32_t njobs, const Twiddles& tw,
                                  hipStream_t s);
hipError_t launch_schoolbook(const JobList<SchoolJob>& jobs, uint32_t njobs, hipStream_t s);
hipError_t launch_crt_carry(const JobList<CarryJob>& jobs, uint32_t njobs, uint32_t nblocks, int log2n,
                            const CrtConst& cc, hipStream_t s);

}  // namespace pintt
