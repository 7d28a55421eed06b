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
};

// Constants for Garner's CRT at one transform length N.  The residues coming
// out of the inverse transform equal N * x * 2^-32 (Montgomery pointwise
// product), so the first multiplier of each prime also carries 2^32 / N.
struct CrtConst {
    uint2 k0;        // 2^32 N^-1                  mod p0
    uint2 k1;        // 2^32 N^-1 p0^-1            mod p1
    uint2 c1;        // p0^-1                      mod p1
    uint2 k2;        // 2^32 N^-1 (p0 p1)^-1       mod p2
    uint2 c20;       // (p0 p1)^-1                 mod p2
    uint2 c21;       // p1^-1                      mod p2
};

constexpr int kElemsPerThread = 16;
constexpr int kMinLog2N = 10;
constexpr int kCarryThreads = 256;
constexpr int kCarryWordsPerThread = 32;
constexpr int kCarryWordsPerBlock = kCarryThreads * kCarryWordsPerThread;

// Device-side state of the carry pass: block tickets and the epoch that tags
// status words (reset/incremented on the device by the inverse column pass of
// each product, so a captured graph can be replayed).  `error` is set if a look-back ever gave up.
struct CarrySync {
    unsigned long long ticket;
    uint32_t epoch;
    uint32_t error;
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

// Tile shapes compiled into the library (ntt_cols.hip / ntt_rows.hip).
bool col_case_available(int log2r, int log2w);
bool row_case_available(int log2c, int log2rows);

// A plan with an explicit split, checked against the compiled tile shapes.
__host__ inline bool plan_valid(const Plan& pl) {
    return pl.log2r + pl.log2c == pl.log2n && pl.log2w <= pl.log2c && pl.log2rows <= pl.log2r &&
           col_case_available(pl.log2r, pl.log2w) && row_case_available(pl.log2c, pl.log2rows);
}

// Launch entry points (ntt_cols.hip, ntt_rows.hip, ntt_carry.hip).
hipError_t launch_build_twiddles(uint2* fwd, uint2* inv, int prime, int log2n, hipStream_t s);
hipError_t launch_forward_columns(const Plan& pl, const uint32_t* src0, uint32_t ncoef0,
                                  const uint32_t* src1, uint32_t ncoef1, uint32_t* dst0,
                                  uint32_t* dst1, int nops, const Twiddles& tw, hipStream_t s);
hipError_t launch_row_convolution(const Plan& pl, uint32_t* bufA, const uint32_t* bufB, int square,
                                  const Twiddles& tw, hipStream_t s);
// The inverse column pass also arms `sync` for the CRT/carry pass after it.
hipError_t launch_inverse_columns(const Plan& pl, uint32_t* bufA, const Twiddles& tw, CarrySync* sync,
                                  hipStream_t s);
// `status` holds one word per carry block (zero-initialised once).
hipError_t launch_crt_carry(const uint32_t* res, size_t n, uint32_t* out_words, size_t nwords,
                            const CrtConst& cc, CarrySync* sync, uint64_t* status, hipStream_t s);

}  // namespace pintt
