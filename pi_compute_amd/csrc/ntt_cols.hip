// ntt_cols.hip - column passes of the four-step NTT (gfx950).
//   k_col_fwd  limbs -> residues (unpack + reduce), log2(R) forward layers per
//              column, every operand chunk of a bucket in one launch
//   k_col_inv  log2(R) inverse layers per column, in place
#include "ntt_regs.h"

#include <cstdlib>

namespace pintt {
namespace {

__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t bx, uint32_t nb) {
    return (nb & 7u) ? bx : (bx & 7u) * (nb >> 3) + (bx >> 3);
}

// Pass 1: forward column layers, fused with limb unpacking and reduction.
// grid = (column tiles, prime, job).  SET: the prime set of the transform length.
template <int LOGR, int LOGW, int SET>
__global__ void __launch_bounds__((Tile<LOGR, LOGW, true>::NT))
    k_col_fwd(const JobList<FwdJob> jobs, int log2c, Twiddles tw) {
    using TL = Tile<LOGR, LOGW, true>;
    __shared__ uint32_t lds[TL::WORDS];
    __shared__ uint2 zt[TL::L];  // the R - 1 twiddles of the column layers (entry 0 unused)
    // Tiles of 2^5 columns or more read whole 128-byte lines of their own: the
    // three primes of a tile run on one XCD, one after another, and share its
    // L2 copy of the limbs.  Narrower tiles share lines with their neighbours:
    // each XCD takes a contiguous range of tiles.
    constexpr bool kGroupPrimes = LOGW >= 5;
    const ColBlock cb = kGroupPrimes ? prime_grouped_block(blockIdx.x + gridDim.x * (blockIdx.y + kPrimes * blockIdx.z),
                                                           gridDim.x, gridDim.z)
                                     : ColBlock{xcd_swizzle(blockIdx.x, gridDim.x), blockIdx.y, blockIdx.z};
    const FwdJob job = PINTT_JOB(FwdJob, jobs, cb.job);
    const int prime = (int)cb.prime;
    const uint32_t p = prime_of<SET>(prime);
    global_ptr<uint32_t> dst = as_global(job.dst) + ((size_t)prime << (LOGR + log2c));
    global_ptr<const uint32_t> src = as_global(job.src);
    const TwTable z = PINTT_FWD_TABLE(tw, prime);
    for (int i = threadIdx.x; i < TL::L; i += TL::NT) zt[i] = tw_at(z, i);
    const uint32_t c0 = cb.tile << LOGW;
    const uint32_t w0 = TL::word(threadIdx.x);  // element e: tile position e NT + threadIdx.x
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        uint32_t k = (r << log2c) + c0 + w;
        uint32_t v = k < job.ncoef ? src[k] : 0u;
        lds[TL::word_at(w0, e * TL::NT)] = Arith<SET>::load_coeff(v, p);
    }
    __syncthreads();
    uint32_t* const tiles[1] = {lds};
    ct_run<SET, TL, 0, 1>(tiles, zt, 0, 0, p);
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        dst[(r << log2c) + c0 + w] = lds[TL::word_at(w0, e * TL::NT)];
    }
}

// Pass 3: inverse column layers, one prime of one chunk product per workgroup,
// in place.  grid = (column tiles, prime, job).
template <int LOGR, int LOGW, int SET>
__global__ void __launch_bounds__((Tile<LOGR, LOGW, true>::NT))
    k_col_inv(const JobList<InvJob> jobs, int log2c, Twiddles tw) {
    using TL = Tile<LOGR, LOGW, true>;
    __shared__ uint32_t lds[TL::WORDS];
    __shared__ uint2 zt[TL::L];  // inverse twiddles of the column layers
    const InvJob job = PINTT_JOB(InvJob, jobs, blockIdx.z);
    const int prime = blockIdx.y;
    const uint32_t p = prime_of<SET>(prime);
    global_ptr<uint32_t> buf = as_global(job.buf) + ((size_t)prime << (LOGR + log2c));
    const uint32_t c0 = xcd_swizzle(blockIdx.x, gridDim.x) << LOGW;
    const TwTable z = PINTT_INV_TABLE(tw, prime);
    for (int i = threadIdx.x; i < TL::L; i += TL::NT) zt[i] = tw_at(z, i);
    const uint32_t w0 = TL::word(threadIdx.x);  // element e: tile position e NT + threadIdx.x
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        lds[TL::word_at(w0, e * TL::NT)] = buf[(r << log2c) + c0 + w];
    }
    __syncthreads();
    uint32_t* const tiles[1] = {lds};
    gs_run<SET, TL, 0, 1>(tiles, zt, 0, 0, p);
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        buf[(r << log2c) + c0 + w] = lds[TL::word_at(w0, e * TL::NT)];
    }
}

// Column passes with register-resident stages (ntt_regs.h) for tiles of 2^LR
// rows (LR = 7 .. 11) by 8 columns, 2^(LR - 1) threads.  Row r of column w:
// stage A runs layers 0..5 on r's top six bits for the groups g = (r_low, w)
// (r_low = r's low LR - 6 bits; g = (wave << 4) | (lane & 15), w = g & 7);
// stage B runs the last LR - 6 layers on the groups (r >> (LR - 6), w): inside
// each lane for 1 to 4 layers (a lane holds 16 consecutive rows of one column,
// 16 >> (LR - 6) groups), as a five-layer lane-swap stage (lanes l and l + 32,
// row bit 0 on lane bit 5) for LR = 11.  Global memory is read and
// written in the stage A entry layout (forward) or the stage B layout
// (inverse) as whole 32-byte column rows; LDS (word k = (r << 3) | w, padded
// by S words per 2^PSH) carries the one exchange between the stages, at 32
// distinct banks per half-wave both ways (padding and the lane bits of the
// stage B groups chosen per shape; tests/test_reg_layouts.py checks them).
template <int LR>
struct RegCols {
    static_assert(LR >= 7 && LR <= 11, "register column tiles: 2^7 .. 2^11 rows");
    static constexpr int NT = 1 << (LR - 1);
    static constexpr int KB = LR - 6;  // stage B layers
    static constexpr int SB = LR == 7 ? 2 : 1;  // stage A lane swap: register bit SB <-> lane bit 4
    static constexpr uint32_t S = LR <= 8 ? 4u : (LR == 9 ? 2u : 1u);
    static constexpr int PSH = LR == 11 ? 7 : 6;
    static constexpr int WORDS = (8 << LR) + (int)S * ((8 << LR) >> PSH);
    __device__ static __forceinline__ uint32_t lds_at(uint32_t r, uint32_t w) {
        const uint32_t k = (r << 3) | w;
        return k + S * (k >> PSH);
    }
    // Stage A: row of entry-layout register i / post-layout register i.
    __device__ static __forceinline__ uint32_t row_entry(int i, uint32_t lane, uint32_t rlow) {
        return (entry_t(i, lane) << (LR - 6)) | rlow;
    }
    __device__ static __forceinline__ uint32_t row_post(int i, uint32_t lane, uint32_t rlow) {
        return (post_t<SB>(i, lane) << (LR - 6)) | rlow;
    }
    // Stage B: the lane's rows are hi = r >> 4 (LR <= 10; 16 >> (LR - 6)
    // groups per lane, told apart by the register bits above LR - 6) or
    // r >> 5 (LR = 11).
    __device__ static __forceinline__ uint32_t hi_bits(uint32_t lane, uint32_t wave) {
        const uint32_t l3 = (lane >> 3) & 1u, l4 = (lane >> 4) & 1u, l5 = (lane >> 5) & 1u;
        if constexpr (LR == 11)  // r5, r6 = wave0, wave1; r7 = L3; r8 = L4; r9, r10 = wave2, wave3 (r0 = L5)
            return (wave & 3u) | (l3 << 2) | (l4 << 3) | ((wave >> 2) << 4);
        else if constexpr (LR == 10)  // r4 = L5, r5 = wave0, r6 = L3, r7 = L4, r8, r9 = wave1, wave2
            return l5 | ((wave & 1u) << 1) | (l3 << 2) | (l4 << 3) | ((wave >> 1) << 4);
        else if constexpr (LR == 9)  // r4 = L5, r5 = L3, r6 = L4, r7, r8 = wave0, wave1 (r3 = register bit 3)
            return l5 | (l3 << 1) | (l4 << 2) | (wave << 3);
        else  // LR = 7, 8: r4 = L3, r5 = L4, r6 = L5, r7 = wave0
            return l3 | (l4 << 1) | (l5 << 2) | (wave << 3);
    }
    // Stage B row of register i, in the entry layout and (LR = 11) the post one.
    __device__ static __forceinline__ uint32_t row_b(int i, uint32_t lane, uint32_t hi) {
        if constexpr (LR == 11) return (hi << 5) | ((uint32_t)i << 1) | ((lane >> 5) & 1u);
        else return (hi << 4) | (uint32_t)i;
    }
    __device__ static __forceinline__ uint32_t row_b_post(int i, uint32_t lane, uint32_t hi) {
        if constexpr (LR == 11)
            return (hi << 5) | (((lane >> 5) & 1u) << 4) | (((uint32_t)i & 7u) << 1) | ((uint32_t)i >> 3);
        else return (hi << 4) | (uint32_t)i;
    }
};

// Stage B twiddles: at layer 6 + l, block (outer << l) | blk with outer the
// group's block at layer 6, r >> KB: hi (LR = 10, 11) or (hi << (4 - KB)) | g
// (LR = 7 .. 9, g = the register bits above KB).
struct TwColB {
    ConstTable z;
    uint32_t hi;
    int KB;
    __device__ __forceinline__ uint2 operator()(int g, int l, uint32_t blk) const {
        const uint32_t outer = KB >= 4 ? hi : ((hi << (4 - KB)) | (uint32_t)g);
        return tw_at(z, (1u << (6 + l)) + (outer << l) + blk);
    }
    __device__ __forceinline__ uint2 operator()(int l, uint32_t blk) const { return (*this)(0, l, blk); }
};
struct TwColA {
    ConstTable z;
    __device__ __forceinline__ uint2 operator()(int l, uint32_t blk) const { return tw_at(z, (1u << l) + blk); }
};

// Stage B of a register column kernel on the lane's 16 residues.  FRESH
// (inverse): the values are below 2p, as scaled rows store them.
template <int LR, int SET, bool INV, bool FRESH = false>
__device__ __forceinline__ void col_stage_b(uint32_t (&v)[1][16], const TwColB& tw, uint32_t p) {
    if constexpr (LR == 11) {
        if constexpr (INV) gs5<SET, 1, FRESH>(v, tw, p);
        else ct5<SET, 1>(v, tw, p);
    } else {
        if constexpr (INV) gs_lane<SET, LR - 6, FRESH>(v[0], tw, p);
        else ct_lane<SET, LR - 6>(v[0], tw, p);
    }
}

template <int LR, int SET>
__global__ void __launch_bounds__(RegCols<LR>::NT) k_col_fwd_reg(const JobList<FwdJob> jobs, int log2c, Twiddles tw) {
    using RC = RegCols<LR>;
    __shared__ uint32_t lds[RC::WORDS];
    const FwdJob job = PINTT_JOB(FwdJob, jobs, blockIdx.z);
    const int prime = blockIdx.y;
    const uint32_t p = prime_of<SET>(prime);
    global_ptr<uint32_t> dst = as_global(job.dst) + ((size_t)prime << (LR + log2c));
    global_ptr<const uint32_t> src = as_global(job.src);
    const TwTable z = PINTT_FWD_TABLE(tw, prime);
    const uint32_t c0 = xcd_swizzle(blockIdx.x, gridDim.x) << 3;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint32_t g = (wave << 4) | (lane & 15u), w = g & 7u, rlow = g >> 3;
    uint32_t v[1][16];
    // Rows from R/2 on hold no coefficient when the operand fills at most half
    // the transform (every operand of a product of two halves, every square):
    // then layer 0 is a copy and the top half is not read.
    if (job.ncoef <= (1u << (LR + log2c - 1))) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t k = (RC::row_entry(i, lane, rlow) << log2c) + c0 + w;
            v[0][i] = Arith<SET>::load_coeff(k < job.ncoef ? src[k] : 0u, p);
        }
        ct6<SET, 1, true, RC::SB>(v, TwColA{as_const(z)}, p);
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t k = (RC::row_entry(i, lane, rlow) << log2c) + c0 + w;
            v[0][i] = Arith<SET>::load_coeff(k < job.ncoef ? src[k] : 0u, p);
        }
        ct6<SET, 1, false, RC::SB>(v, TwColA{as_const(z)}, p);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) lds[RC::lds_at(RC::row_post(i, lane, rlow), w)] = v[0][i];
    __syncthreads();
    const uint32_t hi = RC::hi_bits(lane, wave), wb = lane & 7u;
#pragma unroll
    for (int i = 0; i < 16; ++i) v[0][i] = lds[RC::lds_at(RC::row_b(i, lane, hi), wb)];
    col_stage_b<LR, SET, false>(v, TwColB{as_const(z), hi, RC::KB}, p);
#pragma unroll
    for (int i = 0; i < 16; ++i) dst[(RC::row_b_post(i, lane, hi) << log2c) + c0 + wb] = v[0][i];
}

// FRESH: every product row arrives below 2p (the row pass scaled its rows).
template <int LR, int SET, bool FRESH>
__global__ void __launch_bounds__(RegCols<LR>::NT) k_col_inv_reg(const JobList<InvJob> jobs, int log2c, Twiddles tw) {
    using RC = RegCols<LR>;
    __shared__ uint32_t lds[RC::WORDS];
    const InvJob job = PINTT_JOB(InvJob, jobs, blockIdx.z);
    const int prime = blockIdx.y;
    const uint32_t p = prime_of<SET>(prime);
    global_ptr<uint32_t> buf = as_global(job.buf) + ((size_t)prime << (LR + log2c));
    const TwTable z = PINTT_INV_TABLE(tw, prime);
    const uint32_t c0 = xcd_swizzle(blockIdx.x, gridDim.x) << 3;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint32_t hi = RC::hi_bits(lane, wave), wb = lane & 7u;
    uint32_t v[1][16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[0][i] = buf[(RC::row_b_post(i, lane, hi) << log2c) + c0 + wb];
    col_stage_b<LR, SET, true, FRESH>(v, TwColB{as_const(z), hi, RC::KB}, p);
#pragma unroll
    for (int i = 0; i < 16; ++i) lds[RC::lds_at(RC::row_b(i, lane, hi), wb)] = v[0][i];
    __syncthreads();
    const uint32_t g = (wave << 4) | (lane & 15u), w = g & 7u, rlow = g >> 3;
#pragma unroll
    for (int i = 0; i < 16; ++i) v[0][i] = lds[RC::lds_at(RC::row_post(i, lane, rlow), w)];
    gs6<SET, 1, RC::SB>(v, TwColA{as_const(z)}, p);
#pragma unroll
    for (int i = 0; i < 16; ++i) buf[(RC::row_entry(i, lane, rlow) << log2c) + c0 + w] = v[0][i];
}

// Column tiles of 2^9 .. 2^11 rows by 8 columns go to the register kernels
// unless PINTT_REG_COLS=0 (the forward pass also for the narrow set's 2^7-
// and 2^8-row tiles).
bool reg_cols_enabled() {
    static const bool on = [] {
        const char* v = std::getenv("PINTT_REG_COLS");
        return !(v && v[0] == '0');
    }();
    return on;
}

}  // namespace

// Tile shapes per prime set.  The wide set only serves 2^23 .. 2^25 points,
// where rows of at most 2^14 leave 2^9 .. 2^11 rows.
#define PINTT_COL_CASES_NARROW(X) X(5, 5) X(5, 6) X(6, 4) X(6, 5) X(6, 6) X(7, 3) X(7, 4) X(7, 5) X(8, 3) X(8, 4) X(8, 5) X(8, 6) X(9, 3) X(9, 4) X(9, 5) X(10, 3) X(10, 4) X(11, 3)
#define PINTT_COL_CASES_WIDE(X) X(9, 3) X(9, 4) X(9, 5) X(10, 3) X(10, 4) X(11, 3)

bool col_case_available(int set, int log2r, int log2w) {
#define PINTT_CASE(LR, LW) \
    if (log2r == LR && log2w == LW) return true;
    if (set == kSetNarrow) {
        PINTT_COL_CASES_NARROW(PINTT_CASE)
    } else {
        PINTT_COL_CASES_WIDE(PINTT_CASE)
    }
#undef PINTT_CASE
    return false;
}

hipError_t launch_forward_columns(const Plan& pl, const JobList<FwdJob>& jobs, uint32_t njobs, const Twiddles& tw,
                                  hipStream_t s) {
    dim3 grid(1u << (pl.log2c - pl.log2w), kPrimes, njobs);
    if (pl.log2w == 3 && pl.log2r >= 7 && pl.log2r <= 11 && reg_cols_enabled()) {
#define PINTT_REG_COL(LR, SET)                                                                                   \
    if (pl.log2r == LR) {                                                                                         \
        hipLaunchKernelGGL((k_col_fwd_reg<LR, SET>), grid, dim3(RegCols<LR>::NT), 0, s, jobs, pl.log2c, tw); \
        return hipGetLastError();                                                                                \
    }
        if (set_for(pl.log2n) == kSetNarrow) {
            PINTT_REG_COL(7, kSetNarrow)
            PINTT_REG_COL(8, kSetNarrow)
            PINTT_REG_COL(9, kSetNarrow)
            PINTT_REG_COL(10, kSetNarrow)
            PINTT_REG_COL(11, kSetNarrow)
        } else {
            PINTT_REG_COL(9, kSetWide)
            PINTT_REG_COL(10, kSetWide)
            PINTT_REG_COL(11, kSetWide)
        }
#undef PINTT_REG_COL
    }
#define PINTT_CASE_SET(LR, LW, SET)                                                                  \
    if (pl.log2r == LR && pl.log2w == LW) {                                                          \
        hipLaunchKernelGGL((k_col_fwd<LR, LW, SET>), grid, dim3(Tile<LR, LW, true>::NT), 0, s, jobs, \
                           pl.log2c, tw);                                                            \
        return hipGetLastError();                                                                    \
    }
#define PINTT_CASE_N(LR, LW) PINTT_CASE_SET(LR, LW, kSetNarrow)
#define PINTT_CASE_W(LR, LW) PINTT_CASE_SET(LR, LW, kSetWide)
    if (set_for(pl.log2n) == kSetNarrow) {
        PINTT_COL_CASES_NARROW(PINTT_CASE_N)
    } else {
        PINTT_COL_CASES_WIDE(PINTT_CASE_W)
    }
#undef PINTT_CASE_W
#undef PINTT_CASE_N
#undef PINTT_CASE_SET
    return hipErrorInvalidValue;
}

hipError_t launch_inverse_columns(const Plan& pl, const JobList<InvJob>& jobs, uint32_t njobs, const Twiddles& tw,
                                  hipStream_t s) {
    dim3 grid(1u << (pl.log2c - pl.log2w), kPrimes, njobs);
    // 2^7- and 2^8-row tiles keep the LDS kernel for the inverse: reading and
    // writing 32-byte column rows from one-wave workgroups, the register
    // kernel moved a 2^15-point bucket at 2.6 TB/s against 4.1 TB/s for the
    // LDS kernel with 16-column tiles (the forward pass, which writes only,
    // is faster on the register kernel).
    if (pl.log2w == 3 && pl.log2r >= 9 && pl.log2r <= 11 && reg_cols_enabled()) {
        const bool fresh = rows_scaled(pl);
#define PINTT_REG_COL(LR, SET)                                                                                   \
    if (pl.log2r == LR) {                                                                                         \
        if (fresh)                                                                                                \
            hipLaunchKernelGGL((k_col_inv_reg<LR, SET, true>), grid, dim3(RegCols<LR>::NT), 0, s, jobs,       \
                               pl.log2c, tw);                                                                    \
        else                                                                                                      \
            hipLaunchKernelGGL((k_col_inv_reg<LR, SET, false>), grid, dim3(RegCols<LR>::NT), 0, s, jobs,      \
                               pl.log2c, tw);                                                                    \
        return hipGetLastError();                                                                                \
    }
        if (set_for(pl.log2n) == kSetNarrow) {
            PINTT_REG_COL(9, kSetNarrow)
            PINTT_REG_COL(10, kSetNarrow)
            PINTT_REG_COL(11, kSetNarrow)
        } else {
            PINTT_REG_COL(9, kSetWide)
            PINTT_REG_COL(10, kSetWide)
            PINTT_REG_COL(11, kSetWide)
        }
#undef PINTT_REG_COL
    }
#define PINTT_CASE_SET(LR, LW, SET)                                                                  \
    if (pl.log2r == LR && pl.log2w == LW) {                                                          \
        hipLaunchKernelGGL((k_col_inv<LR, LW, SET>), grid, dim3(Tile<LR, LW, true>::NT), 0, s, jobs, \
                           pl.log2c, tw);                                                            \
        return hipGetLastError();                                                                    \
    }
#define PINTT_CASE_N(LR, LW) PINTT_CASE_SET(LR, LW, kSetNarrow)
#define PINTT_CASE_W(LR, LW) PINTT_CASE_SET(LR, LW, kSetWide)
    if (set_for(pl.log2n) == kSetNarrow) {
        PINTT_COL_CASES_NARROW(PINTT_CASE_N)
    } else {
        PINTT_COL_CASES_WIDE(PINTT_CASE_W)
    }
#undef PINTT_CASE_W
#undef PINTT_CASE_N
#undef PINTT_CASE_SET
    return hipErrorInvalidValue;
}

}  // namespace pintt
