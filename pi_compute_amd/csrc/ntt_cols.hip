// ntt_cols.hip - column passes of the four-step NTT (gfx950).
//   k_col_fwd  limbs -> residues (unpack + reduce), log2(R) forward layers per
//              column, every operand chunk of a bucket in one launch
//   k_col_inv  log2(R) inverse layers per column, in place
#include "ntt_device.h"

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
    const FwdJob job = PINTT_JOB(FwdJob, jobs, blockIdx.z);
    const int prime = blockIdx.y;
    const uint32_t p = prime_of<SET>(prime);
    global_ptr<uint32_t> dst = as_global(job.dst) + ((size_t)prime << (LOGR + log2c));
    global_ptr<const uint32_t> src = as_global(job.src);
    const TwTable z = PINTT_FWD_TABLE(tw, prime);
    for (int i = threadIdx.x; i < TL::L; i += TL::NT) zt[i] = tw_at(z, i);
    // Neighbouring column tiles share 128-byte lines; give each XCD (workgroups
    // are dealt to the 8 XCDs round-robin) a contiguous range of tiles.
    const uint32_t c0 = xcd_swizzle(blockIdx.x, gridDim.x) << LOGW;
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        uint32_t k = (r << log2c) + c0 + w;
        uint32_t v = k < job.ncoef ? src[k] : 0u;
        lds[TL::addr(w, r)] = Arith<SET>::load_coeff(v, p);
    }
    __syncthreads();
    uint32_t* const tiles[1] = {lds};
    ct_run<SET, TL, 0, 1>(tiles, zt, 0, 0, p);
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        dst[(r << log2c) + c0 + w] = lds[TL::addr(w, r)];
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
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        lds[TL::addr(w, r)] = buf[(r << log2c) + c0 + w];
    }
    __syncthreads();
    uint32_t* const tiles[1] = {lds};
    gs_run<SET, TL, 0, 1>(tiles, zt, 0, 0, p);
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        buf[(r << log2c) + c0 + w] = lds[TL::addr(w, r)];
    }
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
