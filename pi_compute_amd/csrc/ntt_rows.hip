// ntt_rows.hip - row pass of the four-step NTT (gfx950): forward layers on the
// rows of a and b, Montgomery pointwise product, inverse layers, in place.
#include "ntt_device.h"

namespace pintt {
namespace {

// Pass 2: forward row layers of a (and b), pointwise product, inverse row layers.
template <int LOGC, int LOGROWS, bool SQUARE>
__global__ void __launch_bounds__((Tile<LOGC, LOGROWS, false>::NT))
    k_row_conv(uint32_t* __restrict__ bufA, const uint32_t* __restrict__ bufB, int log2r,
               Twiddles tw) {
    using TL = Tile<LOGC, LOGROWS, false>;
    __shared__ uint32_t ldsA[TL::WORDS];
    __shared__ uint32_t ldsB[SQUARE ? 1 : TL::WORDS];
    const int prime = blockIdx.y;
    const uint32_t p = prime_of(prime);
    const uint32_t pinv = neg_inv_mod_2_32(p);
    const size_t base = ((size_t)prime << (log2r + LOGC)) + ((size_t)blockIdx.x << (LOGROWS + LOGC));
    const uint32_t row0 = blockIdx.x << LOGROWS;
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        ldsA[pad(idx)] = bufA[base + idx];
        if constexpr (!SQUARE) ldsB[pad(idx)] = bufB[base + idx];
    }
    __syncthreads();
    ct_run<TL, 0, SQUARE ? 1 : 2>(ldsA, ldsB, tw.fwd[prime], log2r, row0, p);
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t a = ldsA[pad(e * TL::NT + threadIdx.x)];
        uint32_t b = SQUARE ? a : ldsB[pad(e * TL::NT + threadIdx.x)];
        ldsA[pad(e * TL::NT + threadIdx.x)] = mul_mont(a, b, p, pinv);
    }
    __syncthreads();
    gs_run<TL, 0>(ldsA, tw.inv[prime], log2r, row0, p);
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        bufA[base + idx] = ldsA[pad(idx)];
    }
}


}  // namespace

#define PINTT_ROW_CASES(X) \
    X(5, 5) X(5, 6) X(5, 7) X(5, 8) X(6, 5) X(6, 6) X(6, 7) X(7, 4) \
    X(7, 5) X(7, 6) X(8, 3) X(8, 4) X(8, 5) X(9, 2) X(9, 3) X(9, 4) \
    X(10, 1) X(10, 2) X(10, 3) X(11, 0) X(11, 1) X(11, 2) X(12, 0) X(12, 1) \
    X(13, 0) X(14, 0)

bool row_case_available(int log2c, int log2rows) {
#define PINTT_CASE(LC, LROWS) \
    if (log2c == LC && log2rows == LROWS) return true;
    PINTT_ROW_CASES(PINTT_CASE)
#undef PINTT_CASE
    return false;
}

hipError_t launch_row_convolution(const Plan& pl, uint32_t* bufA, const uint32_t* bufB, int square,
                                  const Twiddles& tw, hipStream_t s) {
    dim3 grid(1u << (pl.log2r - pl.log2rows), kPrimes);
#define PINTT_CASE(LC, LROWS)                                                                        \
    if (pl.log2c == LC && pl.log2rows == LROWS) {                                                    \
        if (square)                                                                                  \
            hipLaunchKernelGGL((k_row_conv<LC, LROWS, true>), grid, dim3(Tile<LC, LROWS, false>::NT), \
                               0, s, bufA, bufB, pl.log2r, tw);                                      \
        else                                                                                         \
            hipLaunchKernelGGL((k_row_conv<LC, LROWS, false>), grid,                                 \
                               dim3(Tile<LC, LROWS, false>::NT), 0, s, bufA, bufB, pl.log2r, tw);    \
        return hipGetLastError();                                                                    \
    }
    PINTT_ROW_CASES(PINTT_CASE)
#undef PINTT_CASE
    return hipErrorInvalidValue;
}

}  // namespace pintt
