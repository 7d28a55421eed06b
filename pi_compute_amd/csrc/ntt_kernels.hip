// ntt_kernels.hip - HIP kernels of the three-prime NTT product for gfx950.
//
// Passes per product (N = R * C, residues as u32, one plane of N per prime):
//   1. k_col_fwd      limbs -> residues, log2(R) forward layers per column
//                     (both operands in one launch, grid.z = operand)
//   2. k_row_conv     forward layers of each row of a and b, Montgomery
//                     pointwise product, inverse layers of the row
//   3. k_col_inv_crt  inverse layers per column for all three primes, Garner
//                     CRT to a 96-bit value per position, three word planes
//   4. k_carry_fn / k_carry_scan / k_carry_apply
//                     word sums s_j = X0[j] + X1[j-1] + X2[j-2] and a
//                     parallel carry scan to the output limbs
// See ntt_kernels.h for the transform layout.
#include "ntt_kernels.h"

namespace pintt {

namespace {

__device__ __forceinline__ uint32_t prime_of(int i) {
    return i == 0 ? kP[0] : (i == 1 ? kP[1] : kP[2]);
}

// One pad word after every 16 keeps radix-16 strides off a single LDS bank.
__device__ __forceinline__ uint32_t pad(uint32_t x) { return x + (x >> 4); }

// A tile of SEQ sequences of length L in LDS.  COLS: the column pass, where
// element (seq w, index r) sits at r * SEQ + w so that neighbouring threads
// touch neighbouring words; otherwise sequences are stored one after another.
template <int LOGL_, int LOGSEQ_, bool COLS_>
struct Tile {
    static constexpr int LOGL = LOGL_;
    static constexpr int LOGSEQ = LOGSEQ_;
    static constexpr bool COLS = COLS_;
    static constexpr int L = 1 << LOGL;
    static constexpr int SEQ = 1 << LOGSEQ;
    static constexpr int NT = L * SEQ / kElemsPerThread;
    static constexpr int WORDS = L * SEQ + ((L * SEQ) >> 4);
    __device__ static __forceinline__ uint32_t addr(uint32_t seq, uint32_t idx) {
        return COLS ? pad((idx << LOGSEQ) + seq) : pad((seq << LOGL) + idx);
    }
};

// Group g of a stage covering local layers [S0, S0 + K): the 2^K elements
// base + t * DD (t < 2^K) of one sequence, DD = L / 2^(S0+K).  `bblk` is the
// block index of the group at local layer S0.
template <class TL, int S0, int K>
struct Stage {
    static constexpr int G = 1 << K;
    static constexpr int NG = kElemsPerThread / G;
    static constexpr int DD = TL::L >> (S0 + K);
    static constexpr int GPS = TL::L / G;  // groups per sequence
    __device__ static __forceinline__ void locate(uint32_t gi, uint32_t& seq, uint32_t& base,
                                                  uint32_t& bblk) {
        uint32_t grp;
        if constexpr (TL::COLS) {
            seq = gi & (TL::SEQ - 1);
            grp = gi >> TL::LOGSEQ;
        } else {
            grp = gi & (GPS - 1);
            seq = gi / GPS;
        }
        uint32_t j = grp % DD;
        bblk = grp / DD;
        base = bblk * (DD * G) + j;
    }
};

// Forward (Cooley-Tukey) layers [S0, S0+K) on one or two tiles.  Twiddle of
// local layer S0+l, sub-block `sub` of the group: global layer sb + S0 + l,
// global block ((row0 + seq) << (S0 + l)) + (bblk << l) + sub, where row0 is
// the block index of sequence 0 at layer sb (0 for columns, the row for rows).
template <class TL, int S0, int K, int NB>
__device__ __forceinline__ void ct_stage(uint32_t* t0, uint32_t* t1, const uint2* __restrict__ Z,
                                         int sb, uint32_t row0, uint32_t p) {
    using ST = Stage<TL, S0, K>;
#pragma unroll
    for (int q = 0; q < ST::NG; ++q) {
        uint32_t seq, base, bblk;
        ST::locate(q * TL::NT + threadIdx.x, seq, base, bblk);
        uint32_t blk = TL::COLS ? row0 : row0 + seq;
        uint32_t v0[ST::G], v1[ST::G];
#pragma unroll
        for (int t = 0; t < ST::G; ++t) {
            v0[t] = t0[TL::addr(seq, base + t * ST::DD)];
            if constexpr (NB == 2) v1[t] = t1[TL::addr(seq, base + t * ST::DD)];
        }
#pragma unroll
        for (int l = 0; l < K; ++l) {
            const int h = ST::G >> (l + 1);
#pragma unroll
            for (int sub = 0; sub < (1 << l); ++sub) {
                uint2 z = Z[(1u << (sb + S0 + l)) + (blk << (S0 + l)) + (bblk << l) + sub];
#pragma unroll
                for (int t = 0; t < h; ++t) {
                    ct_bfly(v0[sub * 2 * h + t], v0[sub * 2 * h + t + h], z, p);
                    if constexpr (NB == 2) ct_bfly(v1[sub * 2 * h + t], v1[sub * 2 * h + t + h], z, p);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < ST::G; ++t) {
            t0[TL::addr(seq, base + t * ST::DD)] = v0[t];
            if constexpr (NB == 2) t1[TL::addr(seq, base + t * ST::DD)] = v1[t];
        }
    }
}

// Inverse (Gentleman-Sande) layers S0+K-1 down to S0 on one tile.
template <class TL, int S0, int K>
__device__ __forceinline__ void gs_stage(uint32_t* t0, const uint2* __restrict__ Zi, int sb,
                                         uint32_t row0, uint32_t p) {
    using ST = Stage<TL, S0, K>;
#pragma unroll
    for (int q = 0; q < ST::NG; ++q) {
        uint32_t seq, base, bblk;
        ST::locate(q * TL::NT + threadIdx.x, seq, base, bblk);
        uint32_t blk = TL::COLS ? row0 : row0 + seq;
        uint32_t v[ST::G];
#pragma unroll
        for (int t = 0; t < ST::G; ++t) v[t] = t0[TL::addr(seq, base + t * ST::DD)];
#pragma unroll
        for (int l = K - 1; l >= 0; --l) {
            const int h = ST::G >> (l + 1);
#pragma unroll
            for (int sub = 0; sub < (1 << l); ++sub) {
                uint2 z = Zi[(1u << (sb + S0 + l)) + (blk << (S0 + l)) + (bblk << l) + sub];
#pragma unroll
                for (int t = 0; t < h; ++t) gs_bfly(v[sub * 2 * h + t], v[sub * 2 * h + t + h], z, p);
            }
        }
#pragma unroll
        for (int t = 0; t < ST::G; ++t) t0[TL::addr(seq, base + t * ST::DD)] = v[t];
    }
}

template <class TL, int S0, int NB>
__device__ __forceinline__ void ct_run(uint32_t* t0, uint32_t* t1, const uint2* __restrict__ Z,
                                       int sb, uint32_t row0, uint32_t p) {
    if constexpr (S0 < TL::LOGL) {
        constexpr int K = (TL::LOGL - S0) < 4 ? (TL::LOGL - S0) : 4;
        ct_stage<TL, S0, K, NB>(t0, t1, Z, sb, row0, p);
        __syncthreads();
        ct_run<TL, S0 + K, NB>(t0, t1, Z, sb, row0, p);
    }
}

template <class TL, int S0>
__device__ __forceinline__ void gs_run(uint32_t* t0, const uint2* __restrict__ Zi, int sb,
                                       uint32_t row0, uint32_t p) {
    if constexpr (S0 < TL::LOGL) {
        constexpr int K = (TL::LOGL - S0) < 4 ? (TL::LOGL - S0) : 4;
        gs_run<TL, S0 + K>(t0, Zi, sb, row0, p);
        gs_stage<TL, S0, K>(t0, Zi, sb, row0, p);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Pass 1: forward column layers, fused with limb unpacking and reduction.
template <int LOGR, int LOGW>
__global__ void __launch_bounds__((Tile<LOGR, LOGW, true>::NT))
    k_col_fwd(const uint32_t* __restrict__ src0, uint32_t ncoef0, const uint32_t* __restrict__ src1,
              uint32_t ncoef1, uint32_t* __restrict__ dst0, uint32_t* __restrict__ dst1, int log2c,
              Twiddles tw) {
    using TL = Tile<LOGR, LOGW, true>;
    __shared__ uint32_t lds[TL::WORDS];
    const int prime = blockIdx.y;
    const uint32_t p = prime_of(prime);
    const uint32_t* src = blockIdx.z ? src1 : src0;
    const uint32_t ncoef = blockIdx.z ? ncoef1 : ncoef0;
    uint32_t* dst = (blockIdx.z ? dst1 : dst0) + ((size_t)prime << (LOGR + log2c));
    const uint32_t c0 = blockIdx.x << LOGW;
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        uint32_t k = (r << log2c) + c0 + w;
        uint32_t v = k < ncoef ? src[k] : 0u;
        lds[TL::addr(w, r)] = reduce_u32(v, p);
    }
    __syncthreads();
    ct_run<TL, 0, 1>(lds, nullptr, tw.fwd[prime], 0, 0, p);
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        dst[(r << log2c) + c0 + w] = lds[TL::addr(w, r)];
    }
}

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

__device__ __forceinline__ void crt3(uint32_t r0, uint32_t r1, uint32_t r2, const CrtConst& cc,
                                     uint32_t& x0, uint32_t& x1, uint32_t& x2) {
    constexpr uint32_t p0 = kP[0], p1 = kP[1], p2 = kP[2];
    uint32_t y0 = mul_shoup(r0, cc.k0.x, cc.k0.y, p0);
    uint32_t y1 = sub_mod(mul_shoup(r1, cc.k1.x, cc.k1.y, p1), mul_shoup(y0, cc.c1.x, cc.c1.y, p1), p1);
    uint32_t y2 = sub_mod(mul_shoup(r2, cc.k2.x, cc.k2.y, p2), mul_shoup(y0, cc.c20.x, cc.c20.y, p2), p2);
    y2 = sub_mod(y2, mul_shoup(y1, cc.c21.x, cc.c21.y, p2), p2);
    // x = y0 + p0 * (y1 + p1 * y2) < p0 p1 p2 < 2^93
    uint64_t t = (uint64_t)y2 * p1 + y1;
    uint64_t lo = (uint64_t)(uint32_t)t * p0 + y0;
    uint64_t hi = (t >> 32) * (uint64_t)p0 + (lo >> 32);
    x0 = (uint32_t)lo;
    x1 = (uint32_t)hi;
    x2 = (uint32_t)(hi >> 32);
}

// Pass 3: inverse column layers for the three primes, then CRT.
template <int LOGR, int LOGW>
__global__ void __launch_bounds__((Tile<LOGR, LOGW, true>::NT))
    k_col_inv_crt(const uint32_t* __restrict__ bufA, uint32_t* __restrict__ planes, int log2c,
                  Twiddles tw, CrtConst cc) {
    using TL = Tile<LOGR, LOGW, true>;
    __shared__ uint32_t lds[TL::WORDS];
    const size_t n = (size_t)1 << (LOGR + log2c);
    const uint32_t c0 = blockIdx.x << LOGW;
    uint32_t res[kPrimes][kElemsPerThread];
#pragma unroll
    for (int prime = 0; prime < kPrimes; ++prime) {
        const uint32_t p = prime_of(prime);
        const uint32_t* src = bufA + prime * n;
#pragma unroll
        for (int e = 0; e < kElemsPerThread; ++e) {
            uint32_t idx = e * TL::NT + threadIdx.x;
            uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
            lds[TL::addr(w, r)] = src[(r << log2c) + c0 + w];
        }
        __syncthreads();
        gs_run<TL, 0>(lds, tw.inv[prime], 0, 0, p);
#pragma unroll
        for (int e = 0; e < kElemsPerThread; ++e) {
            uint32_t idx = e * TL::NT + threadIdx.x;
            uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
            res[prime][e] = lds[TL::addr(w, r)];
        }
        __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        uint32_t w = idx & (TL::SEQ - 1), r = idx >> LOGW;
        size_t k = ((size_t)r << log2c) + c0 + w;
        uint32_t x0, x1, x2;
        crt3(res[0][e], res[1][e], res[2][e], cc, x0, x1, x2);
        planes[k] = x0;
        planes[n + k] = x1;
        planes[2 * n + k] = x2;
    }
}

// ---------------------------------------------------------------------------
// Carry propagation.  Word j of the product before carries is
// s_j = X0[j] + X1[j-1] + X2[j-2] < 3 * 2^32.  Along any run of words the
// carry stays below 4, so the effect of a run on an incoming carry c in
// {0,1,2,3} is a map {0..3} -> {0..3}, packed as four 2-bit entries.  Maps
// compose associatively, which gives a parallel scan.

__device__ __forceinline__ uint32_t fn_apply(uint32_t f, uint32_t c) { return (f >> (2 * c)) & 3u; }

// later o earlier
__device__ __forceinline__ uint32_t fn_compose(uint32_t later, uint32_t earlier) {
    uint32_t r = 0;
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c) r |= fn_apply(later, fn_apply(earlier, c)) << (2 * c);
    return r;
}

constexpr uint32_t kFnIdentity = 0u | (1u << 2) | (2u << 4) | (3u << 6);

__device__ __forceinline__ uint64_t word_sum(const uint32_t* __restrict__ planes, size_t n, size_t j) {
    uint64_t s = 0;
    if (j < n) s += planes[j];
    if (j >= 1 && j - 1 < n) s += planes[n + j - 1];
    if (j >= 2 && j - 2 < n) s += planes[2 * n + j - 2];
    return s;
}

// The carry map of words [j0, j0 + 16) (words at or past nwords are zero).
__device__ __forceinline__ uint32_t thread_fn(const uint32_t* __restrict__ planes, size_t n,
                                              size_t nwords, size_t j0) {
    uint64_t c = 0;
    uint32_t w0 = 0;
    bool ones = true;
#pragma unroll
    for (int e = 0; e < kCarryWordsPerThread; ++e) {
        size_t j = j0 + e;
        uint64_t v = (j < nwords ? word_sum(planes, n, j) : 0) + c;
        uint32_t o = (uint32_t)v;
        c = v >> 32;
        if (e == 0) w0 = o;
        else ones = ones && (o == 0xFFFFFFFFu);
    }
    uint32_t f = 0;
#pragma unroll
    for (uint32_t cin = 0; cin < 4; ++cin) {
        uint32_t over = (ones && (uint64_t)w0 + cin > 0xFFFFFFFFull) ? 1u : 0u;
        f |= ((uint32_t)c + over) << (2 * cin);
    }
    return f;
}

// Inclusive scan of carry maps over the block in thread order; returns the
// thread's inclusive prefix and leaves the block total in *total.
template <int NT>
__device__ __forceinline__ uint32_t block_scan_fn(uint32_t f, uint32_t* lds_waves, uint32_t* total) {
    constexpr int W = 64;
    const int lane = threadIdx.x & (W - 1), wave = threadIdx.x / W;
#pragma unroll
    for (int off = 1; off < W; off <<= 1) {
        uint32_t other = __shfl_up(f, off, W);
        if (lane >= off) f = fn_compose(f, other);
    }
    if (lane == W - 1) lds_waves[wave] = f;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = kFnIdentity;
        for (int i = 0; i < NT / W; ++i) {
            acc = fn_compose(lds_waves[i], acc);
            lds_waves[i] = acc;  // inclusive prefix over waves
        }
    }
    __syncthreads();
    if (wave > 0) f = fn_compose(f, lds_waves[wave - 1]);
    *total = lds_waves[NT / W - 1];
    return f;
}

__global__ void __launch_bounds__(kCarryThreads)
    k_carry_fn(const uint32_t* __restrict__ planes, size_t n, size_t nwords, uint32_t* __restrict__ block_fn) {
    __shared__ uint32_t waves[kCarryThreads / 64];
    size_t j0 = (size_t)blockIdx.x * kCarryWordsPerBlock + (size_t)threadIdx.x * kCarryWordsPerThread;
    uint32_t f = thread_fn(planes, n, nwords, j0);
    uint32_t total;
    block_scan_fn<kCarryThreads>(f, waves, &total);
    if (threadIdx.x == 0) block_fn[blockIdx.x] = total;
}

// One workgroup: carry into every block from the maps of the blocks before it.
__global__ void __launch_bounds__(kScanThreads)
    k_carry_scan(const uint32_t* __restrict__ block_fn, uint32_t nblk, uint32_t* __restrict__ block_cin) {
    __shared__ uint32_t waves[kScanThreads / 64];
    const uint32_t per = (nblk + kScanThreads - 1) / kScanThreads;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t f = kFnIdentity;
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < nblk) f = fn_compose(block_fn[b0 + i], f);
    uint32_t total;
    uint32_t incl = block_scan_fn<kScanThreads>(f, waves, &total);
    // exclusive prefix = inclusive prefix of the previous thread
    uint32_t excl = __shfl_up(incl, 1, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    if (lane == 0) excl = wave > 0 ? waves[wave - 1] : kFnIdentity;
    uint32_t c = fn_apply(excl, 0);
    for (uint32_t i = 0; i < per; ++i) {
        if (b0 + i < nblk) {
            block_cin[b0 + i] = c;
            c = fn_apply(block_fn[b0 + i], c);
        }
    }
}

__global__ void __launch_bounds__(kCarryThreads)
    k_carry_apply(const uint32_t* __restrict__ planes, size_t n, size_t nwords,
                  const uint32_t* __restrict__ block_cin, uint32_t* __restrict__ out) {
    __shared__ uint32_t waves[kCarryThreads / 64];
    size_t j0 = (size_t)blockIdx.x * kCarryWordsPerBlock + (size_t)threadIdx.x * kCarryWordsPerThread;
    uint32_t f = thread_fn(planes, n, nwords, j0);
    uint32_t total;
    uint32_t incl = block_scan_fn<kCarryThreads>(f, waves, &total);
    uint32_t excl = __shfl_up(incl, 1, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    if (lane == 0) excl = wave > 0 ? waves[wave - 1] : kFnIdentity;
    uint64_t c = fn_apply(excl, block_cin[blockIdx.x]);
#pragma unroll
    for (int e = 0; e < kCarryWordsPerThread; ++e) {
        size_t j = j0 + e;
        if (j < nwords) {
            uint64_t v = word_sum(planes, n, j) + c;
            out[j] = (uint32_t)v;
            c = v >> 32;
        }
    }
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t powmod_dev(uint32_t b, uint32_t e, uint32_t p) {
    uint64_t r = 1, x = b;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

// Z[i], i = 2^s + b: w_{2^(s+1)}^bitrev_s(b) = w25^(bitrev_s(b) * 2^(24 - s)).
__global__ void k_build_twiddles(uint2* __restrict__ fwd, uint2* __restrict__ inv, uint32_t p,
                                 uint32_t w25, uint32_t count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t ef = 0;
    if (i > 1) {
        int s = 31 - __clz(i);
        uint32_t b = i - (1u << s);
        uint32_t rev = __brev(b) >> (32 - s);
        ef = rev << (24 - s);
    }
    uint32_t ei = (ef == 0) ? 0u : ((1u << 25) - ef);
    uint32_t z = (i == 0) ? 1u : powmod_dev(w25, ef, p);
    uint32_t zi = (i == 0) ? 1u : powmod_dev(w25, ei, p);
    fwd[i] = make_uint2(z, (uint32_t)(((uint64_t)z << 32) / p));
    inv[i] = make_uint2(zi, (uint32_t)(((uint64_t)zi << 32) / p));
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.  Template dispatch over the tile shapes make_plan() produces.

hipError_t launch_build_twiddles(uint2* fwd, uint2* inv, int prime, int log2n, hipStream_t s) {
    const uint32_t p = kP[prime];
    // w25: a primitive 2^25-th root of unity, g^((p-1)/2^25).
    uint64_t r = 1, x = kGen[prime];
    uint64_t e = (uint64_t)(p - 1) >> 25;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    uint32_t count = 1u << log2n;
    hipLaunchKernelGGL(k_build_twiddles, dim3((count + 255) / 256), dim3(256), 0, s, fwd, inv, p,
                       (uint32_t)r, count);
    return hipGetLastError();
}

#define PINTT_COL_CASES(X) X(5, 5) X(5, 6) X(6, 6) X(7, 5) X(8, 4) X(9, 4) X(10, 4) X(11, 3)
#define PINTT_ROW_CASES(X) \
    X(5, 5) X(6, 5) X(6, 6) X(7, 5) X(8, 4) X(9, 3) X(10, 2) X(11, 1) X(12, 0) X(13, 0) X(14, 0)

hipError_t launch_forward_columns(const Plan& pl, const uint32_t* src0, uint32_t ncoef0,
                                  const uint32_t* src1, uint32_t ncoef1, uint32_t* dst0,
                                  uint32_t* dst1, int nops, const Twiddles& tw, hipStream_t s) {
    dim3 grid(1u << (pl.log2c - pl.log2w), kPrimes, nops);
#define PINTT_CASE(LR, LW)                                                                         \
    if (pl.log2r == LR && pl.log2w == LW) {                                                        \
        hipLaunchKernelGGL((k_col_fwd<LR, LW>), grid, dim3(Tile<LR, LW, true>::NT), 0, s, src0,   \
                           ncoef0, src1, ncoef1, dst0, dst1, pl.log2c, tw);                        \
        return hipGetLastError();                                                                  \
    }
    PINTT_COL_CASES(PINTT_CASE)
#undef PINTT_CASE
    return hipErrorInvalidValue;
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

hipError_t launch_inverse_columns_crt(const Plan& pl, const uint32_t* bufA, uint32_t* planes,
                                      const Twiddles& tw, const CrtConst& cc, hipStream_t s) {
    dim3 grid(1u << (pl.log2c - pl.log2w));
#define PINTT_CASE(LR, LW)                                                                          \
    if (pl.log2r == LR && pl.log2w == LW) {                                                         \
        hipLaunchKernelGGL((k_col_inv_crt<LR, LW>), grid, dim3(Tile<LR, LW, true>::NT), 0, s, bufA, \
                           planes, pl.log2c, tw, cc);                                               \
        return hipGetLastError();                                                                   \
    }
    PINTT_COL_CASES(PINTT_CASE)
#undef PINTT_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_carry(const uint32_t* planes, size_t n, uint32_t* out_words, size_t nwords,
                        uint32_t* block_fn, uint32_t* block_cin, hipStream_t s) {
    uint32_t nblk = (uint32_t)((nwords + kCarryWordsPerBlock - 1) / kCarryWordsPerBlock);
    hipLaunchKernelGGL(k_carry_fn, dim3(nblk), dim3(kCarryThreads), 0, s, planes, n, nwords, block_fn);
    hipLaunchKernelGGL(k_carry_scan, dim3(1), dim3(kScanThreads), 0, s, block_fn, nblk, block_cin);
    hipLaunchKernelGGL(k_carry_apply, dim3(nblk), dim3(kCarryThreads), 0, s, planes, n, nwords,
                       block_cin, out_words);
    return hipGetLastError();
}

}  // namespace pintt
