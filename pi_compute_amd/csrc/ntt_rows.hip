// ntt_rows.hip - row pass of the four-step NTT (gfx950): forward layers on the
// rows of the operands, Montgomery pointwise products, inverse layers, in place.
#include "ntt_regs.h"

#include <cstdlib>

namespace pintt {
namespace {

// Pass 2, per row tile: forward row layers of the chunk(s) (and of b unless
// they are already done), Montgomery pointwise products, inverse row layers,
// written back in place.  grid = (row tiles, prime, job).  SET: the prime set
// of the transform length.
// Twiddles: a row's own twiddles are the table entries of its blocks (one
// (w, w') pair per element and direction: N pairs per direction in all,
// which outgrow the caches from 2^22 points on).  SCALED kernels (one row per
// workgroup) instead scale the row by a_r^c as it is loaded and by a_r^-c as
// it is stored and run the layers of block 0, whose twiddles every workgroup
// shares (row_factor); a scaling costs three multiplications per element and
// direction (mul_fixed) plus one Shoup product per element position for the
// factor.  Measured on MI355X against per-row twiddles: squares 5% faster at
// 2^22 points, paired rows 2% faster at 2^21 and 7% faster at 2^20 (the k = 25
// product), one-chunk rows 6% slower at 2^20.  launch_row_convolution picks
// SCALED from 2^21 points on for squares, one-chunk products and later chunks,
// and from 2^20 on for pairs.
// Transformed rows of b, stored for later chunks or read back, are never
// scaled.
// Three waves per SIMD (at most 168 VGPRs): left alone, the compiler spends
// ~250 registers on the paired transforms and halves occupancy.  The pair
// kind (three tiles in flight) gets two waves and 256 registers instead.
template <int LOGC, int LOGROWS, int KIND, int SET, bool SCALED>
__global__ void __launch_bounds__((Tile<LOGC, LOGROWS, false>::NT)) __attribute__((amdgpu_waves_per_eu(KIND == kRowPair ? 2 : 3)))
    k_row_conv(const JobList<RowJob> jobs, int log2r, Twiddles tw) {
    using TL = Tile<LOGC, LOGROWS, false>;
    constexpr bool kPair = KIND == kRowPair;
    constexpr bool kScaled = SCALED;
    static_assert(!SCALED || LOGROWS == 0, "scaled rows: one row per workgroup");
    __shared__ uint32_t ldsA[TL::WORDS];
    __shared__ uint32_t ldsB[KIND == kRowSquare ? 1 : TL::WORDS];
    __shared__ uint32_t ldsC[kPair ? TL::WORDS : 1];
    const RowJob job = PINTT_JOB(RowJob, jobs, blockIdx.z);
    const int prime = blockIdx.y;
    const uint32_t p = prime_of<SET>(prime);
    const uint32_t pinv = neg_inv_mod_2_32(p);
    const size_t base = ((size_t)prime << (log2r + LOGC)) + ((size_t)blockIdx.x << (LOGROWS + LOGC));
    const uint32_t row0 = blockIdx.x << LOGROWS;
    global_ptr<uint32_t> a = as_global(job.a) + base;
    global_ptr<uint32_t> b = as_global(job.b) + base;
    global_ptr<uint32_t> a2 = as_global(job.a2) + base;
    // Scaled rows: element e of a thread sits at c = e NT + tid, and
    // a_r^c = a_r^tid * a_r^(e NT).  Exponents bitrev_log2r(r) c 2^(K - n)
    // stay below 2^K.  The second factor, beta_e, is the same for the whole
    // workgroup: 32 threads put the 16 forward and 16 inverse ones in LDS as
    // plain residues with their Shoup quotients.  The first, h = a_r^(+-tid)
    // in Montgomery form, is the thread's own; h beta_e (one Shoup product)
    // is then the Montgomery form f of a_r^(+-c), and an element is scaled
    // by mul_fixed with f and f p^-1.
    const RowTable rt = PINTT_ROW_TABLE(tw, prime);
    constexpr uint32_t kMask = (1u << kSetMaxLog2N[SET]) - 1;
    const uint32_t rbits = __brev(row0) >> (32 - log2r);
    const int eshift = kSetMaxLog2N[SET] - (log2r + LOGC);
    const uint32_t pinv_pos = 0u - pinv;  // p^-1 mod 2^32
    __shared__ uint2 beta[kScaled ? 2 * kElemsPerThread : 1];
    uint32_t h_fwd = 0, h_inv = 0;
    if constexpr (kScaled) {
        const uint32_t et = (rbits * threadIdx.x) << eshift;
        h_fwd = row_factor(rt, et, p);
        h_inv = row_factor(rt, (0u - et) & kMask, p);
        if (threadIdx.x < 2 * kElemsPerThread) {
            const uint32_t e = threadIdx.x % kElemsPerThread;
            const uint32_t eu = (rbits * (e * TL::NT)) << eshift;
            const uint32_t bm = row_factor(rt, threadIdx.x < kElemsPerThread ? eu : (0u - eu) & kMask, p);
            const uint32_t bv = mul_mont(bm, 1u, p, pinv);  // out of Montgomery form
            beta[threadIdx.x] = make_uint2(bv, (uint32_t)(((uint64_t)bv << 32) / p));
        }
        __syncthreads();
    }
    // Element e of a thread has tile position e NT + threadIdx.x.
    const uint32_t w0 = TL::word(threadIdx.x);
    auto at = [&](int e) { return TL::word_at(w0, e * TL::NT); };
    auto factor = [&](int e, bool inverse) {
        const uint2 bq = beta[(inverse ? kElemsPerThread : 0) + e];
        const uint32_t f = mul_shoup(inverse ? h_inv : h_fwd, bq.x, bq.y, p);
        return make_uint2(f, f * pinv_pos);
    };
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        if constexpr (kScaled) {
            const uint2 f = factor(e, false);
            ldsA[at(e)] = Arith<SET>::scale(a[idx], f.x, f.y, p);
            if constexpr (KIND == kRowTransformB || kPair) ldsB[at(e)] = Arith<SET>::scale(b[idx], f.x, f.y, p);
            if constexpr (KIND == kRowReuseB) ldsB[at(e)] = b[idx];
            if constexpr (kPair) ldsC[at(e)] = Arith<SET>::scale(a2[idx], f.x, f.y, p);
        } else {
            ldsA[at(e)] = a[idx];
            if constexpr (KIND != kRowSquare) ldsB[at(e)] = b[idx];
            if constexpr (kPair) ldsC[at(e)] = a2[idx];
        }
    }
    __syncthreads();
    // Scaled rows run block 0's layers: sb = row0 = 0.
    const int sb = kScaled ? 0 : log2r;
    const uint32_t rb = kScaled ? 0u : row0;
    const auto zf = PINTT_FWD_TABLE(tw, prime);
    if constexpr (kPair) {
        uint32_t* const t[3] = {ldsA, ldsB, ldsC};
        ct_run<SET, TL, 0, 3, kScaled ? 2 : 8>(t, zf, sb, rb, p);
    } else if constexpr (KIND == kRowTransformB) {
        uint32_t* const t[2] = {ldsA, ldsB};
        ct_run<SET, TL, 0, 2, kScaled ? 2 : 8>(t, zf, sb, rb, p);
    } else {
        uint32_t* const t[1] = {ldsA};
        ct_run<SET, TL, 0, 1, kScaled ? 2 : 8>(t, zf, sb, rb, p);
    }
    if constexpr (KIND == kRowTransformB || kPair) {
        if (job.write_b) {
#pragma unroll
            for (int e = 0; e < kElemsPerThread; ++e) {
                uint32_t idx = e * TL::NT + threadIdx.x;
                b[idx] = ldsB[at(e)];
            }
        }
    }
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = at(e);
        uint32_t x = ldsA[idx];
        if constexpr (KIND == kRowSquare) {
            ldsA[idx] = Arith<SET>::pointwise(x, x, p, pinv);
        } else {
            const uint32_t ys = Arith<SET>::shared_operand(ldsB[idx], p);
            ldsA[idx] = Arith<SET>::pointwise_shared(x, ys, p, pinv);
            if constexpr (kPair) ldsC[idx] = Arith<SET>::pointwise_shared(ldsC[idx], ys, p, pinv);
        }
    }
    __syncthreads();
    const auto zi = PINTT_INV_TABLE(tw, prime);
    if constexpr (kPair) {
        uint32_t* const t[2] = {ldsA, ldsC};
        gs_run<SET, TL, 0, 2, true>(t, zi, sb, rb, p);
    } else {
        uint32_t* const t[1] = {ldsA};
        gs_run<SET, TL, 0, 1, true>(t, zi, sb, rb, p);
    }
#pragma unroll
    for (int e = 0; e < kElemsPerThread; ++e) {
        uint32_t idx = e * TL::NT + threadIdx.x;
        if constexpr (kScaled) {
            const uint2 f = factor(e, true);
            a[idx] = Arith<SET>::scale(ldsA[at(e)], f.x, f.y, p);
            if constexpr (kPair) a2[idx] = Arith<SET>::scale(ldsC[at(e)], f.x, f.y, p);
        } else {
            a[idx] = ldsA[at(e)];
            if constexpr (kPair) a2[idx] = ldsC[at(e)];
        }
    }
}

// Twiddle of stage layer l, block blk, for a stage whose first layer is
// global layer L0 and whose group sits in block `outer` there.
struct TwStage {
    ConstTable z;
    int L0;
    uint32_t outer;
    __device__ __forceinline__ uint2 operator()(int l, uint32_t blk) const {
        return tw_at(z, (1u << (L0 + l)) + (outer << l) + blk);
    }
};

// Geometry of the register row kernels: rows of 2^LC points (LC = 10, 11,
// 12), 16 residues per thread, 256 threads: a row takes 2^(LC-4) threads (one,
// two or four waves) and a workgroup 2^(12-LC) rows.  Stage A runs layers
// 0..5 of a row on its 2^(LC-6) groups of 64 elements 2^(LC-6) apart, stage B
// layers 6..LC-1 on its 64 groups of 2^(LC-6) contiguous elements: a six-layer
// lane-swap stage (LC = 12), a five-layer one (LC = 11) or four layers inside
// each lane (LC = 10).  Positions in a row, for the thread's row-local index
// ell (wave w_r = ell >> 6 within the row), lane L and register i / r:
//   j = (w_r << 4) | (L & 15)                 stage A group
//   lanepos = ((L >> 4) << (LC - 6)) | j
//   stage A entry: k = (i << (LC - 4)) | lanepos   (global loads and stores)
//   stage A post:  k = (post_t<SB>(r) << (LC - 6)) | j
//   stage B entry: k = (u << (LC - 6)) | t', with
//     LC = 12: u = j,                    t' = (i << 2) | (L >> 4)
//     LC = 11: u = (w_r << 5) | (L & 31), t' = (i << 1) | (L >> 5)
//     LC = 10: u = L,                    t' = i
// LDS, padded by PW words per 2^PSH, is visited twice (after stage A and after
// stage B's inverse); SB and the padding are chosen so that both layouts hit
// 32 distinct banks per half-wave (tests/test_reg_layouts.py checks them).
template <int LC>
struct RegRows {
    static_assert(LC >= 10 && LC <= 12, "register rows of 2^10 .. 2^12 points");
    static constexpr int GA = LC - 6;
    static constexpr int LPR = 1 << (LC - 4);  // threads per row
    static constexpr int RPW = 256 / LPR;      // rows per workgroup
    static constexpr int SB = LC == 12 ? 1 : 2;
    static constexpr int PSH = LC == 12 ? 6 : (LC == 11 ? 5 : 4);
    static constexpr uint32_t PW = LC == 12 ? 2u : 1u;
    static constexpr int WORDS = (1 << LC) + (int)PW * ((1 << LC) >> PSH);
    __device__ static __forceinline__ uint32_t lds_at(uint32_t k) { return k + PW * (k >> PSH); }
    // Stage B entry position of register i.
    __device__ static __forceinline__ uint32_t b_entry(int i, uint32_t lane, uint32_t wr) {
        if constexpr (LC == 12) return (((wr << 4) | (lane & 15u)) << 6) | ((uint32_t)i << 2) | (lane >> 4);
        else if constexpr (LC == 11) return (((wr << 5) | (lane & 31u)) << 5) | ((uint32_t)i << 1) | (lane >> 5);
        else return (lane << 4) | (uint32_t)i;
    }
    // Stage B group index u.
    __device__ static __forceinline__ uint32_t b_group(uint32_t lane, uint32_t wr) {
        if constexpr (LC == 12) return (wr << 4) | (lane & 15u);
        else if constexpr (LC == 11) return (wr << 5) | (lane & 31u);
        else return lane;
    }
};

// Stage B of a register row kernel, forward or inverse, on NB tiles;
// forward values enter below B0 p.
template <int LC, int SET, int NB, bool INV, int B0, class TW>
__device__ __forceinline__ void row_stage_b(uint32_t (&v)[NB][16], const TW& tw, uint32_t p) {
    // The inverse stage B takes the pointwise products, below 2p.
    if constexpr (LC == 12) {
        if constexpr (INV) gs6<SET, NB, 1, true>(v, tw, p);
        else ct6<SET, NB, false, 1, B0>(v, tw, p);
    } else if constexpr (LC == 11) {
        if constexpr (INV) gs5<SET, NB, true>(v, tw, p);
        else ct5<SET, NB, B0>(v, tw, p);
    } else {
        auto tl = [&](int, int l, uint32_t blk) { return tw(l, blk); };
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            if constexpr (INV) gs_lane<SET, 4, true>(v[u], tl, p);
            else ct_lane<SET, 4, B0>(v[u], tl, p);
        }
    }
}

// Row pass with register-resident stages (ntt_regs.h) for rows of 2^LC
// points (RegRows).  The pointwise product is formed in registers between the
// forward and the inverse stage B.  Transformed rows of b (kRowTransformB with
// write_b, kRowReuseB, kRowPair with write_b) are stored in register order,
// word (r << (LC - 4)) | ell for register r of the stage B post layout: every
// kernel of this shape reads and writes them the same way.
// SCALED as in k_row_conv: rows scaled by a_r^(+-k) on load and store, block
// 0's twiddles; otherwise the row's own twiddles.
template <int LC, int KIND, int SET, bool SCALED>
__global__ void __launch_bounds__(256) k_row_reg(const JobList<RowJob> jobs, int log2r, Twiddles tw) {
    using RG = RegRows<LC>;
    // Forward values enter stage A below 2p when scaled (mul_fixed), below 8p
    // as the column pass stored them otherwise; stage B starts where A ends.
    constexpr int kB0A = SCALED ? 2 : 8;
    constexpr int kB0B = ct6_exit(kB0A, false);
    constexpr bool kPair = KIND == kRowPair;
    constexpr bool kTransB = KIND == kRowTransformB;
    constexpr int NA = kPair ? 2 : 1;                       // chunk tiles
    constexpr int NL = NA + ((kPair || kTransB) ? 1 : 0);   // tiles through stage A forward
    __shared__ uint32_t lds[NL][RG::RPW][RG::WORDS];
    const RowJob job = PINTT_JOB(RowJob, jobs, blockIdx.z);
    const int prime = blockIdx.y;
    const uint32_t p = prime_of<SET>(prime);
    const uint32_t pinv = neg_inv_mod_2_32(p);
    const uint32_t rw = threadIdx.x >> (LC - 4);            // row within the workgroup
    const uint32_t ell = threadIdx.x & (RG::LPR - 1);        // thread within the row
    const uint32_t row0 = blockIdx.x * RG::RPW + rw;
    const size_t base = ((size_t)prime << (log2r + LC)) + ((size_t)row0 << LC);
    global_ptr<uint32_t> ga[2] = {as_global(job.a) + base, as_global(job.a2) + base};
    global_ptr<uint32_t> gb = as_global(job.b) + base;
    const uint32_t lane = lane_id(), wr = ell >> 6;
    const uint32_t j = (wr << 4) | (lane & 15u);
    const uint32_t lanepos = ((lane >> 4) << RG::GA) | j;
    const int sb = SCALED ? 0 : log2r;
    const uint32_t rb = SCALED ? 0u : row0;
    const uint32_t u = RG::b_group(lane, wr);
    const TwTable zf = PINTT_FWD_TABLE(tw, prime);
    const TwTable zi = PINTT_INV_TABLE(tw, prime);
    const TwStage fA{as_const(zf), sb, rb}, fB{as_const(zf), sb + 6, (rb << 6) | u};
    const TwStage iA{as_const(zi), sb, rb}, iB{as_const(zi), sb + 6, (rb << 6) | u};

    // Scaling factors (SCALED): element k = (i << (LC - 4)) | lanepos of row
    // r is scaled by a_r^k = a_r^lanepos * a_r^(i << (LC - 4)); the second
    // factor is uniform over the row and sits in LDS with its Shoup quotient,
    // the first is the thread's own (Montgomery form).  As in k_row_conv.
    __shared__ uint2 beta[SCALED ? RG::RPW : 1][SCALED ? 32 : 1];
    uint32_t h_fwd = 0, h_inv = 0;
    const uint32_t pinv_pos = 0u - pinv;
    if constexpr (SCALED) {
        const RowTable rt = PINTT_ROW_TABLE(tw, prime);
        constexpr uint32_t kMask = (1u << kSetMaxLog2N[SET]) - 1;
        const int eshift = kSetMaxLog2N[SET] - (log2r + LC);
        const uint32_t rbits = __brev(row0) >> (32 - log2r);
        const uint32_t et = (rbits * lanepos) << eshift;
        h_fwd = row_factor(rt, et, p);
        h_inv = row_factor(rt, (0u - et) & kMask, p);
        if (threadIdx.x < 32u * RG::RPW) {
            const uint32_t br = threadIdx.x >> 5;  // the row whose factor this thread makes
            const uint32_t rbr = __brev(blockIdx.x * RG::RPW + br) >> (32 - log2r);
            const uint32_t e = threadIdx.x & 15u;
            const uint32_t eu = (rbr * (e << (LC - 4))) << eshift;
            const uint32_t bm = row_factor(rt, (threadIdx.x & 16u) ? (0u - eu) & kMask : eu, p);
            const uint32_t bv = mul_mont(bm, 1u, p, pinv);  // out of Montgomery form
            beta[br][threadIdx.x & 31u] = make_uint2(bv, (uint32_t)(((uint64_t)bv << 32) / p));
        }
        __syncthreads();
    }
    auto factor = [&](int e, bool inverse) {
        const uint2 bq = beta[SCALED ? rw : 0][(inverse ? 16 : 0) + e];
        const uint32_t f = mul_shoup(inverse ? h_inv : h_fwd, bq.x, bq.y, p);
        return make_uint2(f, f * pinv_pos);
    };

    // Stage A forward on every tile at once (chunk tiles, then b): their loads
    // are all in flight together and the tiles share each twiddle.
    {
        uint32_t v[NL][16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t k = ((uint32_t)i << (LC - 4)) | lanepos;
#pragma unroll
            for (int t = 0; t < NL; ++t) v[t][i] = t < NA ? ga[t][k] : gb[k];
            if constexpr (SCALED) {
                const uint2 f = factor(i, false);
#pragma unroll
                for (int t = 0; t < NL; ++t) v[t][i] = Arith<SET>::scale(v[t][i], f.x, f.y, p);
            }
        }
        ct6<SET, NL, false, RG::SB, kB0A>(v, fA, p);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t k = (post_t<RG::SB>(r, lane) << RG::GA) | j;
#pragma unroll
            for (int t = 0; t < NL; ++t) lds[t][rw][RG::lds_at(k)] = v[t][r];
        }
    }
    __syncthreads();
    // Stage B: b's transformed row (post layout), then the chunk tiles:
    // forward, pointwise, inverse, back to LDS in the stage B entry layout.
    uint32_t vb[1][16];
    if constexpr (KIND == kRowReuseB) {
#pragma unroll
        for (int r = 0; r < 16; ++r) vb[0][r] = gb[((uint32_t)r << (LC - 4)) | ell];
    } else if constexpr (kPair || kTransB) {
#pragma unroll
        for (int i = 0; i < 16; ++i) vb[0][i] = lds[NL - 1][rw][RG::lds_at(RG::b_entry(i, lane, wr))];
        row_stage_b<LC, SET, 1, false, kB0B>(vb, fB, p);
        if (job.write_b) {
#pragma unroll
            for (int r = 0; r < 16; ++r) gb[((uint32_t)r << (LC - 4)) | ell] = vb[0][r];
        }
    }
    {
        uint32_t v[NA][16];
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int t = 0; t < NA; ++t) v[t][i] = lds[t][rw][RG::lds_at(RG::b_entry(i, lane, wr))];
        row_stage_b<LC, SET, NA, false, kB0B>(v, fB, p);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            if constexpr (KIND == kRowSquare) {
#pragma unroll
                for (int t = 0; t < NA; ++t) v[t][r] = Arith<SET>::pointwise(v[t][r], v[t][r], p, pinv);
            } else {
                const uint32_t ys = Arith<SET>::shared_operand(vb[0][r], p);
#pragma unroll
                for (int t = 0; t < NA; ++t) v[t][r] = Arith<SET>::pointwise_shared(v[t][r], ys, p, pinv);
            }
        }
        row_stage_b<LC, SET, NA, true, kB0B>(v, iB, p);
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int t = 0; t < NA; ++t) lds[t][rw][RG::lds_at(RG::b_entry(i, lane, wr))] = v[t][i];
    }
    __syncthreads();
    // Stage A inverse on the chunk tiles together, then back in place.
    {
        uint32_t v[NA][16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t k = (post_t<RG::SB>(r, lane) << RG::GA) | j;
#pragma unroll
            for (int t = 0; t < NA; ++t) v[t][r] = lds[t][rw][RG::lds_at(k)];
        }
        gs6<SET, NA, RG::SB>(v, iA, p);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t k = ((uint32_t)i << (LC - 4)) | lanepos;
            if constexpr (SCALED) {
                const uint2 f = factor(i, true);
#pragma unroll
                for (int t = 0; t < NA; ++t) ga[t][k] = Arith<SET>::scale(v[t][i], f.x, f.y, p);
            } else {
#pragma unroll
                for (int t = 0; t < NA; ++t) ga[t][k] = v[t][i];
            }
        }
    }
}

// Rows of 2^10 .. 2^12 points on the register kernels (k_row_reg) unless
// PINTT_REG_ROWS=0.
bool reg_rows_enabled() {
    static const bool on = [] {
        const char* v = std::getenv("PINTT_REG_ROWS");
        return !(v && v[0] == '0');
    }();
    return on;
}

// Scaled rows from this transform length on (PINTT_SCALED_MIN), for pairs
// between PINTT_SCALED_PAIR_MIN and PINTT_SCALED_PAIR_MAX.
int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}
const int kScaledRowsMinLog2N = env_int("PINTT_SCALED_MIN", 21);
const int kScaledPairMinLog2N = env_int("PINTT_SCALED_PAIR_MIN", 20);
const int kScaledPairMaxLog2N = env_int("PINTT_SCALED_PAIR_MAX", kMaxLog2N);

template <int LC, int LROWS, int K, int SET>
void launch_rows(const Plan& pl, dim3 grid, const JobList<RowJob>& jobs, const Twiddles& tw, hipStream_t s) {
    // Rows of 2^10 (four per workgroup) and 2^11 points (two per workgroup):
    // the register kernels, with the row's own twiddles below the scaling
    // thresholds.
    if constexpr ((LC == 10 && LROWS == 2) || (LC == 11 && LROWS == 1)) {
        if (set_for(pl.log2n) == kSetNarrow && reg_rows_enabled()) {
            const bool scaled = K == kRowPair ? pl.log2n >= kScaledPairMinLog2N && pl.log2n <= kScaledPairMaxLog2N
                                              : pl.log2n >= kScaledRowsMinLog2N;
            if (scaled)
                hipLaunchKernelGGL((k_row_reg<LC, K, SET, true>), grid, dim3(256), 0, s, jobs, pl.log2r, tw);
            else
                hipLaunchKernelGGL((k_row_reg<LC, K, SET, false>), grid, dim3(256), 0, s, jobs, pl.log2r, tw);
            return;
        }
    }
    if constexpr (LROWS == 0) {
        const bool scaled = K == kRowPair ? pl.log2n >= kScaledPairMinLog2N && pl.log2n <= kScaledPairMaxLog2N
                                          : pl.log2n >= kScaledRowsMinLog2N;
        if constexpr (LC == 12) {
            if (reg_rows_enabled()) {
                if (scaled)
                    hipLaunchKernelGGL((k_row_reg<12, K, SET, true>), grid, dim3(256), 0, s, jobs, pl.log2r, tw);
                else
                    hipLaunchKernelGGL((k_row_reg<12, K, SET, false>), grid, dim3(256), 0, s, jobs, pl.log2r, tw);
                return;
            }
        }
        if (scaled) {
            hipLaunchKernelGGL((k_row_conv<LC, LROWS, K, SET, true>), grid, dim3(Tile<LC, LROWS, false>::NT), 0, s,
                               jobs, pl.log2r, tw);
            return;
        }
    }
    hipLaunchKernelGGL((k_row_conv<LC, LROWS, K, SET, false>), grid, dim3(Tile<LC, LROWS, false>::NT), 0, s, jobs,
                       pl.log2r, tw);
}

}  // namespace

// Whether every row kernel at this plan scales its rows (then it stores its
// products through Arith::scale, below 2p), whatever the row kinds of the
// bucket: the same thresholds and shapes launch_rows applies.
bool rows_scaled(const Plan& pl) {
    if (set_for(pl.log2n) != kSetNarrow) return false;
    const bool shape = pl.log2rows == 0 || (reg_rows_enabled() && ((pl.log2c == 10 && pl.log2rows == 2) ||
                                                                   (pl.log2c == 11 && pl.log2rows == 1)));
    return shape && pl.log2n >= kScaledRowsMinLog2N && pl.log2n >= kScaledPairMinLog2N &&
           pl.log2n <= kScaledPairMaxLog2N;
}

// Tile shapes per prime set.  The wide set only serves 2^23 .. 2^25 points,
// rows of 2^12 .. 2^14 (columns hold at most 2^11 rows).
#define PINTT_ROW_CASES_NARROW(X) \
    X(5, 5) X(5, 6) X(5, 7) X(5, 8) X(6, 5) X(6, 6) X(6, 7) X(7, 4) \
    X(7, 5) X(7, 6) X(8, 3) X(8, 4) X(8, 5) X(9, 2) X(9, 3) X(9, 4) \
    X(10, 1) X(10, 2) X(10, 3) X(11, 0) X(11, 1) X(11, 2) X(12, 0) X(12, 1) \
    X(13, 0) X(14, 0)
#define PINTT_ROW_CASES_WIDE(X) X(12, 0) X(12, 1) X(13, 0) X(14, 0)

// Pair kernels: the row cases with at most 2^13 words per workgroup.
#define PINTT_PAIR_CASES_NARROW(X) \
    X(5, 5) X(5, 6) X(5, 7) X(5, 8) X(6, 5) X(6, 6) X(6, 7) X(7, 4) \
    X(7, 5) X(7, 6) X(8, 3) X(8, 4) X(8, 5) X(9, 2) X(9, 3) X(9, 4) \
    X(10, 1) X(10, 2) X(10, 3) X(11, 0) X(11, 1) X(11, 2) X(12, 0) X(12, 1) \
    X(13, 0)
#define PINTT_PAIR_CASES_WIDE(X) X(12, 0) X(12, 1) X(13, 0)

bool row_case_available(int set, int log2c, int log2rows) {
#define PINTT_CASE(LC, LROWS) \
    if (log2c == LC && log2rows == LROWS) return true;
    if (set == kSetNarrow) {
        PINTT_ROW_CASES_NARROW(PINTT_CASE)
    } else {
        PINTT_ROW_CASES_WIDE(PINTT_CASE)
    }
#undef PINTT_CASE
    return false;
}

bool pair_case_available(int set, int log2c, int log2rows) {
#define PINTT_CASE(LC, LROWS) \
    if (log2c == LC && log2rows == LROWS) return true;
    if (set == kSetNarrow) {
        PINTT_PAIR_CASES_NARROW(PINTT_CASE)
    } else {
        PINTT_PAIR_CASES_WIDE(PINTT_CASE)
    }
#undef PINTT_CASE
    return false;
}

hipError_t launch_row_convolution(const Plan& pl, const JobList<RowJob>& jobs, uint32_t njobs, RowKind kind,
                                  const Twiddles& tw, hipStream_t s) {
    dim3 grid(1u << (pl.log2r - pl.log2rows), kPrimes, njobs);
#define PINTT_LAUNCH(LC, LROWS, K, SET) launch_rows<LC, LROWS, K, SET>(pl, grid, jobs, tw, s)
#define PINTT_CASE_SET(LC, LROWS, SET)                                                \
    if (pl.log2c == LC && pl.log2rows == LROWS) {                                     \
        if (kind == kRowSquare) PINTT_LAUNCH(LC, LROWS, kRowSquare, SET);             \
        else if (kind == kRowTransformB) PINTT_LAUNCH(LC, LROWS, kRowTransformB, SET); \
        else if (kind == kRowReuseB) PINTT_LAUNCH(LC, LROWS, kRowReuseB, SET);       \
        else return hipErrorInvalidValue;                                             \
        return hipGetLastError();                                                     \
    }
#define PINTT_PAIR_CASE_SET(LC, LROWS, SET)                                 \
    if (kind == kRowPair && pl.log2c == LC && pl.log2rows == LROWS) {      \
        PINTT_LAUNCH(LC, LROWS, kRowPair, SET);                             \
        return hipGetLastError();                                           \
    }
#define PINTT_CASE_N(LC, LROWS) PINTT_CASE_SET(LC, LROWS, kSetNarrow)
#define PINTT_CASE_W(LC, LROWS) PINTT_CASE_SET(LC, LROWS, kSetWide)
#define PINTT_PAIR_CASE_N(LC, LROWS) PINTT_PAIR_CASE_SET(LC, LROWS, kSetNarrow)
#define PINTT_PAIR_CASE_W(LC, LROWS) PINTT_PAIR_CASE_SET(LC, LROWS, kSetWide)
    if (set_for(pl.log2n) == kSetNarrow) {
        PINTT_PAIR_CASES_NARROW(PINTT_PAIR_CASE_N)
        PINTT_ROW_CASES_NARROW(PINTT_CASE_N)
    } else {
        PINTT_PAIR_CASES_WIDE(PINTT_PAIR_CASE_W)
        PINTT_ROW_CASES_WIDE(PINTT_CASE_W)
    }
#undef PINTT_PAIR_CASE_W
#undef PINTT_PAIR_CASE_N
#undef PINTT_CASE_W
#undef PINTT_CASE_N
#undef PINTT_PAIR_CASE_SET
#undef PINTT_CASE_SET
#undef PINTT_LAUNCH
    return hipErrorInvalidValue;
}

}  // namespace pintt
