// ntt_carry.hip - CRT + carry pass and the twiddle-table builder (gfx950).
#include "ntt_device.h"

#include <cstdlib>

namespace pintt {
namespace {

// Pass 4: CRT and carry propagation, per group, in two kernels.
//
// Word j of a group's product before carries is the sum, over its chunk
// products i, of X0_i[k] + X1_i[k-1] + X2_i[k-2] with k = j - off_i, where
// X0..X2 are the 32-bit words of the CRT value of chunk product i at position k.
// At most two chunk products overlap at any word, so a word sum is below
// 6 * 2^32 and the carry between words stays at most 6.  The effect of a run of
// words on an incoming carry c in {0..7} is c -> b + [c >= t]: b is the run's
// carry-out for carry-in 0, and a larger carry-in adds exactly one more carry
// when it overflows the run's first word while every later word is all ones
// (t = 8: never).  These step maps are closed under composition and compose
// with a few compares.  A word's map is constant unless its sum's low half is
// within 7 of 2^32, and a run's map is constant as soon as one of its words'
// maps is.
//
// k_crt_carry: a workgroup of CT threads sums 8 CT consecutive words (the
// kCarryWindowWords of the window, then its block); each thread loads the residues of its own 8 positions of
// every chunk product that reaches them (two 16-byte loads per plane), forms
// their CRT values and adds them into its word sums; the terms that spill into
// the next thread's first two words pass by lane shuffles (and through LDS
// across waves).  The threads' carry maps are scanned across the workgroup.
// The window's map is almost always constant: then it gives the block's exact
// carry-in and the block's words are written final.  Otherwise they are
// written as if the carry-in were 0 and the block's map is flagged.  Every
// block stores its map (the effect of its own words) in the group's map array.
// k_carry_fixup: one wave per block; a flagged block composes the maps of the
// group's earlier blocks into its true carry-in and adds it at its first word,
// rippling through words that are all ones, never past its last word (what
// spills out is already counted in the next block's carry-in).  No workgroup
// ever waits for another.

// Map encoding: b in bits 0-2, t in bits 3-6; bit 7 marks the identity.
constexpr uint32_t kFnIdentity = 1u << 7;
__device__ __forceinline__ uint32_t fn_make(uint32_t b, uint32_t t) { return b | (t << 3); }
__device__ __forceinline__ uint32_t fn_apply(uint32_t f, uint32_t c) {
    return (f & kFnIdentity) ? c : (f & 7u) + (c >= ((f >> 3) & 15u) ? 1u : 0u);
}

// later o earlier
__device__ __forceinline__ uint32_t fn_compose(uint32_t later, uint32_t earlier) {
    const uint32_t gb = earlier & 7u, gt = (earlier >> 3) & 15u;
    const uint32_t fb = later & 7u, ft = (later >> 3) & 15u;
    const uint32_t lo = fb + (gb >= ft ? 1u : 0u), hi = fb + (gb + 1 >= ft ? 1u : 0u);
    const uint32_t r = fn_make(lo, (gt >= 8 || lo == hi) ? 8u : gt);
    return (earlier & kFnIdentity) ? later : ((later & kFnIdentity) ? earlier : r);
}

// Inclusive scan of carry maps over the block in thread order.  Returns the
// thread's exclusive prefix; *total receives the block aggregate.
template <int NT>
__device__ __forceinline__ uint32_t block_scan_fn(uint32_t f, uint32_t* lds_waves, uint32_t* total) {
    constexpr int W = 64;
    const int lane = threadIdx.x & (W - 1), wave = threadIdx.x / W;
    uint32_t incl = f;
#pragma unroll
    for (int off = 1; off < W; off <<= 1) {
        uint32_t other = __shfl_up(incl, off, W);
        if (lane >= off) incl = fn_compose(incl, other);
    }
    if (lane == W - 1) lds_waves[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = kFnIdentity;
        for (int i = 0; i < NT / W; ++i) {
            acc = fn_compose(lds_waves[i], acc);
            lds_waves[i] = acc;  // inclusive prefix over waves
        }
    }
    __syncthreads();
    uint32_t excl = __shfl_up(incl, 1, W);
    if (lane == 0) excl = kFnIdentity;
    if (wave > 0) excl = fn_compose(excl, lds_waves[wave - 1]);
    *total = lds_waves[NT / W - 1];
    return excl;
}

// x r mod p in [0, 2p) for any 32-bit r (Shoup's product without its last
// reduction).
__device__ __forceinline__ uint32_t shoup_lazy(uint32_t r, uint2 w, uint32_t p) {
    return r * w.x - __umulhi(r, w.y) * p;
}

// CRT of one position: the exact value x < p0 p1 p2 as 3 words.  Residues may
// be any 32-bit values (each is first multiplied by Shoup).
//
// Narrow set: y_i in [0, 2 p_i) and sum_i y_i M_i = x + k M exactly, with
// sum_i y_i / p_i = k + x / M.  Every product the set serves has x < 2^85 <
// 0.526 M, so the sum lies in [k, k + 0.526); with 0.25 added its distance to
// the integers around it is above 0.22, and the single-precision estimate
// (three conversions and three fused multiply-adds, relative errors near 2^-24
// on terms below 2, about 2e-6 in all) truncates to k.  The 96-bit sum then
// needs no reduction: x = sum_i y_i M_i + k (2^96 - M) mod 2^96.  Its partial
// sums fit 64 bits: y_i mlo_i < 2^30 2^32, three of them below 2^63.6, plus
// k negm_0 < 6 2^32.
template <int SET>
__device__ __forceinline__ void crt_words(uint32_t r0, uint32_t r1, uint32_t r2, const CrtConst& cc,
                                          uint32_t& x0, uint32_t& x1, uint32_t& x2) {
    constexpr uint32_t p0 = kPrimeSet[SET][0], p1 = kPrimeSet[SET][1], p2 = kPrimeSet[SET][2];
    if constexpr (SET == kSetNarrow) {
        const uint32_t y0 = shoup_lazy(r0, cc.e[0], p0);
        const uint32_t y1 = shoup_lazy(r1, cc.e[1], p1);
        const uint32_t y2 = shoup_lazy(r2, cc.e[2], p2);
        float f = __builtin_fmaf((float)y0, cc.inv[0], 0.25f);
        f = __builtin_fmaf((float)y1, cc.inv[1], f);
        f = __builtin_fmaf((float)y2, cc.inv[2], f);
        const uint32_t k = (uint32_t)f;
        const uint64_t lo = (uint64_t)y0 * cc.mlo[0] + (uint64_t)y1 * cc.mlo[1] + (uint64_t)y2 * cc.mlo[2] +
                            (uint64_t)k * cc.negm[0];
        const uint64_t hi = (uint64_t)y0 * cc.mhi[0] + (uint64_t)y1 * cc.mhi[1] + (uint64_t)y2 * cc.mhi[2] +
                            (lo >> 32) + (uint64_t)k * cc.negm[1] + ((uint64_t)(k * cc.negm[2]) << 32);
        x0 = (uint32_t)lo;
        x1 = (uint32_t)hi;
        x2 = (uint32_t)(hi >> 32);
        return;
    }
    uint32_t y0 = mul_shoup(r0, cc.k0.x, cc.k0.y, p0);
    uint32_t y1 = sub_mod(mul_shoup(r1, cc.k1.x, cc.k1.y, p1), mul_shoup(y0, cc.c1.x, cc.c1.y, p1), p1);
    uint32_t y2 = sub_mod(mul_shoup(r2, cc.k2.x, cc.k2.y, p2), mul_shoup(y0, cc.c20.x, cc.c20.y, p2), p2);
    y2 = sub_mod(y2, mul_shoup(y1, cc.c21.x, cc.c21.y, p2), p2);
    // x = y0 + p0 * (y1 + p1 * y2)
    uint64_t t = (uint64_t)y2 * p1 + y1;
    uint64_t lo = (uint64_t)(uint32_t)t * p0 + y0;
    uint64_t hi = (t >> 32) * (uint64_t)p0 + (lo >> 32);
    x0 = (uint32_t)lo;
    x1 = (uint32_t)hi;
    x2 = (uint32_t)(hi >> 32);
}

// Index of the group that owns carry block `blk` (jobs sorted by blk0, job 0
// at block 0), called by every lane of a wave.
__device__ __forceinline__ uint32_t find_job(const JobList<CarryJob>& jobs, uint32_t njobs, uint32_t blk) {
    uint32_t lo = 0;  // last job with blk0 <= blk
    if (jobs.ninl) {
#pragma unroll
        for (int i = 1; i < kInlineJobs; ++i)
            if ((uint32_t)i < jobs.ninl && jobs.inl[i].blk0 <= blk) lo = i;
        return lo;
    }
    const CarryJob* dev = jobs.dev;
    return wave_search(njobs, blk, [dev](uint32_t i) { return dev[i].blk0; });
}

constexpr int kWT = kCarryWordsPerThread;
constexpr int kWinThreads = kCarryWindowWords / kWT;      // threads whose words form the window
static_assert(kWT == 8 && kCarryWindowWords % kWT == 0 && kWinThreads <= 32, "carry layout");
constexpr uint32_t kFnUnresolved = 1u << 8;  // map flag: the block's carry-in is not yet in its words

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, int d) {
    const uint32_t lo = __shfl_up((uint32_t)v, d, 64), hi = __shfl_up((uint32_t)(v >> 32), d, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Whether chunk product i of `job` reaches the words [first, first + CT kWT)
// of a block (its positions first - 2 .. spill into them).
template <int CT>
__device__ __forceinline__ bool chunk_reaches(const CarryJob& job, int i, long long first) {
    const long long k0 = first - (long long)job.off[i];
    return !(k0 + CT * kWT <= 0 || k0 - 2 >= (long long)job.len[i]);
}

// The residues of chunk product i at the thread's kWT positions (two 16-byte
// loads per plane where all are inside the chunk) and, for threads 0 and 1,
// at position first - 2 + t before the block's first word.
template <int CT>
__device__ __forceinline__ void load_chunk(const CarryJob& job, int i, long long first, int t, size_t n,
                                           uint32_t (&r)[3][kWT], uint32_t (&h)[3]) {
    const long long k0 = first - (long long)job.off[i];
    const long long len = job.len[i];
    const global_ptr<const uint32_t> res = as_global(job.res[i]);
    const long long k = k0 + kWT * t;
    if (k >= 0 && k + kWT <= len) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int e = 0; e < kWT; e += 4) {
                const u32x4 x = *(global_ptr<const u32x4>)(res + q * n + k + e);
                r[q][e] = x.x;
                r[q][e + 1] = x.y;
                r[q][e + 2] = x.z;
                r[q][e + 3] = x.w;
            }
    } else {
#pragma unroll
        for (int u = 0; u < kWT; ++u) {
            const bool in = k + u >= 0 && k + u < len;
#pragma unroll
            for (int q = 0; q < 3; ++q) r[q][u] = in ? res[q * n + k + u] : 0u;
        }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) h[q] = 0u;
    if (t < 2) {
        const long long kk = k0 - 2 + t;
        if (kk >= 0 && kk < len) {
#pragma unroll
            for (int q = 0; q < 3; ++q) h[q] = res[q * n + kk];
        }
    }
}

// CRT words of the thread's positions added into its word sums s; the terms
// of the next thread's words 0 and 1 into ea, eb; threads 0 and 1 add the
// terms of the positions before the block into ha, hb.
template <int SET>
__device__ __forceinline__ void add_chunk(const uint32_t (&r)[3][kWT], const uint32_t (&h)[3], int t,
                                          const CrtConst& cc, uint64_t (&s)[kWT], uint64_t& ea, uint64_t& eb,
                                          uint64_t& ha, uint64_t& hb) {
#pragma unroll
    for (int u = 0; u < kWT; ++u) {
        uint32_t x0, x1, x2;
        crt_words<SET>(r[0][u], r[1][u], r[2][u], cc, x0, x1, x2);
        s[u] += x0;
        if (u + 1 < kWT) s[u + 1] += x1;
        else ea += x1;
        if (u + 2 < kWT) s[u + 2] += x2;
        else if (u + 2 == kWT) ea += x2;
        else eb += x2;
    }
    if (t < 2) {  // positions first - 2 and first - 1
        uint32_t x0, x1, x2;
        crt_words<SET>(h[0], h[1], h[2], cc, x0, x1, x2);
        if (t == 0) {
            ha += x2;
        } else {
            ha += x1;
            hb += x2;
        }
    }
}

// grid = carry workgroups of the launch; jobs sorted by blk0.  SET: the prime
// set the residues belong to; CT: threads per workgroup.
template <int SET, int CT>
__global__ void __launch_bounds__(CT)
    k_crt_carry(const JobList<CarryJob> jobs, uint32_t njobs, int log2n, CrtConst cc) {
    constexpr int kCT = CT;
    constexpr int kWB = carry_block_words(CT);  // words the workgroup writes
    static_assert(kCT % 64 == 0, "whole waves");
    __shared__ uint32_t waves[kCT / 64];
    __shared__ uint64_t edge[kCT / 64][2];  // a wave's last lane's spill terms, for the next wave's lane 0
    __shared__ uint64_t head[2][2];         // the spill terms of the two positions before the first word
    __shared__ uint32_t sh_win;
    const CarryJob job = PINTT_JOB(CarryJob, jobs, find_job(jobs, njobs, blockIdx.x));
    const uint32_t lb = blockIdx.x - job.blk0;  // block within the group
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const long long first = (long long)lb * kWB - kCarryWindowWords;  // first word summed
    const long long P = first + kWT * t;                               // the thread's first word
    const size_t n = (size_t)1 << log2n;

    uint64_t s[kWT];
#pragma unroll
    for (int e = 0; e < kWT; ++e) s[e] = 0;
    uint64_t ea = 0, eb = 0;  // terms of the next thread's words 0 and 1
    uint64_t ha = 0, hb = 0;  // threads 0, 1: terms of words `first` and `first` + 1
    // The first two chunk products that reach the words: every load issued
    // before any is used, so that a wave waits for memory once.  Products cut
    // into three or four chunks load the others after them.  Constant chunk
    // indices keep the job (a kernel-argument copy) in registers.
    {
        uint32_t r[2][3][kWT], h[2][3];
        bool use[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            use[i] = (uint32_t)i < job.m && chunk_reaches<kCT>(job, i, first);
            if (use[i]) load_chunk<kCT>(job, i, first, t, n, r[i], h[i]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (use[i]) add_chunk<SET>(r[i], h[i], t, cc, s, ea, eb, ha, hb);
    }
#pragma unroll
    for (int i = 2; i < kMaxChunks; ++i) {
        if ((uint32_t)i >= job.m) break;
        if (!chunk_reaches<kCT>(job, i, first)) continue;
        uint32_t r[3][kWT], h[3];
        load_chunk<kCT>(job, i, first, t, n, r, h);
        add_chunk<SET>(r, h, t, cc, s, ea, eb, ha, hb);
    }
    // Spill terms from the thread before: lane shuffles inside a wave, LDS
    // across waves and, for the first thread, the two positions before.
    const uint64_t pa = shfl_up_u64(ea, 1), pb = shfl_up_u64(eb, 1);
    if (lane == 63) {
        edge[wave][0] = ea;
        edge[wave][1] = eb;
    }
    if (t < 2) {
        head[t][0] = ha;
        head[t][1] = hb;
    }
    __syncthreads();
    uint64_t ia = pa, ib = pb;
    if (lane == 0) {
        ia = wave ? edge[wave - 1][0] : head[0][0] + head[1][0];
        ib = wave ? edge[wave - 1][1] : head[0][1] + head[1][1];
    }
    s[0] += ia;
    s[1] += ib;
#pragma unroll
    for (int e = 0; e < kWT; ++e)
        if (P + e < 0 || (uint64_t)(P + e) >= job.nwords) s[e] = 0;

    // Carry map of this thread's run (computed with carry-in 0).
    uint32_t f;
    {
        uint64_t c = 0;
        uint32_t w0 = 0;
        bool ones = true;
#pragma unroll
        for (int e = 0; e < kWT; ++e) {
            uint64_t v = s[e] + c;
            uint32_t o = (uint32_t)v;
            c = v >> 32;
            if (e == 0) w0 = o;
            else ones = ones && (o == 0xFFFFFFFFu);
        }
        // smallest carry-in that overflows word 0: 2^32 - w0 (8 stands for never)
        const uint32_t room = 0u - w0;  // 2^32 - w0 mod 2^32; 0 when w0 == 0
        const uint32_t tt = (ones && w0 != 0 && room <= 7u) ? room : 8u;
        f = fn_make((uint32_t)c, tt);
    }
    // The window's map (threads 0 .. kWinThreads - 1, in wave 0).
    uint32_t wm = f;
#pragma unroll
    for (int off = 1; off < kWinThreads; off <<= 1) {
        const uint32_t other = __shfl_up(wm, off, 64);
        if (lane >= off) wm = fn_compose(wm, other);
    }
    if (t == kWinThreads - 1) sh_win = wm;
    // Maps of the block's own threads; the window counts as the identity.
    uint32_t agg;
    const bool in_window = t < kWinThreads;
    const uint32_t excl = block_scan_fn<kCT>(in_window ? kFnIdentity : f, waves, &agg);
    const uint32_t win = sh_win;
    const bool resolved = ((win >> 3) & 15u) >= 8u;  // constant: the carry-in is the same for every carry into the window
    const uint32_t cin = resolved ? (win & 7u) : 0u;
    if (t == 0) as_global(job.maps)[lb] = agg | (resolved ? 0u : kFnUnresolved);
    if (in_window) return;
    uint32_t w[kWT];
    {
        uint64_t c = fn_apply(excl, cin);
#pragma unroll
        for (int e = 0; e < kWT; ++e) {
            const uint64_t v = s[e] + c;
            w[e] = (uint32_t)v;
            c = v >> 32;
        }
    }
    const global_ptr<uint32_t> out = as_global(job.out);
    const uint64_t j = (uint64_t)P;  // >= 0 outside the window
    if (j + kWT <= job.nwords) {
        *(global_ptr<u32x4>)(out + j) = u32x4{w[0], w[1], w[2], w[3]};
        *(global_ptr<u32x4>)(out + j + 4) = u32x4{w[4], w[5], w[6], w[7]};
    } else {
#pragma unroll
        for (int e = 0; e < kWT; ++e)
            if (j + e < job.nwords) out[j + e] = w[e];
    }
}

// One wave per block of the first kernel: a flagged block's true carry-in,
// rippled into its words (blocks of carry_block_words(CT) words).
template <int CT>
__global__ void __launch_bounds__(64) k_carry_fixup(const JobList<CarryJob> jobs, uint32_t njobs) {
    constexpr int kWB = carry_block_words(CT);
    const uint32_t lane = threadIdx.x;
    {
        // The groups' maps lie one after another in block order (maps - blk0
        // is the same for every job of a launch, and job 0 starts at block
        // 0), so the flag needs no job lookup: almost every block stops here.
        const CarryJob j0 = PINTT_JOB(CarryJob, jobs, 0u);
        if (!(as_global(j0.maps)[blockIdx.x] & kFnUnresolved)) return;
    }
    const CarryJob job = PINTT_JOB(CarryJob, jobs, find_job(jobs, njobs, blockIdx.x));
    const uint32_t lb = blockIdx.x - job.blk0;
    if (lb == 0) return;  // never flagged: its window lies before word 0
    const global_ptr<const uint32_t> maps = as_global(job.maps);
    // Maps of blocks 0 .. lb-1 in order, 1024 per round: 16 consecutive
    // maps per lane, all loaded before any is composed.
    uint32_t total = kFnIdentity;  // composition of the maps of the rounds so far
    for (uint32_t base = 0; base < lb; base += 1024) {
        const uint32_t lo = base + 16 * lane;
        uint32_t m[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) m[i] = lo + i < lb ? maps[lo + i] : kFnIdentity;
        uint32_t acc = kFnIdentity;
#pragma unroll
        for (int i = 0; i < 16; ++i) acc = fn_compose(m[i], acc);
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {  // lane 0 ends with lane 63's o ... o lane 0's
            const uint32_t other = __shfl_down(acc, off, 64);
            if (lane + off < 64) acc = fn_compose(other, acc);
        }
        total = fn_compose(__shfl(acc, 0, 64), total);
    }
    uint32_t c = fn_apply(total, 0);
    const global_ptr<uint32_t> out = as_global(job.out);
    const uint64_t j0 = (uint64_t)lb * kWB;
    const uint64_t jend = min(j0 + (uint64_t)kWB, job.nwords);
    for (uint64_t j = j0; c != 0 && j < jend; j += 64) {
        const uint64_t jw = j + lane;
        const bool in = jw < jend;
        const uint32_t w = in ? out[jw] : 0u;
        // Word j takes c; a carry out of it runs through the all-ones words after it.
        const uint32_t w0 = __shfl(w, 0, 64) + c;
        const bool over = w0 < c;
        const uint64_t stop = __ballot(in && lane > 0 && w != 0xFFFFFFFFu);
        const uint32_t first = stop ? (uint32_t)__builtin_ctzll(stop) : 64u;  // first word the carry ends in
        if (lane == 0) out[jw] = w0;
        if (over && in && lane > 0 && lane <= first) out[jw] = lane < first ? 0u : w + 1u;
        c = (over && first == 64u) ? 1u : 0u;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
namespace {
__device__ __forceinline__ uint32_t powmod_dev(uint32_t b, uint32_t e, uint32_t p) {
    uint64_t r = 1, x = b;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

// Powers of wK in two tables, lo[i] = wK^i (i < 2^12) and hi[i] = wK^(i 2^12)
// (i <= 2^(K-12)): every exponent e <= 2^K is hi[e >> 12] * lo[e & 4095].
__global__ void k_power_tables(uint32_t* __restrict__ lo, uint32_t* __restrict__ hi, uint32_t p, uint32_t wk,
                               uint32_t nhi) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kPowLo) lo[i] = powmod_dev(wk, i, p);
    if (i < nhi) hi[i] = powmod_dev(wk, i << kPowLoBits, p);
}

// Z[i], i = 2^s + b: w_{2^(s+1)}^bitrev_s(b) = wK^(bitrev_s(b) * 2^(K - 1 - s)),
// wK a primitive 2^K-th root of unity (K: the set's longest transform); the
// inverse table holds wK^(2^K - e).
__global__ void k_build_twiddles(uint2* __restrict__ fwd, uint2* __restrict__ inv, const uint32_t* __restrict__ lo,
                                 const uint32_t* __restrict__ hi, uint32_t p, int log2k, uint32_t count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t ef = 0;
    if (i > 1) {
        int s = 31 - __clz(i);
        uint32_t b = i - (1u << s);
        uint32_t rev = __brev(b) >> (32 - s);
        ef = rev << (log2k - 1 - s);
    }
    uint32_t ei = (ef == 0) ? 0u : ((1u << log2k) - ef);
    auto power = [&](uint32_t e) {
        return (uint32_t)((uint64_t)hi[e >> kPowLoBits] * lo[e & (kPowLo - 1)] % p);
    };
    uint32_t z = (i == 0) ? 1u : power(ef);
    uint32_t zi = (i == 0) ? 1u : power(ei);
    fwd[i] = make_uint2(z, (uint32_t)(((uint64_t)z << 32) / p));
    inv[i] = make_uint2(zi, (uint32_t)(((uint64_t)zi << 32) / p));
}

// Row factor table: lo[i] = wK^i 2^32 mod p (i < 2^12), then hi[i] =
// (wK^(i 2^12), floor(wK^(i 2^12) 2^32 / p)) for i < 2^(K-12).
__global__ void k_row_factor_table(uint32_t* __restrict__ row, const uint32_t* __restrict__ lo,
                                   const uint32_t* __restrict__ hi, uint32_t p, uint32_t nhi) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kPowLo) row[i] = (uint32_t)(((uint64_t)lo[i] << 32) % p);
    if (i < nhi) {
        uint2* h = reinterpret_cast<uint2*>(row + kPowLo);
        h[i] = make_uint2(hi[i], (uint32_t)(((uint64_t)hi[i] << 32) / p));
    }
}

}  // namespace

hipError_t launch_build_twiddles(uint2* fwd, uint2* inv, uint32_t* row, uint32_t* pow_scratch, int set, int prime,
                                 int log2count, hipStream_t s) {
    const uint32_t p = kPrimeSet[set][prime];
    const int log2k = kSetMaxLog2N[set];
    if (log2count < 0 || log2count > log2k) return hipErrorInvalidValue;
    // wK = g^((p-1)/2^K).
    uint64_t r = 1, x = kGenSet[set][prime];
    uint64_t e = (uint64_t)(p - 1) >> log2k;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    const uint32_t nhi = (1u << (log2k - kPowLoBits)) + 1;
    uint32_t* lo = pow_scratch;
    uint32_t* hi = pow_scratch + kPowLo;
    const uint32_t nt = nhi > kPowLo ? nhi : kPowLo;
    hipLaunchKernelGGL(k_power_tables, dim3((nt + 255) / 256), dim3(256), 0, s, lo, hi, p, (uint32_t)r, nhi);
    const hipError_t e1 = hipGetLastError();
    if (e1 != hipSuccess) return e1;
    const uint32_t nrow = 1u << (log2k - kPowLoBits);
    hipLaunchKernelGGL(k_row_factor_table, dim3((nt + 255) / 256), dim3(256), 0, s, row, lo, hi, p, nrow);
    const hipError_t e2 = hipGetLastError();
    if (e2 != hipSuccess) return e2;
    uint32_t count = 1u << log2count;
    hipLaunchKernelGGL(k_build_twiddles, dim3((count + 255) / 256), dim3(256), 0, s, fwd, inv, lo, hi, p, log2k,
                       count);
    return hipGetLastError();
}

int carry_threads() {
    static const int t = [] {
        const char* e = std::getenv("PINTT_CARRY_THREADS");
        const int v = e ? std::atoi(e) : 256;
        return v == 64 || v == 128 ? v : 256;
    }();
    return t;
}

template <int CT>
hipError_t launch_crt_carry_ct(const JobList<CarryJob>& jobs, uint32_t njobs, uint32_t nblocks, int log2n,
                               const CrtConst& cc, hipStream_t s) {
    if (set_for(log2n) == kSetNarrow)
        hipLaunchKernelGGL((k_crt_carry<kSetNarrow, CT>), dim3(nblocks), dim3(CT), 0, s, jobs, njobs, log2n, cc);
    else
        hipLaunchKernelGGL((k_crt_carry<kSetWide, CT>), dim3(nblocks), dim3(CT), 0, s, jobs, njobs, log2n, cc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_carry_fixup<CT>, dim3(nblocks), dim3(64), 0, s, jobs, njobs);
    return hipGetLastError();
}

hipError_t launch_crt_carry(const JobList<CarryJob>& jobs, uint32_t njobs, uint32_t nblocks, int log2n,
                            const CrtConst& cc, hipStream_t s) {
    switch (carry_threads()) {
        case 64: return launch_crt_carry_ct<64>(jobs, njobs, nblocks, log2n, cc, s);
        case 128: return launch_crt_carry_ct<128>(jobs, njobs, nblocks, log2n, cc, s);
        default: return launch_crt_carry_ct<256>(jobs, njobs, nblocks, log2n, cc, s);
    }
}

}  // namespace pintt
