// ntt_carry.hip - CRT + carry pass and the twiddle-table builder (gfx950).
#include "ntt_device.h"

namespace pintt {
namespace {

// Pass 4: CRT and carry propagation in one pass, per group.
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
// with a few compares: a wave-shuffle scan inside
// the block, then a decoupled look-back across the group's blocks (numbered in
// the order they start, from the group's ticket counter, so a block only waits
// for blocks that already run and publish their aggregate before waiting).
// A status word carries epoch, flag and map in one 64-bit atomic, so relaxed
// atomics suffice: release/acquire would add L2 write-backs and invalidations
// on every publication and look-back round.

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

constexpr uint64_t kFlagAggregate = 1, kFlagPrefix = 2;
constexpr uint32_t kLookbackSpinLimit = 1u << 26;

__device__ __forceinline__ uint64_t status_word(uint32_t epoch, uint64_t flag, uint32_t fn) {
    return ((uint64_t)epoch << 32) | (flag << 24) | fn;
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

// Garner CRT of one position: the exact value x < p0 p1 p2 < 2^93 as 3 words.
// Residues may be any 32-bit values (each is first multiplied by Shoup).
template <int SET>
__device__ __forceinline__ void crt_words(uint32_t r0, uint32_t r1, uint32_t r2, const CrtConst& cc,
                                          uint32_t& x0, uint32_t& x1, uint32_t& x2) {
    constexpr uint32_t p0 = kPrimeSet[SET][0], p1 = kPrimeSet[SET][1], p2 = kPrimeSet[SET][2];
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

// Ordered composition over the block of one map per thread, thread 0 being the
// latest: returns f_0 o f_1 o ... o f_255 to every thread.
__device__ __forceinline__ uint32_t block_compose_desc(uint32_t g, uint32_t* lds_waves) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t o = __shfl_down(g, off, 64);
        if (lane + off < 64) g = fn_compose(g, o);
    }
    if (lane == 0) lds_waves[wave] = g;
    __syncthreads();
    uint32_t r = kFnIdentity;
#pragma unroll
    for (int w = kCarryThreads / 64 - 1; w >= 0; --w) r = fn_compose(lds_waves[w], r);
    __syncthreads();
    return r;
}

// Residues of one chunk product that a thread needs: positions k0-2, k0-1 (the
// words they spill into) and k0 .. k0+WT-1, three primes each; zero outside the
// chunk product (residues 0 give x = 0).
template <int WT>
struct ChunkRes {
    uint32_t h[2][3];
    uint32_t r[3][WT];
    bool any;  // some position in range
};

// Loads only, so that every chunk's loads are in flight before any CRT.
// k0: position of the thread's first word inside the chunk product.
template <int WT>
__device__ __forceinline__ void load_chunk(ChunkRes<WT>& c, global_ptr<const uint32_t> res, size_t n, uint32_t len,
                                           long long k0) {
    c.any = !(k0 + WT <= 0 || k0 - 2 >= (long long)len);
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        const long long k = k0 - 2 + d;
        const bool in = c.any && k >= 0 && k < (long long)len;
#pragma unroll
        for (int q = 0; q < 3; ++q) c.h[d][q] = in ? res[q * n + k] : 0u;
    }
    const bool full = c.any && k0 >= 0 && k0 + WT <= (long long)len;
    if (full) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int v = 0; v < WT / 4; ++v) {
                const u32x4 x = *(global_ptr<const u32x4>)(res + q * n + k0 + 4 * v);
                c.r[q][4 * v] = x.x;
                c.r[q][4 * v + 1] = x.y;
                c.r[q][4 * v + 2] = x.z;
                c.r[q][4 * v + 3] = x.w;
            }
    } else {
#pragma unroll
        for (int e = 0; e < WT; ++e) {
            const long long k = k0 + e;
            const bool in = c.any && k >= 0 && k < (long long)len;
#pragma unroll
            for (int q = 0; q < 3; ++q) c.r[q][e] = in ? res[q * n + k] : 0u;
        }
    }
}

// Adds the chunk product's word contributions X0[k] + X1[k-1] + X2[k-2] to the
// thread's word sums s[0 .. WT).
template <int WT, int SET>
__device__ __forceinline__ void add_chunk_words(uint64_t* s, const ChunkRes<WT>& c, const CrtConst& cc) {
    uint32_t xm1[3], xm2[3];  // x at positions k-1, k-2
    crt_words<SET>(c.h[0][0], c.h[0][1], c.h[0][2], cc, xm2[0], xm2[1], xm2[2]);
    crt_words<SET>(c.h[1][0], c.h[1][1], c.h[1][2], cc, xm1[0], xm1[1], xm1[2]);
#pragma unroll
    for (int e = 0; e < WT; ++e) {
        uint32_t x0, x1, x2;
        crt_words<SET>(c.r[0][e], c.r[1][e], c.r[2][e], cc, x0, x1, x2);
        s[e] += (uint64_t)x0 + xm1[1] + xm2[2];
        xm2[0] = xm1[0]; xm2[1] = xm1[1]; xm2[2] = xm1[2];
        xm1[0] = x0; xm1[1] = x1; xm1[2] = x2;
    }
}

// grid = total carry blocks of the launch; jobs sorted by blk0.  SET: the
// prime set the residues belong to.
template <int SET>
__global__ void __launch_bounds__(kCarryThreads)
    k_crt_carry(const JobList<CarryJob> jobs, uint32_t njobs, int log2n, CrtConst cc,
                CarryGlobal* __restrict__ glob) {
    constexpr int WB = kCarryWordsPerBlock;
    constexpr int WT = kCarryWordsPerThread;
    __shared__ uint32_t waves[kCarryThreads / 64];
    __shared__ uint32_t sh_blk, sh_epoch, sh_job, sh_first_prefix, sh_first_wait;
    if (threadIdx.x == 0) {
        uint32_t lo = 0;  // last job with blk0 <= blockIdx.x
        if (jobs.ninl) {
#pragma unroll
            for (int i = 1; i < kInlineJobs; ++i)
                if ((uint32_t)i < jobs.ninl && jobs.inl[i].blk0 <= blockIdx.x) lo = i;
        } else {
            uint32_t hi = njobs - 1;
            while (lo < hi) {
                uint32_t mid = (lo + hi + 1) >> 1;
                if (jobs.dev[mid].blk0 <= blockIdx.x) lo = mid;
                else hi = mid - 1;
            }
        }
        sh_job = lo;
        unsigned long long* ticket = jobs.ninl ? nullptr : jobs.dev[lo].ticket;
        if (jobs.ninl) {
#pragma unroll
            for (int i = 0; i < kInlineJobs; ++i)
                if ((uint32_t)i == lo) ticket = jobs.inl[i].ticket;
        }
        sh_blk = (uint32_t)__hip_atomic_fetch_add(as_global(ticket), 1ull, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        sh_epoch = __hip_atomic_load(&glob->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const CarryJob job = PINTT_JOB(CarryJob, jobs, sh_job);
    const uint32_t blk = sh_blk, epoch = sh_epoch;
    const size_t n = (size_t)1 << log2n;
    const uint64_t j0 = (uint64_t)blk * WB + (uint64_t)WT * threadIdx.x;  // this thread's first word
    const global_ptr<uint64_t> status = as_global(job.status);

    uint64_t s[WT];
#pragma unroll
    for (int e = 0; e < WT; ++e) s[e] = 0;
    // Chunk products two at a time: both loads issued, then both CRTs.
    // Constant chunk indices keep the job (a kernel-argument copy) in registers.
#pragma unroll
    for (uint32_t i = 0; i < (uint32_t)kMaxChunks; i += 2) {
        if (i < job.m) {
            ChunkRes<WT> c0, c1;
            load_chunk<WT>(c0, as_global(job.res[i]), n, job.len[i], (long long)j0 - (long long)job.off[i]);
            const bool two = i + 1 < job.m;
            if (two)
                load_chunk<WT>(c1, as_global(job.res[i + 1]), n, job.len[i + 1],
                               (long long)j0 - (long long)job.off[i + 1]);
            if (c0.any) add_chunk_words<WT, SET>(s, c0, cc);
            if (two && c1.any) add_chunk_words<WT, SET>(s, c1, cc);
        }
    }
#pragma unroll
    for (int e = 0; e < WT; ++e)
        if (j0 + e >= job.nwords) s[e] = 0;

    // Carry map of this thread's run (computed with carry-in 0).
    uint32_t f;
    {
        uint64_t c = 0;
        uint32_t w0 = 0;
        bool ones = true;
#pragma unroll
        for (int e = 0; e < WT; ++e) {
            uint64_t v = s[e] + c;
            uint32_t o = (uint32_t)v;
            c = v >> 32;
            if (e == 0) w0 = o;
            else ones = ones && (o == 0xFFFFFFFFu);
        }
        // smallest carry-in that overflows word 0: 2^32 - w0 (8 stands for never)
        const uint32_t room = 0u - w0;  // 2^32 - w0 mod 2^32; 0 when w0 == 0
        const uint32_t t = (ones && w0 != 0 && room <= 7u) ? room : 8u;
        f = fn_make((uint32_t)c, t);
    }
    uint32_t agg;
    const uint32_t excl = block_scan_fn<kCarryThreads>(f, waves, &agg);
    __syncthreads();  // block_scan_fn's lds_waves are reused below

    // Look-back: each round, thread t reads the status of block hi - t; the maps
    // up to the nearest published inclusive prefix are composed in order.
    if (threadIdx.x == 0)
        __hip_atomic_store(&status[blk], status_word(epoch, blk == 0 ? kFlagPrefix : kFlagAggregate, agg),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t before = kFnIdentity;  // composition of every earlier block's map
    long long hi = (long long)blk - 1;
    uint32_t spins = 0;
    while (hi >= 0) {
        const long long idx = hi - (long long)threadIdx.x;
        uint32_t flag = (uint32_t)kFlagPrefix, fm = kFnIdentity;  // before block 0: identity
        if (idx >= 0) {
            uint64_t v = __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            flag = ((uint32_t)(v >> 32) == epoch) ? (uint32_t)((v >> 24) & 0xFF) : 0u;
            fm = (uint32_t)(v & 0xFFFFFF);
        }
        if (threadIdx.x == 0) {
            sh_first_prefix = kCarryThreads;
            sh_first_wait = kCarryThreads;
        }
        __syncthreads();
        if (flag == (uint32_t)kFlagPrefix) atomicMin(&sh_first_prefix, threadIdx.x);
        if (flag == 0) atomicMin(&sh_first_wait, threadIdx.x);
        __syncthreads();
        const uint32_t first_prefix = sh_first_prefix, first_wait = sh_first_wait;
        __syncthreads();
        if (first_wait <= first_prefix && first_wait < kCarryThreads) {
            if (++spins >= kLookbackSpinLimit) {  // cannot happen while blocks make progress
                if (threadIdx.x == 0) glob->error = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        const uint32_t g = threadIdx.x <= first_prefix ? fm : kFnIdentity;
        before = fn_compose(before, block_compose_desc(g, waves));
        if (first_prefix < kCarryThreads) break;
        hi -= kCarryThreads;
    }
    if (threadIdx.x == 0 && blk > 0)
        __hip_atomic_store(&status[blk], status_word(epoch, kFlagPrefix, fn_compose(agg, before)),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    // Final words with the thread's true carry-in.
    uint64_t c = fn_apply(excl, fn_apply(before, 0));
    uint32_t w[WT];
#pragma unroll
    for (int e = 0; e < WT; ++e) {
        uint64_t v = s[e] + c;
        w[e] = (uint32_t)v;
        c = v >> 32;
    }
    const global_ptr<uint32_t> out = as_global(job.out);
    if (j0 + WT <= job.nwords) {
#pragma unroll
        for (int q = 0; q < WT / 4; ++q)
            *(global_ptr<u32x4>)(out + j0 + 4 * q) = u32x4{w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]};
    } else {
#pragma unroll
        for (int e = 0; e < WT; ++e)
            if (j0 + e < job.nwords) out[j0 + e] = w[e];
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

// Z[i], i = 2^s + b: w_{2^(s+1)}^bitrev_s(b) = wK^(bitrev_s(b) * 2^(K - 1 - s)),
// wK a primitive 2^K-th root of unity (K: the set's longest transform).
__global__ void k_build_twiddles(uint2* __restrict__ fwd, uint2* __restrict__ inv, uint32_t p,
                                 uint32_t wk, int log2k, uint32_t count) {
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
    uint32_t z = (i == 0) ? 1u : powmod_dev(wk, ef, p);
    uint32_t zi = (i == 0) ? 1u : powmod_dev(wk, ei, p);
    fwd[i] = make_uint2(z, (uint32_t)(((uint64_t)z << 32) / p));
    inv[i] = make_uint2(zi, (uint32_t)(((uint64_t)zi << 32) / p));
}

}  // namespace

hipError_t launch_build_twiddles(uint2* fwd, uint2* inv, int set, int prime, hipStream_t s) {
    const uint32_t p = kPrimeSet[set][prime];
    const int log2k = kSetMaxLog2N[set];
    // wK = g^((p-1)/2^K).
    uint64_t r = 1, x = kGenSet[set][prime];
    uint64_t e = (uint64_t)(p - 1) >> log2k;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    uint32_t count = 1u << log2k;
    hipLaunchKernelGGL(k_build_twiddles, dim3((count + 255) / 256), dim3(256), 0, s, fwd, inv, p,
                       (uint32_t)r, log2k, count);
    return hipGetLastError();
}

hipError_t launch_crt_carry(const JobList<CarryJob>& jobs, uint32_t njobs, uint32_t nblocks, int log2n,
                            const CrtConst& cc, CarryGlobal* glob, hipStream_t s) {
    if (set_for(log2n) == kSetNarrow)
        hipLaunchKernelGGL(k_crt_carry<kSetNarrow>, dim3(nblocks), dim3(kCarryThreads), 0, s, jobs, njobs, log2n,
                           cc, glob);
    else
        hipLaunchKernelGGL(k_crt_carry<kSetWide>, dim3(nblocks), dim3(kCarryThreads), 0, s, jobs, njobs, log2n, cc,
                           glob);
    return hipGetLastError();
}

}  // namespace pintt
