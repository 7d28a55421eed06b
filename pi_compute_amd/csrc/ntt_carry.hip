// ntt_carry.hip - CRT + carry pass and the twiddle-table builder (gfx950).
#include "ntt_device.h"

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
// with a few compares.
//
// k_crt_carry: a workgroup owns kCarryWordsPerBlock words of one group.  It
// reads the residues of every chunk product that reaches them with coalesced
// 16-byte loads, writes their CRT values to LDS (three words per position),
// and each thread then sums the kCarryWordsPerThread consecutive words it
// owns.  The threads' carry maps are scanned across the workgroup; the words
// are written as if the workgroup's carry-in were 0, and its own map goes to
// the group's map array.
// k_carry_fixup: one wave per workgroup of the first kernel composes the maps
// of the group's earlier workgroups into the true carry-in and adds it at the
// workgroup's first word, rippling through words that are all ones, never past
// the workgroup's last word (what spills out is already counted in the next
// workgroup's carry-in).  No workgroup ever waits for another.

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

// Index of the group that owns carry workgroup `blk` (jobs sorted by blk0).
__device__ __forceinline__ uint32_t find_job(const JobList<CarryJob>& jobs, uint32_t njobs, uint32_t blk) {
    uint32_t lo = 0;  // last job with blk0 <= blk
    if (jobs.ninl) {
#pragma unroll
        for (int i = 1; i < kInlineJobs; ++i)
            if ((uint32_t)i < jobs.ninl && jobs.inl[i].blk0 <= blk) lo = i;
    } else {
        uint32_t hi = njobs - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (jobs.dev[mid].blk0 <= blk) lo = mid;
            else hi = mid - 1;
        }
    }
    return lo;
}

constexpr int kCT = kCarryThreads;
constexpr int kWT = kCarryWordsPerThread;
constexpr int kWB = kCarryWordsPerBlock;
constexpr int kGroups = kWT / 4;  // groups of four consecutive positions per thread
static_assert(kWT % 4 == 0 && kWB == kWT * kCT, "each thread loads kWT / 4 groups of four positions");
// LDS slot of word j of the workgroup, j in [-2, kWB): the two words before the
// workgroup first, then one pad word after every kWT so that the threads'
// kWT-word runs start on different banks.
__device__ __forceinline__ uint32_t xslot(int j) {
    const uint32_t i = (uint32_t)(j + 2);
    return i + i / kWT;
}
constexpr int kXWords = (kWB + 2) + (kWB + 2) / kWT + 1;

// grid = carry workgroups of the launch; jobs sorted by blk0.  SET: the prime
// set the residues belong to.
template <int SET>
__global__ void __launch_bounds__(kCarryThreads)
    k_crt_carry(const JobList<CarryJob> jobs, uint32_t njobs, int log2n, CrtConst cc) {
    __shared__ uint32_t X[3][kXWords];
    __shared__ uint32_t waves[kCT / 64];
    __shared__ uint32_t sh_job;
    if (threadIdx.x == 0) sh_job = find_job(jobs, njobs, blockIdx.x);
    __syncthreads();
    const CarryJob job = PINTT_JOB(CarryJob, jobs, sh_job);
    const uint32_t lb = blockIdx.x - job.blk0;  // workgroup within the group
    const long long j0 = (long long)lb * kWB;   // its first word
    const size_t n = (size_t)1 << log2n;
    const int t = threadIdx.x;

    uint64_t s[kWT];
#pragma unroll
    for (int e = 0; e < kWT; ++e) s[e] = 0;
    // Constant chunk indices keep the job (a kernel-argument copy) in registers.
#pragma unroll
    for (uint32_t i = 0; i < (uint32_t)kMaxChunks; ++i) {
        if (i >= job.m) break;
        const long long k0 = j0 - (long long)job.off[i];  // chunk position of word 0
        const long long len = job.len[i];
        if (k0 + kWB <= 0 || k0 - 2 >= len) continue;  // no position of this chunk reaches the words
        const global_ptr<const uint32_t> res = as_global(job.res[i]);
        // Positions k0 + 4t + 4 kCT v + u (v < kGroups, u < 4): all loads first, then the CRTs.
        uint32_t r[kGroups][3][4];
#pragma unroll
        for (int v = 0; v < kGroups; ++v) {
            const long long k = k0 + 4 * t + 4 * kCT * v;
            if (k >= 0 && k + 4 <= len) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const u32x4 x = *(global_ptr<const u32x4>)(res + q * n + k);
                    r[v][q][0] = x.x;
                    r[v][q][1] = x.y;
                    r[v][q][2] = x.z;
                    r[v][q][3] = x.w;
                }
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const bool in = k + u >= 0 && k + u < len;
#pragma unroll
                    for (int q = 0; q < 3; ++q) r[v][q][u] = in ? res[q * n + k + u] : 0u;
                }
            }
        }
#pragma unroll
        for (int v = 0; v < kGroups; ++v)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                uint32_t x0, x1, x2;
                crt_words<SET>(r[v][0][u], r[v][1][u], r[v][2][u], cc, x0, x1, x2);
                const uint32_t sl = xslot(4 * t + 4 * kCT * v + u);
                X[0][sl] = x0;
                X[1][sl] = x1;
                X[2][sl] = x2;
            }
        if (t < 2) {  // the two positions before the workgroup's first word
            const long long k = k0 - 2 + t;
            const bool in = k >= 0 && k < len;
            uint32_t x0, x1, x2;
            crt_words<SET>(in ? res[k] : 0u, in ? res[n + k] : 0u, in ? res[2 * n + k] : 0u, cc, x0, x1, x2);
            const uint32_t sl = xslot(t - 2);
            X[0][sl] = x0;
            X[1][sl] = x1;
            X[2][sl] = x2;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < kWT; ++e) {
            const int jj = kWT * t + e;
            s[e] += (uint64_t)X[0][xslot(jj)] + X[1][xslot(jj - 1)] + X[2][xslot(jj - 2)];
        }
        __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < kWT; ++e)
        if ((uint64_t)(j0 + kWT * t + e) >= job.nwords) s[e] = 0;

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
    uint32_t agg;
    const uint32_t excl = block_scan_fn<kCT>(f, waves, &agg);
    if (t == 0) as_global(job.maps)[lb] = agg;

    // Words with the workgroup's carry-in taken as 0, through LDS so that the
    // stores are coalesced.
    {
        uint64_t c = fn_apply(excl, 0);
#pragma unroll
        for (int e = 0; e < kWT; ++e) {
            const uint64_t v = s[e] + c;
            X[0][xslot(kWT * t + e)] = (uint32_t)v;
            c = v >> 32;
        }
    }
    __syncthreads();
    const global_ptr<uint32_t> out = as_global(job.out);
#pragma unroll
    for (int v = 0; v < kGroups; ++v) {
        const int jj = 4 * t + 4 * kCT * v;
        const uint64_t j = (uint64_t)j0 + jj;
        if (j + 4 <= job.nwords) {
            *(global_ptr<u32x4>)(out + j) =
                u32x4{X[0][xslot(jj)], X[0][xslot(jj + 1)], X[0][xslot(jj + 2)], X[0][xslot(jj + 3)]};
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (j + u < job.nwords) out[j + u] = X[0][xslot(jj + u)];
        }
    }
}

// One wave per carry workgroup: its true carry-in, rippled into its words.
__global__ void __launch_bounds__(64) k_carry_fixup(const JobList<CarryJob> jobs, uint32_t njobs) {
    const uint32_t lane = threadIdx.x;
    const CarryJob job = PINTT_JOB(CarryJob, jobs, find_job(jobs, njobs, blockIdx.x));
    const uint32_t lb = blockIdx.x - job.blk0;
    if (lb == 0) return;
    // Maps of workgroups 0 .. lb-1 in order, 1024 per round: 16 consecutive
    // maps per lane, all loaded before any is composed.
    const global_ptr<const uint32_t> maps = as_global(job.maps);
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

hipError_t launch_crt_carry(const JobList<CarryJob>& jobs, uint32_t njobs, uint32_t nblocks, int log2n,
                            const CrtConst& cc, hipStream_t s) {
    if (set_for(log2n) == kSetNarrow)
        hipLaunchKernelGGL(k_crt_carry<kSetNarrow>, dim3(nblocks), dim3(kCarryThreads), 0, s, jobs, njobs, log2n,
                           cc);
    else
        hipLaunchKernelGGL(k_crt_carry<kSetWide>, dim3(nblocks), dim3(kCarryThreads), 0, s, jobs, njobs, log2n, cc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_carry_fixup, dim3(nblocks), dim3(64), 0, s, jobs, njobs);
    return hipGetLastError();
}

}  // namespace pintt
