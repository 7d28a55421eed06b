// bigint_add.hip - batched addition / subtraction of limb arrays on the device
// (gfx950): the signed add R = Qr*Rl + Pl*Rr of binary-splitting merges, kept
// on the device between tree levels.
//
// With w = max(xn, yn): out = |x| + |y| (w + 1 limbs) or |x| - |y| (w limbs;
// when |x| < |y| it is |x| - |y| + 2^(64 w) and the job is flagged negative;
// lengths may include zero top limbs).  Subtraction is x + ~y + 1 over the width.  A run of limbs maps
// an incoming carry c in {0, 1} to its carry-out: two bits (f(0), f(1)).
// Three launches per batch:
//   k_add_maps   each block of 4096 limbs: per-thread maps, block composition
//   k_add_scan   one workgroup per job: carry-in of every block of the job
//   k_add_apply  each block: in-block scan from its carry-in, limbs written
#include "ntt_device.h"
#include "bigint_add.h"

namespace pintt {
namespace {

constexpr int kAT = kAddThreads;
constexpr int kAL = kAddLimbsPerThread;

// map: bit 0 = carry-out for carry-in 0, bit 1 = for carry-in 1; later o earlier
__device__ __forceinline__ uint32_t cm_compose(uint32_t later, uint32_t earlier) {
    const uint32_t e0 = earlier & 1u, e1 = (earlier >> 1) & 1u;
    return ((later >> e0) & 1u) | (((later >> e1) & 1u) << 1);
}
constexpr uint32_t kCmIdentity = 2u;  // f(0) = 0, f(1) = 1

// Last job with blk0 <= blk (jobs sorted by blk0), by every lane of a wave.
__device__ __forceinline__ uint32_t find_job(const AddJob* __restrict__ jobs, uint32_t njobs, uint32_t blk) {
    return wave_search(njobs, blk, [jobs](uint32_t i) { return jobs[i].blk0; });
}

__device__ __forceinline__ uint64_t width_of(const AddJob& j) { return j.xn > j.yn ? j.xn : j.yn; }

// The thread's limbs i0 .. i0+kAL-1 of x and of y (complemented when
// subtracting), zero-extended; positions at or past the width are zero.
__device__ __forceinline__ void load_limbs(const AddJob& j, uint64_t i0, uint64_t w, uint64_t* xs, uint64_t* ys) {
    const global_ptr<const uint64_t> x = as_global(j.x), y = as_global(j.y);
#pragma unroll
    for (int e = 0; e < kAL; ++e) {
        const uint64_t i = i0 + e;
        const uint64_t xv = i < j.xn ? x[i] : 0ull;
        uint64_t yv = i < j.yn ? y[i] : 0ull;
        if (j.sub) yv = ~yv;
        xs[e] = i < w ? xv : 0ull;
        ys[e] = i < w ? yv : 0ull;
    }
}

// Carry map of the thread's run (limbs past the width do not carry).
__device__ __forceinline__ uint32_t run_map(const uint64_t* xs, const uint64_t* ys, uint64_t i0, uint64_t w) {
    uint32_t c0 = 0, c1 = 1;
#pragma unroll
    for (int e = 0; e < kAL; ++e) {
        if (i0 + e >= w) break;
        const uint64_t s = xs[e] + ys[e];
        const uint32_t cs = s < xs[e] ? 1u : 0u;
        c0 = cs | ((s + c0 < s) ? 1u : 0u);
        c1 = cs | ((s + c1 < s) ? 1u : 0u);
    }
    return c0 | (c1 << 1);
}

// Exclusive scan of maps over the block in thread order; *total = composition of all.
__device__ __forceinline__ uint32_t block_scan_cm(uint32_t f, uint32_t* lds, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = f;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off, 64);
        if (lane >= off) incl = cm_compose(incl, o);
    }
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    uint32_t before = kCmIdentity;
    for (int w = 0; w < wave; ++w) before = cm_compose(lds[w], before);
    uint32_t all = kCmIdentity;
    for (int w = 0; w < kAT / 64; ++w) all = cm_compose(lds[w], all);
    uint32_t excl = __shfl_up(incl, 1, 64);
    if (lane == 0) excl = kCmIdentity;
    __syncthreads();
    *total = all;
    return cm_compose(excl, before);
}

__global__ void __launch_bounds__(kAT) k_add_maps(const AddJob* __restrict__ jobs, uint32_t njobs,
                                                  uint8_t* __restrict__ maps) {
    __shared__ uint32_t lds[kAT / 64];
    const uint32_t jz = find_job(jobs, njobs, blockIdx.x);
    const AddJob j = jobs[jz];
    const uint64_t w = width_of(j);
    const uint64_t i0 = (uint64_t)(blockIdx.x - j.blk0) * kAddLimbsPerBlock + (uint64_t)threadIdx.x * kAL;
    uint64_t xs[kAL], ys[kAL];
    load_limbs(j, i0, w, xs, ys);
    uint32_t total;
    (void)block_scan_cm(run_map(xs, ys, i0, w), lds, &total);
    if (threadIdx.x == 0) maps[blockIdx.x] = (uint8_t)total;
}

// One workgroup per job: carry-in of each of its blocks, carry-out of the job.
__global__ void __launch_bounds__(kAT) k_add_scan(const AddJob* __restrict__ jobs, const uint8_t* __restrict__ maps,
                                                  uint8_t* __restrict__ carry_in, uint32_t* __restrict__ error,
                                                  uint32_t* __restrict__ status) {
    __shared__ uint32_t lds[kAT / 64];
    const AddJob j = jobs[blockIdx.x];
    uint32_t c = j.sub ? 1u : 0u;
    for (uint32_t b0 = 0; b0 < j.nblk; b0 += kAT) {
        const uint32_t b = b0 + threadIdx.x;
        const uint32_t f = b < j.nblk ? maps[j.blk0 + b] : kCmIdentity;
        uint32_t total;
        const uint32_t excl = block_scan_cm(f, lds, &total);
        if (b < j.nblk) carry_in[j.blk0 + b] = (uint8_t)((excl >> c) & 1u);
        c = (total >> c) & 1u;
    }
    if (threadIdx.x == 0) {
        const uint32_t negative = (j.sub && c != 1u) ? 1u : 0u;  // |x| < |y|
        if (negative) atomicOr(error, 1u);
        if (!j.sub) as_global(j.out)[width_of(j)] = c;
        status[blockIdx.x] = negative;
    }
}

__global__ void __launch_bounds__(kAT) k_add_apply(const AddJob* __restrict__ jobs, uint32_t njobs,
                                                   const uint8_t* __restrict__ carry_in) {
    __shared__ uint32_t lds[kAT / 64];
    const uint32_t jz = find_job(jobs, njobs, blockIdx.x);
    const AddJob j = jobs[jz];
    const uint64_t w = width_of(j);
    const uint64_t i0 = (uint64_t)(blockIdx.x - j.blk0) * kAddLimbsPerBlock + (uint64_t)threadIdx.x * kAL;
    uint64_t xs[kAL], ys[kAL];
    load_limbs(j, i0, w, xs, ys);
    uint32_t total;
    const uint32_t excl = block_scan_cm(run_map(xs, ys, i0, w), lds, &total);
    uint32_t c = (excl >> carry_in[blockIdx.x]) & 1u;
    const global_ptr<uint64_t> out = as_global(j.out);
#pragma unroll
    for (int e = 0; e < kAL; ++e) {
        const uint64_t i = i0 + e;
        if (i >= w) break;
        const uint64_t s = xs[e] + ys[e];
        const uint64_t r = s + c;
        c = ((s < xs[e]) || (r < s)) ? 1u : 0u;
        out[i] = r;
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_add_batch(const AddJob* jobs, uint32_t njobs, uint32_t nblocks, uint8_t* maps, uint8_t* carry_in,
                            uint32_t* error, uint32_t* status, hipStream_t s) {
    hipLaunchKernelGGL(k_add_maps, dim3(nblocks), dim3(kAT), 0, s, jobs, njobs, maps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_add_scan, dim3(njobs), dim3(kAT), 0, s, jobs, maps, carry_in, error, status);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_add_apply, dim3(nblocks), dim3(kAT), 0, s, jobs, njobs, carry_in);
    return hipGetLastError();
}

}  // namespace pintt
