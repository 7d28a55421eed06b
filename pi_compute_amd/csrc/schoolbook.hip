// schoolbook.hip - small products of a batch by the schoolbook method
// (gfx950): at a few hundred limbs a product costs far less as n^2 64-bit
// multiply-adds than as a three-prime transform (whose passes, launches and
// padding do not shrink with the operands).  One workgroup per product: the
// operands in LDS, each thread sums whole columns of the product (128-bit
// partial products into three-word accumulators), then one lane carries the
// columns from the bottom and the workgroup writes the limbs out.
#include "ntt_device.h"

namespace pintt {
namespace {

constexpr int kST = kSchoolThreads;
constexpr int kSM = kSchoolMaxLimbs;

__global__ void __launch_bounds__(kST) k_schoolbook(const JobList<SchoolJob> jobs) {
    __shared__ uint64_t A[kSM], B[kSM];
    __shared__ uint64_t S0[2 * kSM], S1[2 * kSM], S2[2 * kSM];
    const SchoolJob job = PINTT_JOB(SchoolJob, jobs, blockIdx.x);
    const uint32_t an = job.an, bn = job.bn;
    const global_ptr<const uint64_t> a = as_global(job.a), b = as_global(job.b);
    for (uint32_t i = threadIdx.x; i < an; i += kST) A[i] = a[i];
    for (uint32_t i = threadIdx.x; i < bn; i += kST) B[i] = b[i];
    __syncthreads();
    const uint32_t nc = an + bn - 1;  // columns 0 .. an + bn - 2
    for (uint32_t j = threadIdx.x; j < nc; j += kST) {
        const uint32_t lo = j + 1 > bn ? j + 1 - bn : 0;
        const uint32_t hi = j < an - 1 ? j : an - 1;
        uint64_t s0 = 0, s1 = 0, s2 = 0;
        for (uint32_t i = lo; i <= hi; ++i) {
            const uint64_t x = A[i], y = B[j - i];
            const uint64_t pl = x * y, ph = __umul64hi(x, y);
            s0 += pl;
            const uint64_t c0 = s0 < pl ? 1u : 0u;
            const uint64_t t = s1 + ph;
            const uint64_t c1 = t < s1 ? 1u : 0u;
            s1 = t + c0;
            s2 += c1 + (s1 < c0 ? 1u : 0u);
        }
        S0[j] = s0;
        S1[j] = s1;
        S2[j] = s2;
    }
    __syncthreads();
    // Carries by the first wave: lane l takes K >= 2 consecutive columns and
    // carries them from zero; then the lanes' carry-outs (< 2^74) run through
    // lane 0 in order (a carry-in overflows a lane of two or more limbs only if
    // its limbs above the second are all ones), and each lane adds its own
    // carry-in.
    __shared__ uint64_t LC0[64], LC1[64], LR0[64], LR1[64], CI0[65], CI1[65];
    __shared__ uint32_t LN[64], LONES[64];
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x;
        const uint32_t K = nc > 128 ? (nc + 63) / 64 : 2;
        const uint32_t j0 = lane * K < nc ? lane * K : nc, j1 = j0 + K < nc ? j0 + K : nc;
        uint64_t c0 = 0, c1 = 0;
        uint32_t ones = 1;
        for (uint32_t j = j0; j < j1; ++j) {
            const uint64_t t0 = S0[j] + c0;
            const uint64_t k0 = t0 < c0 ? 1u : 0u;
            const uint64_t t1 = S1[j] + c1;
            const uint64_t k1 = t1 < c1 ? 1u : 0u;
            const uint64_t u1 = t1 + k0;
            const uint64_t k2 = u1 < k0 ? 1u : 0u;
            S0[j] = t0;
            c0 = u1;
            c1 = S2[j] + k1 + k2;
            if (j >= j0 + 2 && t0 != ~0ull) ones = 0;
        }
        LC0[lane] = c0;
        LC1[lane] = c1;
        LR0[lane] = j1 > j0 ? S0[j0] : 0;
        LR1[lane] = j1 > j0 + 1 ? S0[j0 + 1] : 0;
        LN[lane] = j1 - j0;
        LONES[lane] = ones;
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
            uint64_t i0 = 0, i1 = 0;  // carry into lane l
            for (uint32_t l = 0; l < 64; ++l) {
                CI0[l] = i0;
                CI1[l] = i1;
                const uint32_t n = LN[l];
                if (n == 0) continue;  // nothing here: the carry passes on
                const uint64_t s0 = LR0[l] + i0;
                const uint64_t k = s0 < i0 ? 1u : 0u;
                uint64_t add;  // what the carry-in adds to the lane's carry-out
                if (n == 1) {
                    add = i1 + k;  // below 2^11
                } else {
                    const uint64_t s1 = LR1[l] + i1;
                    const uint64_t kk = (s1 < i1 ? 1u : 0u) + ((s1 + k) < s1 ? 1u : 0u);
                    add = (LONES[l] && kk) ? 1u : 0u;
                }
                i0 = LC0[l] + add;
                i1 = LC1[l] + (i0 < add ? 1u : 0u);
            }
            CI0[64] = i0;
            CI1[64] = i1;
        }
        __builtin_amdgcn_wave_barrier();
        uint64_t a0 = CI0[lane], a1 = CI1[lane];  // the lane's carry-in, added to its limbs
        for (uint32_t j = j0; j < j1 && (a0 | a1); ++j) {
            const uint64_t t = S0[j] + a0;
            const uint64_t k = t < a0 ? 1u : 0u;
            S0[j] = t;
            a0 = a1 + k;
            a1 = 0;
        }
        if (lane == 0) S0[nc] = CI0[64];  // the product is below 2^(64 (an + bn)): nothing above
    }
    __syncthreads();
    const global_ptr<uint64_t> out = as_global(job.out);
    for (uint32_t j = threadIdx.x; j <= nc; j += kST) out[j] = S0[j];
}

}  // namespace

hipError_t launch_schoolbook(const JobList<SchoolJob>& jobs, uint32_t njobs, hipStream_t s) {
    hipLaunchKernelGGL(k_schoolbook, dim3(njobs), dim3(kST), 0, s, jobs);
    return hipGetLastError();
}

}  // namespace pintt
