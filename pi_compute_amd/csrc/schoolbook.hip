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
    if (threadIdx.x == 0) {
        // Column j plus the carry from below (under 2^73), limb j out.
        uint64_t c0 = 0, c1 = 0;
        for (uint32_t j = 0; j < nc; ++j) {
            const uint64_t t0 = S0[j] + c0;
            const uint64_t k0 = t0 < c0 ? 1u : 0u;
            const uint64_t t1 = S1[j] + c1;
            const uint64_t k1 = t1 < c1 ? 1u : 0u;
            const uint64_t u1 = t1 + k0;
            const uint64_t k2 = u1 < k0 ? 1u : 0u;
            S0[j] = t0;
            c0 = u1;
            c1 = S2[j] + k1 + k2;
        }
        S0[nc] = c0;  // the product is below 2^(64 (an + bn)): nothing above
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
