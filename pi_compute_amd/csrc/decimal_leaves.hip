// decimal_leaves.hip - the last step of the device-resident decimal
// conversion (gfx950): numbers of at most kLeafMaxDigits decimal digits, one
// per thread, written as ASCII digits (the role of the host leaves, mpz_get_str
// at the bottom of the reference's BaseConverter, src/io/base_converter.cpp).
//
// A thread keeps its number in LDS (limb i of lane t at word i * 64 + t, so a
// wave's accesses fall in distinct banks) and divides it by 10^19 over and over,
// top limb first: each pass turns the remainder into the next 19 digits from
// the right.  Division of a two-limb value by the invariant 10^19 (a normalised
// divisor: 2^63 < 10^19 < 2^64) uses a precomputed reciprocal (Moller and
// Granlund, "Improved division by invariant integers", IEEE Trans. Computers
// 60(2), 2011, algorithm 4).
#include "ntt_device.h"
#include "bigint_add.h"

namespace pintt {
namespace {

constexpr uint64_t kD19 = 10000000000000000000ull;
// floor((2^128 - 1) / 10^19) - 2^64
constexpr uint64_t kD19Inv =
    (uint64_t)((~(unsigned __int128)0) / kD19 - ((unsigned __int128)1 << 64));

// q = floor((u1 2^64 + u0) / 10^19), *r the remainder; requires u1 < 10^19.
__device__ __forceinline__ uint64_t div_d19(uint64_t u1, uint64_t u0, uint64_t* r) {
    const uint64_t p0 = kD19Inv * u1;
    const uint64_t p1 = __umul64hi(kD19Inv, u1);
    const uint64_t q0 = p0 + u0;
    uint64_t q1 = p1 + u1 + 1 + (q0 < p0 ? 1u : 0u);
    uint64_t rr = u0 - q1 * kD19;
    if (rr > q0) {
        q1 -= 1;
        rr += kD19;
    }
    if (rr >= kD19) {
        q1 += 1;
        rr -= kD19;
    }
    *r = rr;
    return q1;
}

// The 19 digits of r (< 10^19, leading zeros kept) at o[end - 19 .. end),
// positions below 0 skipped.
__device__ __forceinline__ void put19(uint8_t* o, int64_t end, uint64_t r) {
    const uint64_t hi = r / 1000000000ull;
    uint32_t c = (uint32_t)(r - hi * 1000000000ull);
    uint32_t b = (uint32_t)(hi % 1000000000ull);
    const uint32_t a = (uint32_t)(hi / 1000000000ull);
    for (int k = 0; k < 9; ++k) {
        const int64_t p = end - 1 - k;
        if (p >= 0) o[p] = (uint8_t)('0' + c % 10u);
        c /= 10u;
    }
    for (int k = 0; k < 9; ++k) {
        const int64_t p = end - 10 - k;
        if (p >= 0) o[p] = (uint8_t)('0' + b % 10u);
        b /= 10u;
    }
    if (end - 19 >= 0) o[end - 19] = (uint8_t)('0' + a);
}

__global__ void __launch_bounds__(kLeafThreads) k_decimal_leaves(const LeafJob* __restrict__ jobs, uint32_t njobs,
                                                                 uint8_t* __restrict__ out) {
    __shared__ uint64_t lds[kLeafMaxLimbs * kLeafThreads];
    const uint32_t jz = blockIdx.x * kLeafThreads + threadIdx.x;
    if (jz >= njobs) return;  // no barriers below: each thread works alone
    const LeafJob job = jobs[jz];
    uint64_t* v = lds + threadIdx.x;
    const global_ptr<const uint64_t> x = as_global(job.x);
    uint32_t n = (uint32_t)job.xn;
    for (uint32_t i = 0; i < n; ++i) v[i * kLeafThreads] = x[i];
    while (n > 0 && v[(n - 1) * kLeafThreads] == 0) --n;
    uint8_t* o = out + job.offset;
    for (int64_t end = (int64_t)job.width; end > 0; end -= 19) {
        uint64_t r = 0;
        for (int32_t i = (int32_t)n - 1; i >= 0; --i) v[i * kLeafThreads] = div_d19(r, v[i * kLeafThreads], &r);
        if (n > 0 && v[(n - 1) * kLeafThreads] == 0) --n;
        put19(o, end, r);
    }
}

}  // namespace

hipError_t launch_decimal_leaves(const LeafJob* jobs, uint32_t njobs, uint8_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_decimal_leaves, dim3((njobs + kLeafThreads - 1) / kLeafThreads), dim3(kLeafThreads), 0, s,
                       jobs, njobs, out);
    return hipGetLastError();
}

}  // namespace pintt
