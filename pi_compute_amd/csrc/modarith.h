// modarith.h - 32-bit modular arithmetic for the three NTT primes (gfx950).
//
// All three primes are below 2^31, so 2p < 2^32 and every sum of two reduced
// residues fits a 32-bit lane.  Twiddle products use Shoup's precomputed
// quotient (w' = floor(w * 2^32 / p)): one high and one low 32-bit product plus
// a fused q*p subtraction, result in [0, 2p) for any 32-bit x, then one
// conditional subtract.  The pointwise product of two variable residues uses
// Montgomery reduction with R = 2^32 (the stray 2^-32 is folded into the CRT
// constants).  Reductions are written as min(x, x - p) on unsigned values:
// x - p wraps above x exactly when x < p.
//
// Replaces the reference's 64-bit '%' arithmetic (src/gpu/int_ntt.cu:47-51).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pintt {

constexpr int kPrimes = 3;
constexpr int kMaxLog2N = 25;  // minimum 2-adic order of p - 1 over the set

// p = c * 2^k + 1.
constexpr uint32_t kP[kPrimes] = {2013265921u /* 15*2^27+1 */, 1811939329u /* 27*2^26+1 */,
                                  2113929217u /* 63*2^25+1 */};
constexpr uint32_t kGen[kPrimes] = {31u, 13u, 5u};

__host__ __device__ constexpr uint32_t neg_inv_mod_2_32(uint32_t p) {
    // Newton iteration for p^-1 mod 2^32, negated.
    uint32_t x = p;  // correct to 3 bits for odd p
    for (int i = 0; i < 5; ++i) x *= 2u - p * x;
    return 0u - x;
}

__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t s = a + b;
    return min(s, s - p);
}

__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t d = a - b;
    return min(d, d + p);
}

// x * w mod p, for any x < 2^32, with wq = floor(w * 2^32 / p).  Result < p.
__device__ __forceinline__ uint32_t mul_shoup(uint32_t x, uint32_t w, uint32_t wq, uint32_t p) {
    uint32_t q = __umulhi(x, wq);
    uint32_t r = x * w - q * p;
    return min(r, r - p);
}

// a * b * 2^-32 mod p for a, b < p (Montgomery REDC, pinv = -p^-1 mod 2^32).
__device__ __forceinline__ uint32_t mul_mont(uint32_t a, uint32_t b, uint32_t p, uint32_t pinv) {
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * pinv;
    uint32_t u = (uint32_t)((t + (uint64_t)m * p) >> 32);
    return min(u, u - p);
}

// Reduce an arbitrary 32-bit coefficient (< 3p for every prime of the set).
__device__ __forceinline__ uint32_t reduce_u32(uint32_t c, uint32_t p) {
    c = min(c, c - 2u * p);
    return min(c, c - p);
}

// Cooley-Tukey butterfly: (a, b) -> (a + z*b, a - z*b).
__device__ __forceinline__ void ct_bfly(uint32_t& a, uint32_t& b, uint2 z, uint32_t p) {
    uint32_t t = mul_shoup(b, z.x, z.y, p);
    uint32_t u = a;
    a = add_mod(u, t, p);
    b = sub_mod(u, t, p);
}

// Gentleman-Sande butterfly: (a, b) -> (a + b, (a - b) * zinv).  a - b + p lies
// in [1, 2p), inside the Shoup input range, so it needs no reduction first.
__device__ __forceinline__ void gs_bfly(uint32_t& a, uint32_t& b, uint2 zi, uint32_t p) {
    uint32_t u = a, v = b;
    a = add_mod(u, v, p);
    b = mul_shoup(u + p - v, zi.x, zi.y, p);
}

}  // namespace pintt
