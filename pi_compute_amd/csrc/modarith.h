// modarith.h - 32-bit modular arithmetic for the NTT primes (gfx950).
//
// Two sets of three primes.  The wide set (primes just below 2^31, 2-adic
// order >= 25) serves transforms up to 2^25 points; every residue is kept fully
// reduced, in [0, p).  The narrow set (primes below 2^29, 2-adic order >= 22)
// serves transforms up to 2^22 points with Harvey's lazy butterflies: 8p still
// fits a 32-bit lane, so forward values live in [0, 8p) -- a forward butterfly
// reduces its first input only when it may exceed 6p, every other layer -- and
// inverse values in [0, 4p) -- an inverse butterfly reduces its sum only when
// an input may exceed 2p.  Their product, 2^85.93, exceeds every
// coefficient of a product the set serves: a coefficient sums at most
// min(la, lb) <= 2^21 terms below 2^64 when la + lb - 1 <= 2^22.
// Twiddle products use Shoup's precomputed quotient (w' = floor(w * 2^32 / p)):
// one high and two low 32-bit products give x * w - q * p in [0, 2p) for any
// 32-bit x.  The pointwise product of two variable residues uses Montgomery
// reduction with R = 2^32 (the stray 2^-32 is folded into the CRT constants).
// Reductions are written as min(x, x - c) on unsigned values: x - c wraps
// above x exactly when x < c.
//
// Replaces the reference's 64-bit '%' arithmetic (src/gpu/int_ntt.cu:47-51).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pintt {

constexpr int kPrimes = 3;
constexpr int kMaxLog2N = 25;        // longest transform (wide set)
constexpr int kNarrowMaxLog2N = 22;  // longest transform of the narrow set

// Prime sets: p = c * 2^k + 1.
constexpr int kSetWide = 0, kSetNarrow = 1, kSets = 2;
constexpr uint32_t kPrimeSet[kSets][kPrimes] = {
    {2013265921u /* 15*2^27+1 */, 1811939329u /* 27*2^26+1 */, 2113929217u /* 63*2^25+1 */},
    {469762049u /* 7*2^26+1 */, 415236097u /* 99*2^22+1 */, 377487361u /* 45*2^23+1 */}};
constexpr uint32_t kGenSet[kSets][kPrimes] = {{31u, 13u, 5u}, {3u, 5u, 7u}};
constexpr int kSetMaxLog2N[kSets] = {kMaxLog2N, kNarrowMaxLog2N};

// The set that serves a transform of 2^log2n points.
__host__ __device__ constexpr int set_for(int log2n) { return log2n <= kNarrowMaxLog2N ? kSetNarrow : kSetWide; }

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

// x * f * 2^-32 mod p for any 32-bit x, f < p and fp = f * p^-1 mod 2^32: a
// Montgomery product with a multiplier fixed for many x (one multiplication
// gives its fp, where Shoup's quotient would need a division).  m = x fp makes
// the low halves of x f and m p equal, so (x f - m p) / 2^32 is the difference
// of the high halves, in (-p, p); the result, plus p, lies in (0, 2p).
__device__ __forceinline__ uint32_t mul_fixed(uint32_t x, uint32_t f, uint32_t fp, uint32_t p) {
    const uint32_t m = x * fp;
    return __umulhi(x, f) - __umulhi(m, p) + p;
}

// Butterflies and the per-set steps around them.  The twiddle z is (w, w').
template <int SET>
struct Arith;

// Wide set: every value in [0, p).
template <>
struct Arith<kSetWide> {
    // A 32-bit coefficient (< 3p for every wide prime) to [0, p).
    __device__ static __forceinline__ uint32_t load_coeff(uint32_t c, uint32_t p) {
        c = min(c, c - 2u * p);
        return min(c, c - p);
    }
    // Cooley-Tukey: (a, b) -> (a + z*b, a - z*b).
    __device__ static __forceinline__ void ct(uint32_t& a, uint32_t& b, uint2 z, uint32_t p) {
        uint32_t t = mul_shoup(b, z.x, z.y, p);
        uint32_t u = a;
        a = add_mod(u, t, p);
        b = sub_mod(u, t, p);
    }
    // Forward butterfly with a bound check the wide set never skips.
    __device__ static __forceinline__ void ct_sel(uint32_t& a, uint32_t& b, uint2 z, uint32_t p, bool) {
        ct(a, b, z, p);
    }
    // Gentleman-Sande: (a, b) -> (a + b, (a - b) * zinv).  a - b + p lies in
    // [1, 2p), inside the Shoup input range, so it needs no reduction first.
    __device__ static __forceinline__ void gs(uint32_t& a, uint32_t& b, uint2 zi, uint32_t p) {
        uint32_t u = a, v = b;
        a = add_mod(u, v, p);
        b = mul_shoup(u + p - v, zi.x, zi.y, p);
    }
    __device__ static __forceinline__ void gs_sel(uint32_t& a, uint32_t& b, uint2 zi, uint32_t p, bool) {
        gs(a, b, zi, p);
    }
    // Forward outputs (in [0, p)) to the inverse input range, multiplied.
    __device__ static __forceinline__ uint32_t pointwise(uint32_t x, uint32_t y, uint32_t p, uint32_t pinv) {
        return mul_mont(x, y, p, pinv);
    }
    __device__ static __forceinline__ uint32_t shared_operand(uint32_t y, uint32_t) { return y; }
    __device__ static __forceinline__ uint32_t pointwise_shared(uint32_t x, uint32_t ys, uint32_t p,
                                                                uint32_t pinv) {
        return mul_mont(x, ys, p, pinv);
    }
    // x * f 2^-32 (mul_fixed), fully reduced.
    __device__ static __forceinline__ uint32_t scale(uint32_t x, uint32_t f, uint32_t fp, uint32_t p) {
        const uint32_t u = mul_fixed(x, f, fp, p);
        return min(u, u - p);
    }
};

// Narrow set (p < 2^29): forward values in [0, 8p), inverse values in [0, 4p).
template <>
struct Arith<kSetNarrow> {
    // Any 32-bit coefficient to [0, 8p): 2^32 < 16p, so one step suffices.
    __device__ static __forceinline__ uint32_t load_coeff(uint32_t c, uint32_t p) { return min(c, c - 8u * p); }
    // a < 8p, b any 32-bit value.  x = a reduced to [0, 4p); r = z*b - q*p in
    // [0, 2p) is kept negated (tn = -r mod 2^32) so that both outputs take one
    // add: x + r in [0, 6p) and x - r + 2p in (0, 6p).
    __device__ static __forceinline__ void ct(uint32_t& a, uint32_t& b, uint2 z, uint32_t p) {
        const uint32_t x = min(a, a - 4u * p);
        const uint32_t q = __umulhi(b, z.y);
        const uint32_t tn = q * p - b * z.x;
        a = x - tn;
        b = x + tn + 2u * p;
    }
    // a < 6p, no reduction: outputs a + r in [0, 8p) and a - r + 2p in (0, 8p).
    __device__ static __forceinline__ void ct_lazy(uint32_t& a, uint32_t& b, uint2 z, uint32_t p) {
        const uint32_t x = a;
        const uint32_t q = __umulhi(b, z.y);
        const uint32_t tn = q * p - b * z.x;
        a = x - tn;
        b = x + tn + 2u * p;
    }
    // lazy: the caller knows a < 6p (a compile-time constant after unrolling).
    __device__ static __forceinline__ void ct_sel(uint32_t& a, uint32_t& b, uint2 z, uint32_t p, bool lazy) {
        if (lazy) ct_lazy(a, b, z, p);
        else ct(a, b, z, p);
    }
    // a, b in [0, 4p): a + b in [0, 8p) reduced to [0, 4p); (a - b + 4p) in
    // (0, 8p) times zinv by Shoup, in [0, 2p).
    __device__ static __forceinline__ void gs(uint32_t& a, uint32_t& b, uint2 zi, uint32_t p) {
        const uint32_t p4 = 4u * p;
        const uint32_t u = a, v = b;
        const uint32_t s = u + v;
        a = min(s, s - p4);
        const uint32_t d = u - v + p4;
        const uint32_t q = __umulhi(d, zi.y);
        b = d * zi.x - q * p;
    }
    // a, b in [0, 2p): a + b in [0, 4p) kept as it is; b as in gs.
    __device__ static __forceinline__ void gs_lazy(uint32_t& a, uint32_t& b, uint2 zi, uint32_t p) {
        const uint32_t u = a, v = b;
        a = u + v;
        const uint32_t d = u - v + 4u * p;
        const uint32_t q = __umulhi(d, zi.y);
        b = d * zi.x - q * p;
    }
    // lazy: the caller knows both inputs are below 2p (a compile-time constant
    // after unrolling).
    __device__ static __forceinline__ void gs_sel(uint32_t& a, uint32_t& b, uint2 zi, uint32_t p, bool lazy) {
        if (lazy) gs_lazy(a, b, zi, p);
        else gs(a, b, zi, p);
    }
    // x, y in [0, 8p) reduced to [0, 2p); their product is below 4p^2 < 2^32 p,
    // so the Montgomery result (t + m p) / 2^32 lies in [0, 2p) unreduced.
    __device__ static __forceinline__ uint32_t pointwise(uint32_t x, uint32_t y, uint32_t p, uint32_t pinv) {
        const uint32_t p2 = 2u * p, p4 = 4u * p;
        x = min(x, x - p4);
        x = min(x, x - p2);
        y = min(y, y - p4);
        y = min(y, y - p2);
        const uint64_t t = (uint64_t)x * y;
        const uint32_t m = (uint32_t)t * pinv;
        return (uint32_t)((t + (uint64_t)m * p) >> 32);
    }
    // The operand of pointwise products that one factor shares (b in a*b):
    // y in [0, 8p) reduced to [0, 2p) once.  Then x needs only [0, 4p): x y <
    // 8p^2 < 2^32 p keeps the Montgomery result in [0, 2p).
    __device__ static __forceinline__ uint32_t shared_operand(uint32_t y, uint32_t p) {
        y = min(y, y - 4u * p);
        return min(y, y - 2u * p);
    }
    __device__ static __forceinline__ uint32_t pointwise_shared(uint32_t x, uint32_t ys, uint32_t p,
                                                                uint32_t pinv) {
        x = min(x, x - 4u * p);
        const uint64_t t = (uint64_t)x * ys;
        const uint32_t m = (uint32_t)t * pinv;
        return (uint32_t)((t + (uint64_t)m * p) >> 32);
    }
    // x * f 2^-32 (mul_fixed) in (0, 2p): inside both the forward [0, 8p) and
    // the inverse [0, 4p) input ranges.
    __device__ static __forceinline__ uint32_t scale(uint32_t x, uint32_t f, uint32_t fp, uint32_t p) {
        return mul_fixed(x, f, fp, p);
    }
};

// Bound (in units of p) of forward values before forward layer l of a run that
// starts below B0 p, when every layer whose first input may exceed 6p reduces it
// to 4p (Arith<kSetNarrow>::ct) and the others skip the reduction (ct_lazy):
// B -> B + 2 while B <= 6, else 6.
__host__ __device__ constexpr int lazy_bound(int B0, int l) {
    int b = B0;
    for (int i = 0; i < l; ++i) b = b <= 6 ? b + 2 : 6;
    return b;
}
__host__ __device__ constexpr bool lazy_ok(int B0, int l) { return lazy_bound(B0, l) <= 6; }
static_assert(lazy_bound(2, 3) == 8 && lazy_bound(8, 1) == 6 && lazy_bound(8, 2) == 8, "lazy schedule");

}  // namespace pintt
