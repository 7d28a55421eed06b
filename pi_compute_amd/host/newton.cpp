// newton.cpp - see newton.h.
//
// One step from precision c to c' <= 2c - 16 (guard g = 32 bits):
//   reciprocal:  E = 2^K - D'*Y   (D' = top c'+g bits of d, K = bitlen(D') + c)
//                Y' = Y*2^(c'-c) + Y*E*2^(c'-c) / 2^K
//   inv. sqrt:   E = 2^(2c) - a*Y^2
//                Y' = Y*2^(c'-c) + Y*E*2^(c'-c) / 2^(2c+1)
// E is cut to its top c'-c+g bits before the product.  With |e| = |E|/2^K
// <= 2^-(c-3), the new relative error is about e^2 <= 2^-(c'+10) plus the
// cuts (2^-(c'+g-1) each) plus one unit from the floor.
#include "newton.h"

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace pi {
namespace {

constexpr size_t kExactBits = 256;  // below this, exact small-number division / square root
constexpr size_t kGuard = 32;

std::vector<size_t> schedule(size_t b) {
    std::vector<size_t> s;  // decreasing; the last entry is the first step's target
    for (size_t c = b; c > kExactBits; c = c / 2 + 16) s.push_back(c);
    return s;
}

// t = floor(x / 2^s) for signed x (s >= 0).
void fshift(mpz_t t, mpz_srcptr x, long s) {
    if (s >= 0)
        mpz_fdiv_q_2exp(t, x, (mp_bitcnt_t)s);
    else
        mpz_mul_2exp(t, x, (mp_bitcnt_t)(-s));
}

}  // namespace

// When |2^K - p| < 2^(L-1) and L < K, the bits of p from L-1 up to K-1 are
// all ones (p < 2^K) or all zeros below a top bit K (p >= 2^K), which a scan
// confirms, and e follows from the low L bits of p.
void pow2_minus(mpz_t e, mpz_srcptr p, size_t K, size_t L) {
    if (L >= 1 && L < K && mpz_sgn(p) > 0) {
        const size_t n = mpz_sizeinbase(p, 2);
        if (n == K + 1 && mpz_scan1(p, L - 1) == K) {  // p = 2^K + (p mod 2^(L-1))
            mpz_tdiv_r_2exp(e, p, L);
            mpz_neg(e, e);
            return;
        }
        if (n == K && mpz_scan0(p, L - 1) >= K) {  // p = 2^K - 2^L + (p mod 2^L)
            mpz_t neg;
            mpz_roinit_n(neg, mpz_limbs_read(p), -(mp_size_t)mpz_size(p));
            mpz_fdiv_r_2exp(e, neg, L);  // 2^L - (p mod 2^L)
            return;
        }
    }
    mpz_set_ui(e, 1);
    mpz_mul_2exp(e, e, K);
    mpz_sub(e, e, p);
}

size_t truncate_top(mpz_t y, mpz_srcptr x, size_t bits) {
    const size_t n = mpz_sgn(x) ? mpz_sizeinbase(x, 2) : 0;
    const size_t s = n > bits ? n - bits : 0;
    // Shift straight from x (truncation is symmetric in the sign), so only
    // the kept bits are copied.
    if (s)
        mpz_tdiv_q_2exp(y, x, s);
    else
        mpz_set(y, x);
    mpz_abs(y, y);
    return s;
}

void newton_reciprocal(mpz_t y, mpz_srcptr d, size_t b, Multiplier& m) {
    const size_t n = mpz_sizeinbase(d, 2);
    auto sched = schedule(b);
    const size_t c0 = sched.empty() ? b : sched.back() / 2 + 16;
    mpz_t t, dt, e, p;
    mpz_inits(t, dt, e, p, NULL);
    {
        // Start: y = floor(2^(n'+c0) / d') with d' the top c0+64 bits of d.
        size_t sd = truncate_top(dt, d, c0 + 64);
        size_t nd = n - sd;
        mpz_set_ui(t, 1);
        mpz_mul_2exp(t, t, nd + c0);
        mpz_tdiv_q(y, t, dt);
    }
    size_t cur = c0;
    for (auto it = sched.rbegin(); it != sched.rend(); ++it) {
        const size_t c = *it;
        const size_t sd = truncate_top(dt, d, c + kGuard);
        const size_t K = (n - sd) + cur;
        m.multiply(p, dt, y);               // ~ 2^K
        pow2_minus(e, p, K, K - cur + 16);  // signed; |e| < 2^(K - cur + 4)
        const int sign = mpz_sgn(e);
        const size_t se = truncate_top(e, e, c - cur + kGuard);
        if (sign < 0) mpz_neg(e, e);
        m.multiply(p, y, e);
        // term = p * 2^(se + c - cur) / 2^K
        fshift(p, p, (long)K - (long)se - (long)(c - cur));
        mpz_mul_2exp(y, y, c - cur);
        mpz_add(y, y, p);
        cur = c;
    }
    mpz_clears(t, dt, e, p, NULL);
}

void newton_reciprocal_limbs(mpz_t y, mpz_srcptr d, size_t B) {
    constexpr size_t G = 3, kExact = 16;
    const size_t Ld = mpz_size(d);
    std::vector<size_t> sched;
    for (size_t c = B; c > kExact; c = (c + 3) / 2) sched.push_back(c);
    size_t c = sched.empty() ? B : (sched.back() + 3) / 2;
    mpz_t dp, p, e, t;
    mpz_inits(dp, p, e, t, NULL);
    auto top = [&](size_t l) {  // the top l limbs of d
        mpz_tdiv_q_2exp(dp, d, 64 * (Ld - l));
    };
    {
        const size_t l = std::min(Ld, c + G);
        top(l);
        mpz_set_ui(t, 1);
        mpz_mul_2exp(t, t, 64 * (l + c));
        mpz_tdiv_q(y, t, dp);
    }
    for (auto it = sched.rbegin(); it != sched.rend(); ++it) {
        const size_t cp = *it, s = cp - c;
        const size_t l = std::min(Ld, cp + G);
        top(l);
        const size_t K = l + c;
        mpz_mul(p, dp, y);
        mpz_set_ui(e, 1);
        mpz_mul_2exp(e, e, 64 * K);
        mpz_sub(e, e, p);
        const int sign = mpz_sgn(e);
        mpz_abs(e, e);
        const size_t m = K - c + 2;
        if (mpz_size(e) > m) throw std::runtime_error("newton_reciprocal_limbs: correction above its bound");
        const size_t T = s + G;
        const size_t lo = m > T ? m - T : 0;
        mpz_tdiv_q_2exp(t, e, 64 * lo);
        mpz_mul(p, y, t);
        mpz_tdiv_q_2exp(p, p, 64 * (K - lo - s));
        mpz_mul_2exp(y, y, 64 * s);
        if (sign >= 0)
            mpz_add(y, y, p);
        else
            mpz_sub(y, y, p);
        c = cp;
    }
    mpz_clears(dp, p, e, t, NULL);
}

void newton_inv_sqrt_limbs(mpz_t y, unsigned long a, size_t B) {
    constexpr size_t G = 3, kExact = 16;
    std::vector<size_t> sched;
    for (size_t c = B; c > kExact; c = (c + 3) / 2) sched.push_back(c);
    size_t c = sched.empty() ? B : (sched.back() + 3) / 2;
    mpz_t t, p, e;
    mpz_inits(t, p, e, NULL);
    mpz_set_ui(t, 1);  // start: floor(sqrt(floor(beta^(2c) / a)))
    mpz_mul_2exp(t, t, 128 * c);
    mpz_tdiv_q_ui(t, t, a);
    mpz_sqrt(y, t);
    for (auto it = sched.rbegin(); it != sched.rend(); ++it) {
        const size_t cp = *it, s = cp - c;
        mpz_mul(p, y, y);
        mpz_mul_ui(p, p, a);  // ~ beta^(2c)
        mpz_set_ui(e, 1);
        mpz_mul_2exp(e, e, 128 * c);
        mpz_sub(e, e, p);
        const int sign = mpz_sgn(e);
        mpz_abs(e, e);
        const size_t m = c + 2;
        if (mpz_size(e) > m) throw std::runtime_error("newton_inv_sqrt_limbs: correction above its bound");
        const size_t T = s + G;
        const size_t lo = m > T ? m - T : 0;
        mpz_tdiv_q_2exp(t, e, 64 * lo);
        mpz_mul_2exp(t, t, 63);  // the halving of the step, as beta / 2
        mpz_mul(p, y, t);
        mpz_tdiv_q_2exp(p, p, 64 * (2 * c - lo - s + 1));
        mpz_mul_2exp(y, y, 64 * s);
        if (sign >= 0)
            mpz_add(y, y, p);
        else
            mpz_sub(y, y, p);
        c = cp;
    }
    mpz_clears(t, p, e, NULL);
}

void newton_inv_sqrt_ui(mpz_t y, unsigned long a, size_t b, Multiplier& m) {
    auto sched = schedule(b);
    const size_t c0 = sched.empty() ? b : sched.back() / 2 + 16;
    mpz_t t, e, p;
    mpz_inits(t, e, p, NULL);
    // Start: y = isqrt(floor(2^(2 c0) / a)).
    mpz_set_ui(t, 1);
    mpz_mul_2exp(t, t, 2 * c0);
    mpz_tdiv_q_ui(t, t, a);
    mpz_sqrt(y, t);
    size_t cur = c0;
    for (auto it = sched.rbegin(); it != sched.rend(); ++it) {
        const size_t c = *it;
        m.square(p, y);
        mpz_mul_ui(p, p, a);                  // ~ 2^(2 cur)
        pow2_minus(e, p, 2 * cur, cur + 16);  // signed; |e| < 2^(cur + 5)
        const int sign = mpz_sgn(e);
        const size_t se = truncate_top(e, e, c - cur + kGuard);
        if (sign < 0) mpz_neg(e, e);
        m.multiply(p, y, e);
        fshift(p, p, (long)(2 * cur + 1) - (long)se - (long)(c - cur));
        mpz_mul_2exp(y, y, c - cur);
        mpz_add(y, y, p);
        cur = c;
    }
    mpz_clears(t, e, p, NULL);
}

void pow_ui(mpz_t y, unsigned long base, unsigned long e, Multiplier& m) {
    mpz_set_ui(y, 1);
    if (e == 0) return;
    int top = 63;
    while (!((e >> top) & 1)) --top;
    for (int i = top; i >= 0; --i) {
        if (mpz_cmp_ui(y, 1) != 0) m.square(y, y);
        if ((e >> i) & 1) mpz_mul_ui(y, y, base);
    }
}

}  // namespace pi
