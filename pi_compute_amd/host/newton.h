// newton.h - reciprocal, inverse square root and powers with every large
// product sent through a Multiplier (the GPU engine when it is one).
//
// The reference computes the final quotient and square root with GMP's mpf
// (NewtonDivider::sqrt_to_precision, src/arithmetic/newton_divider.cpp:34-39;
// src/engine/pi_engine.cpp:433-470), single-threaded.  These routines are the
// fixed-point Newton iterations (precision doubling; Brent & Zimmermann,
// "Modern Computer Arithmetic", sections 3.4 and 3.5) whose large products
// are ordinary Multiplier calls.  Results carry a bounded error in the last
// place; callers keep guard bits.
#pragma once

#include <gmp.h>

#include <cstddef>

#include "multiplier.h"

namespace pi {

// y ~ 2^(n+b) / d, n = bit length of d > 0; |y - 2^(n+b)/d| <= 4.
void newton_reciprocal(mpz_t y, mpz_srcptr d, size_t b, Multiplier& m);

// The same quotient in whole limbs (beta = 2^64): y ~ beta^(Ld + B) / d with
// Ld = limbs of d > 0, every truncation a whole number of limbs (3 guard
// limbs), exact products on the host.  The host statement of the device
// reciprocal in the final step (pi_scaled_newton_device).
void newton_reciprocal_limbs(mpz_t y, mpz_srcptr d, size_t B);

// The same steps on device numbers through the GPU engine's device entry
// points (device_newton.cpp): d uploaded, y downloaded, the same integer as
// newton_reciprocal_limbs(y, d, B).  Its first steps, up to host_limbs
// precision, run on the host (the same integer whichever side a step runs
// on): device_reciprocal_start computes them alone (no device, any thread),
// and `start` hands that result in (null: computed here; the same host_limbs
// both times).  Throws std::runtime_error if the device fails.
constexpr size_t kReciprocalHostLimbs = 16384;
void device_reciprocal_start(mpz_t t, mpz_srcptr d, size_t B, size_t host_limbs = kReciprocalHostLimbs);
void device_reciprocal_limbs(mpz_t y, mpz_srcptr d, size_t B, mpz_srcptr start = nullptr,
                             size_t host_limbs = kReciprocalHostLimbs);

// y ~ beta^B / sqrt(a) in whole limbs, a > 0 (the step: E = beta^(2c) - a y^2,
// y' = y beta^(c'-c) + y E_t 2^63 >> limbs, 2^63 being the step's halving as
// beta / 2), exact products on the host; and the same steps on device numbers
// (device_newton.cpp), the same integer.
void newton_inv_sqrt_limbs(mpz_t y, unsigned long a, size_t B);
void device_inv_sqrt_limbs(mpz_t y, unsigned long a, size_t B);

// y ~ 2^b / sqrt(a), a > 0; |y - 2^b/sqrt(a)| <= 4.
void newton_inv_sqrt_ui(mpz_t y, unsigned long a, size_t b, Multiplier& m);

// y = base^e exactly; the squarings go through m.
void pow_ui(mpz_t y, unsigned long base, unsigned long e, Multiplier& m);

// e = 2^K - p (signed), p >= 0.  `L` bounds the expected size, |2^K - p| <
// 2^(L-1) (a Newton step's product approximates 2^K): then only the low L bits
// of p are read and written; any p gives the exact difference.
void pow2_minus(mpz_t e, mpz_srcptr p, size_t K, size_t L);

// y = |x| >> s with s chosen so that y keeps the top `bits` bits of |x|
// (s = 0 when |x| is shorter); returns s.  y may alias x.
size_t truncate_top(mpz_t y, mpz_srcptr x, size_t bits);

}  // namespace pi
