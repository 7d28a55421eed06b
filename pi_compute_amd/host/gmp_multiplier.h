// gmp_multiplier.h - CPU backend: GMP's mpz_mul (the oracle and CPU baseline).
// Counterpart of src/arithmetic/gmp_multiplier.{h,cpp}:10-19.
#pragma once

#include "multiplier.h"

namespace pi {

class GmpMultiplier : public Multiplier {
public:
    void multiply(mpz_t result, const mpz_t a, const mpz_t b) override { mpz_mul(result, a, b); }
    void square(mpz_t result, const mpz_t a) override { mpz_mul(result, a, a); }
};

}  // namespace pi
