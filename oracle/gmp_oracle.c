/*
 * gmp_oracle.c - the whole-product oracle: GMP mpz_mul on raw magnitudes.
 *
 * TEST INFRASTRUCTURE ONLY.  Restates GmpMultiplier::multiply
 * (src/arithmetic/gmp_multiplier.cpp:10-12: mpz_mul(result, a, b)) for
 * little-endian 64-bit limb arrays, the layout the C ABI under test uses.
 */
#include "ntt_oracle.h"

#include <gmp.h>
#include <string.h>

void orc_gmp_mul(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, uint64_t* out) {
    mpz_t x, y, r;
    mpz_init(x);
    mpz_init(y);
    mpz_init(r);
    mpz_import(x, an, -1, sizeof(uint64_t), 0, 0, a);
    mpz_import(y, bn, -1, sizeof(uint64_t), 0, 0, b);
    mpz_mul(r, x, y);
    memset(out, 0, (an + bn) * sizeof(uint64_t));
    size_t count = 0;
    mpz_export(out, &count, -1, sizeof(uint64_t), 0, 0, r);
    mpz_clear(x);
    mpz_clear(y);
    mpz_clear(r);
}
