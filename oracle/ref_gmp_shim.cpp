// ref_gmp_shim.cpp - C entry point around the reference's own GmpMultiplier
// (src/arithmetic/gmp_multiplier.cpp:10-19), compiled from the reference tree
// by oracle/Makefile into oracle/_ref/libgmpmul_ref.so.  Test and CPU-baseline
// infrastructure only.
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <gmp.h>
#include "arithmetic/gmp_multiplier.h"

namespace {
pi::GmpMultiplier& mult() {
    static pi::GmpMultiplier m;
    return m;
}
}  // namespace

extern "C" {

// out receives an + bn little-endian limbs of |a * b| (square when square != 0).
void ref_gmp_mul(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, uint64_t* out,
                 int square) {
    mpz_t x, y, r;
    mpz_init(x);
    mpz_init(y);
    mpz_init(r);
    mpz_import(x, an, -1, sizeof(uint64_t), 0, 0, a);
    if (square) {
        mult().square(r, x);
    } else {
        mpz_import(y, bn, -1, sizeof(uint64_t), 0, 0, b);
        mult().multiply(r, x, y);
    }
    std::memset(out, 0, (an + bn) * sizeof(uint64_t));
    size_t count = 0;
    mpz_export(out, &count, -1, sizeof(uint64_t), 0, 0, r);
    mpz_clear(x);
    mpz_clear(y);
    mpz_clear(r);
}

// Times `reps` products of the given operands on the calling thread, without
// conversions inside the timed loop; returns seconds per product.
double ref_gmp_time(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, int square,
                    int reps) {
    mpz_t x, y, r;
    mpz_init(x);
    mpz_init(y);
    mpz_init(r);
    mpz_import(x, an, -1, sizeof(uint64_t), 0, 0, a);
    mpz_import(y, bn, -1, sizeof(uint64_t), 0, 0, b);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < reps; ++i) {
        if (square) mult().square(r, x);
        else mult().multiply(r, x, y);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    mpz_clear(x);
    mpz_clear(y);
    mpz_clear(r);
    return ((t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec)) / (reps > 0 ? reps : 1);
}

// Times `reps` products one at a time (each timed on its own, for medians);
// seconds per product into times[0 .. reps).
void ref_gmp_times(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, int square, int reps,
                   double* times) {
    mpz_t x, y, r;
    mpz_init(x);
    mpz_init(y);
    mpz_init(r);
    mpz_import(x, an, -1, sizeof(uint64_t), 0, 0, a);
    mpz_import(y, bn, -1, sizeof(uint64_t), 0, 0, b);
    for (int i = 0; i < reps; ++i) {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        if (square) mult().square(r, x);
        else mult().multiply(r, x, y);
        clock_gettime(CLOCK_MONOTONIC, &t1);
        times[i] = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    }
    mpz_clear(x);
    mpz_clear(y);
    mpz_clear(r);
}

}  // extern "C"
