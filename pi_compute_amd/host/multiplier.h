// multiplier.h - the multiplication strategy interface the engine plugs into.
//
// Same interface as the reference (src/arithmetic/multiplier.h:16-41): callers
// such as binary splitting hold a Multiplier& and never know which backend
// runs.  result must be initialised and may alias either operand.
//
// Addition: multiply_batch() takes a whole set of independent products (one
// binary-splitting tree level).  Its default runs them on host threads through
// multiply(); backends with a batched device path override it.
#pragma once

#include <gmp.h>

#include <cstddef>
#include <memory>

namespace pi {

struct MulJob {
    mpz_ptr result;  // distinct from every operand of the batch
    mpz_srcptr a;
    mpz_srcptr b;    // b == a squares
};

class Multiplier {
public:
    virtual ~Multiplier() = default;

    // result = a * b
    virtual void multiply(mpz_t result, const mpz_t a, const mpz_t b) = 0;

    // result = a * a; backends with a cheaper squaring path override this.
    virtual void square(mpz_t result, const mpz_t a) { multiply(result, a, a); }

    // Independent products; `threads` host threads may be used (0 = all cores).
    virtual void multiply_batch(const MulJob* jobs, size_t count, unsigned threads = 0);
};

using MultiplierPtr = std::unique_ptr<Multiplier>;

// Runs fn(i) for i in [0, count) on up to `threads` threads, largest-first
// when `cost` is given (cost(i) = relative work).
void parallel_for(size_t count, unsigned threads, void (*fn)(void*, size_t), void* ctx,
                  const double* cost = nullptr);

}  // namespace pi
