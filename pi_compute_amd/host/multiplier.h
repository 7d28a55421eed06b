// multiplier.h - the multiplication strategy interface the engine plugs into.
//
// Same interface as the reference (src/arithmetic/multiplier.h:16-41): callers
// such as binary splitting hold a Multiplier& and never know which backend
// runs.  result must be initialised and may alias either operand.
#pragma once

#include <gmp.h>

#include <memory>

namespace pi {

class Multiplier {
public:
    virtual ~Multiplier() = default;

    // result = a * b
    virtual void multiply(mpz_t result, const mpz_t a, const mpz_t b) = 0;

    // result = a * a; backends with a cheaper squaring path override this.
    virtual void square(mpz_t result, const mpz_t a) { multiply(result, a, a); }
};

using MultiplierPtr = std::unique_ptr<Multiplier>;

}  // namespace pi
