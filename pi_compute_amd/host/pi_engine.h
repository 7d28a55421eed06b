// pi_engine.h - Chudnovsky pi over a Multiplier (the reference's PiEngine,
// src/engine/pi_engine.h:16-46, floating-point final step of
// src/engine/pi_engine.cpp:433-503).
//
//   terms   = terms_needed(D + 100)
//   (P,Q,R) = binary splitting over [0, terms)
//   pi      = 426880 * Q * sqrt(10005) / R       at (D+100)*3.3219281 + 64 bits
//   digits  = "3." + the first D decimals of floor(pi * 10^D), no newline
#pragma once

#include <cstddef>
#include <string>

#include "binary_splitting.h"
#include "multiplier.h"

namespace pi {

struct PiConfig {
    size_t digits = 1000;
    std::string output_file;  // empty: do not write
    bool verbose = false;
    unsigned num_threads = 0;  // 0: hardware concurrency
    bool level_batches = true;  // level-synchronous tree with one batch per level
};

struct PiResult {
    std::string digits;
    double elapsed_seconds;
    unsigned long terms_used;
    double bs_seconds, final_seconds, conv_seconds;
};

class PiEngine {
public:
    explicit PiEngine(Multiplier& m) : mult_(m) {}
    PiResult compute(const PiConfig& config);
    std::string compute_digits(size_t digits);

private:
    Multiplier& mult_;
};

}  // namespace pi
