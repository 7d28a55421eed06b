// pi_engine.cpp - see pi_engine.h.
#include "pi_engine.h"

#include <chrono>
#include <cstdio>
#include <iostream>
#include <stdexcept>

#include "decimal.h"

namespace pi {

using Clock = std::chrono::steady_clock;
static double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

PiResult PiEngine::compute(const PiConfig& cfg) {
    if (cfg.digits == 0) throw std::invalid_argument("digits must be > 0");
    const auto t0 = Clock::now();
    const size_t precision = cfg.digits + 100;
    const unsigned long terms = BinarySplitting::terms_needed(precision);
    if (cfg.verbose) std::cout << "  terms: " << terms << std::endl;

    BinarySplitting bs(mult_, cfg.num_threads);
    BSResult r = cfg.level_batches ? bs.compute(0, terms, false) : bs.compute_recursive(0, terms);
    const double bs_s = since(t0);
    if (cfg.verbose) {
        std::cout << "  binary splitting: " << bs_s << " s" << std::endl;
        for (size_t i = 0; i < bs.levels().size(); ++i) {
            const auto& L = bs.levels()[i];
            std::cout << "    level " << i << ": " << L.products << " products, max " << L.max_limbs
                      << " limbs, " << L.seconds << " s" << std::endl;
        }
    }

    const auto t1 = Clock::now();
    mpz_realloc2(r.P, 0);
    const mp_bitcnt_t bits = static_cast<mp_bitcnt_t>(precision * 3.3219281) + 64;
    mpf_t s, num, den, x;
    mpf_init2(s, bits);
    mpf_init2(num, bits);
    mpf_init2(den, bits);
    mpf_init2(x, bits);
    mpf_set_ui(s, 10005);
    mpf_sqrt(s, s);
    mpz_mul_ui(r.Q, r.Q, 426880);
    mpf_set_z(num, r.Q);
    mpz_realloc2(r.Q, 0);
    mpf_set_z(den, r.R);
    mpz_realloc2(r.R, 0);
    mpf_mul(x, num, s);
    mpf_div(x, x, den);
    mpz_t scale, pi_int;
    mpz_init(scale);
    mpz_init(pi_int);
    mpz_ui_pow_ui(scale, 10, cfg.digits);
    mpf_set_z(num, scale);
    mpz_clear(scale);
    mpf_mul(x, x, num);
    mpz_set_f(pi_int, x);
    mpf_clear(s);
    mpf_clear(num);
    mpf_clear(den);
    mpf_clear(x);
    const double fin_s = since(t1);

    const auto t2 = Clock::now();
    std::string d = to_decimal(pi_int, cfg.num_threads);
    mpz_clear(pi_int);
    std::string out;
    out.reserve(cfg.digits + 2);
    out.push_back(d[0]);
    out.push_back('.');
    out.append(d, 1, std::min(cfg.digits, d.size() - 1));
    const double conv_s = since(t2);

    if (!cfg.output_file.empty()) {
        FILE* f = std::fopen(cfg.output_file.c_str(), "wb");
        if (!f) throw std::runtime_error("cannot open output file: " + cfg.output_file);
        size_t w = std::fwrite(out.data(), 1, out.size(), f);
        std::fclose(f);
        if (w != out.size()) throw std::runtime_error("short write: " + cfg.output_file);
    }
    if (cfg.verbose)
        std::cout << "  final step: " << fin_s << " s, decimal conversion: " << conv_s << " s" << std::endl;
    return PiResult{std::move(out), since(t0), terms, bs_s, fin_s, conv_s};
}

std::string PiEngine::compute_digits(size_t digits) {
    PiConfig c;
    c.digits = digits;
    return compute(c).digits;
}

}  // namespace pi
