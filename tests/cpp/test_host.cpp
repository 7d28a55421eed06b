// test_host.cpp - the host layer on the CPU (GMP multiplier only): binary
// splitting known values (reference tests/test_binary_splitting.cpp:30-110),
// tree orders agreeing, one batch per level, decimal conversion, pi digits
// (reference tests/test_validation.cpp:29-60), dispatch rules and the model
// file.  Usage: test_host PI_1000_FILE.  Exit status = number of failures.
#include <gmp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <mutex>
#include <string>
#include <vector>

#include "binary_splitting.h"
#include "decimal.h"
#include "dispatch.h"
#include "gmp_multiplier.h"
#include "mlp_dispatch.h"
#include "newton.h"
#include "operands.h"
#include "pi_engine.h"
#include "pintt_mlp.h"

namespace {

int failures = 0;

void check(bool ok, const std::string& name) {
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", name.c_str());
    std::fflush(stdout);
    if (!ok) ++failures;
}

// GMP multiplier that records how products arrive.
class CountingMultiplier : public pi::GmpMultiplier {
public:
    void multiply(mpz_t r, const mpz_t a, const mpz_t b) override {
        singles.fetch_add(1);
        mpz_mul(r, a, b);
    }
    void multiply_batch(const pi::MulJob* jobs, size_t count, unsigned threads) override {
        {
            std::lock_guard<std::mutex> l(mu);
            batch_sizes.push_back(count);
        }
        (void)threads;
        for (size_t i = 0; i < count; ++i) mpz_mul(jobs[i].result, jobs[i].a, jobs[i].b);
    }
    std::atomic<size_t> singles{0};
    std::vector<size_t> batch_sizes;
    std::mutex mu;
};

bool same(const pi::BSResult& x, const pi::BSResult& y, bool with_p = true) {
    return (!with_p || mpz_cmp(x.P, y.P) == 0) && mpz_cmp(x.Q, y.Q) == 0 && mpz_cmp(x.R, y.R) == 0;
}

std::string slurp(const char* path) {
    std::ifstream f(path, std::ios::binary);
    return std::string(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

}  // namespace

int main(int argc, char** argv) {
    pi::GmpMultiplier gmp;
    using BS = pi::BinarySplitting;

    check(BS::terms_needed(1000000) >= 70000 && BS::terms_needed(1000000) <= 71000, "TermsNeededFor1MillionDigits");
    check(BS::terms_needed(10000100) == 705145, "TermsNeeded10MDigitsWithGuard");
    check(BS::terms_needed(100000100) == 7051375, "TermsNeeded100MDigitsWithGuard");

    {
        BS bs(gmp, 2);
        pi::BSResult r0 = bs.compute(0, 1);
        check(mpz_cmp_ui(r0.P, 1) == 0 && mpz_cmp_ui(r0.Q, 1) == 0 && mpz_cmp_ui(r0.R, 13591409) == 0,
              "BaseCaseTerm0");
        pi::BSResult r1 = bs.compute(1, 2);
        mpz_t q, r;
        mpz_init_set_str(q, "10939058860032000", 10);
        mpz_init_set_str(r, "-2793657715", 10);
        check(mpz_cmp_si(r1.P, -5) == 0 && mpz_cmp(r1.Q, q) == 0 && mpz_cmp(r1.R, r) == 0, "BaseCaseTerm1");
        pi::BSResult r2 = bs.compute(2, 3);
        mpz_set_str(q, "87512470880256000", 10);
        check(mpz_cmp_si(r2.P, -231) == 0 && mpz_cmp(r2.Q, q) == 0, "BaseCaseTerm2");
        // The 128-bit path (k < 2^24) and the mpz path agree with mpz arithmetic at
        // large k on both sides of the switch.
        {
            bool ok = true;
            mpz_t p, qq, rr;
            mpz_inits(p, qq, rr, NULL);
            for (unsigned long k : {1000ul, 7051375ul, (1ul << 24) - 1, 1ul << 24, (1ul << 24) + 7, 1ul << 30}) {
                pi::BSResult t;
                pi::BinarySplitting::base_case(k, t);
                mpz_set_ui(p, 6 * k - 5);
                mpz_mul_ui(p, p, 2 * k - 1);
                mpz_mul_ui(p, p, 6 * k - 1);
                mpz_neg(p, p);
                mpz_ui_pow_ui(qq, k, 3);
                mpz_mul_ui(qq, qq, 10939058860032000ul);
                mpz_set_ui(rr, 545140134ul);
                mpz_mul_ui(rr, rr, k);
                mpz_add_ui(rr, rr, 13591409ul);
                mpz_mul(rr, rr, p);
                ok = ok && mpz_cmp(t.P, p) == 0 && mpz_cmp(t.Q, qq) == 0 && mpz_cmp(t.R, rr) == 0;
            }
            mpz_clears(p, qq, rr, NULL);
            check(ok, "BaseCaseLargeTerms");
        }
        // [0,2): P = P0 P1, Q = Q0 Q1, R = Q1 R0 + P0 R1
        bs.set_leaf_span(1);
        pi::BSResult m = bs.compute(0, 2);
        mpz_set_str(q, "10939058860032000", 10);
        mpz_mul_ui(r, q, 13591409);
        mpz_sub_ui(r, r, 2793657715ul);
        check(mpz_cmp_si(m.P, -5) == 0 && mpz_cmp(m.Q, q) == 0 && mpz_cmp(m.R, r) == 0, "MergeTwoTerms");
        mpz_clears(q, r, NULL);
    }

    // Every tree order yields the same integers.
    for (unsigned long b : {2ul, 3ul, 17ul, 1000ul, 4099ul}) {
        pi::BSResult ref;
        BS::compute_small(0, b, ref);
        for (unsigned long span : {1ul, 16ul, 256ul}) {
            BS bs(gmp, 4);
            bs.set_leaf_span(span);
            pi::BSResult lv = bs.compute(0, b);
            pi::BSResult lvq = bs.compute(0, b, false);
            pi::BSResult rc = bs.compute_recursive(0, b);
            check(same(lv, ref) && same(lvq, ref, false) && same(rc, ref),
                  "TreeOrdersAgree b=" + std::to_string(b) + " span=" + std::to_string(span));
        }
    }

    // One multiply_batch per level, 4 products per merge (3 at the top without P).
    {
        CountingMultiplier cm;
        BS bs(cm, 4);
        bs.set_leaf_span(8);
        pi::BSResult r = bs.compute(0, 8 * 37, false);  // 37 leaves
        size_t expect_merges = 0;
        bool ok = cm.singles == 0 && cm.batch_sizes.size() == bs.levels().size();
        for (size_t i = 0; ok && i < cm.batch_sizes.size(); ++i) {
            const auto& L = bs.levels()[i];
            bool top = i + 1 == cm.batch_sizes.size();
            ok = L.products == cm.batch_sizes[i] && L.products == 4 * L.merges - (top ? 1 : 0);
            expect_merges += L.merges;
        }
        check(ok && expect_merges == 36, "OneBatchPerLevel");
    }

    // Decimal conversion against mpz_get_str.
    {
        pi::SplitMix64 g(99);
        bool ok = true;
        mpz_t x;
        mpz_init(x);
        std::vector<size_t> bits = {1, 5, 64, 1000, 20000, 70000, 333333, 1000003};
        for (size_t nb : bits) {
            pi::draw_operand(x, g, nb);
            char* s = mpz_get_str(nullptr, 10, x);
            ok = ok && pi::to_decimal(x, 4) == s;
            free(s);
        }
        for (unsigned long k : {1ul, 4095ul, 4096ul, 4097ul, 10000ul, 65537ul}) {
            for (int d : {-1, 0, 1}) {
                mpz_ui_pow_ui(x, 10, k);
                if (d < 0) mpz_sub_ui(x, x, 1);
                if (d > 0) mpz_add_ui(x, x, 1);
                char* s = mpz_get_str(nullptr, 10, x);
                ok = ok && pi::to_decimal(x, 3) == s;
                free(s);
            }
        }
        mpz_set_ui(x, 0);
        ok = ok && pi::to_decimal(x) == "0";
        mpz_clear(x);
        check(ok, "DecimalMatchesMpzGetStr");
    }

    // Newton reciprocal / inverse square root within 4 units of the exact floor.
    {
        pi::SplitMix64 g(31);
        bool ok = true;
        mpz_t d, y, want, t;
        mpz_inits(d, y, want, t, NULL);
        const size_t dbits[] = {1, 7, 64, 300, 5000, 70001, 400000};
        const size_t bs[] = {1, 100, 257, 1000, 33333, 300000};
        for (size_t nb : dbits)
            for (size_t b : bs) {
                pi::draw_operand(d, g, nb);
                pi::newton_reciprocal(y, d, b, gmp);
                size_t n = mpz_sizeinbase(d, 2);
                mpz_set_ui(t, 1);
                mpz_mul_2exp(t, t, n + b);
                mpz_tdiv_q(want, t, d);
                mpz_sub(t, y, want);
                mpz_abs(t, t);
                ok = ok && mpz_cmp_ui(t, 4) <= 0;
            }
        check(ok, "NewtonReciprocalWithin4");
        ok = true;
        for (unsigned long a : {2ul, 3ul, 10005ul, 999999937ul})
            for (size_t b : bs) {
                pi::newton_inv_sqrt_ui(y, a, b, gmp);
                mpz_set_ui(t, 1);
                mpz_mul_2exp(t, t, 2 * b);
                mpz_tdiv_q_ui(t, t, a);
                mpz_sqrt(want, t);
                mpz_sub(t, y, want);
                mpz_abs(t, t);
                ok = ok && mpz_cmp_ui(t, 4) <= 0;
            }
        check(ok, "NewtonInverseSqrtWithin4");
        // Whole-limb reciprocal (the device final step's algorithm): random
        // divisors, 2^(64 L - 1) and all-ones divisors, precisions below and far
        // above the divisor's length.
        ok = true;
        for (int it = 0; it < 600; ++it) {
            const size_t Ld = 1 + (size_t)(it * 37) % 260, B = 1 + (size_t)(it * 53) % 700;
            if (it % 3 == 0) {
                pi::draw_operand(d, g, 64 * Ld);
                mpz_setbit(d, 64 * Ld - 1);
            } else if (it % 3 == 1) {
                mpz_set_ui(d, 0);
                mpz_setbit(d, 64 * Ld - 1);
            } else {
                mpz_set_ui(d, 0);
                mpz_setbit(d, 64 * Ld);
                mpz_sub_ui(d, d, 1);
            }
            pi::newton_reciprocal_limbs(y, d, B);
            mpz_set_ui(t, 0);
            mpz_setbit(t, 64 * (Ld + B));
            mpz_tdiv_q(want, t, d);
            mpz_sub(t, y, want);
            mpz_abs(t, t);
            ok = ok && mpz_cmp_ui(t, 4) <= 0;
        }
        check(ok, "NewtonLimbsWithinFew");
        // Whole-limb inverse square root (the device final step's constant).
        ok = true;
        for (unsigned long a : {2ul, 3ul, 10005ul, 999999937ul, 4294967291ul})
            for (size_t B = 1; B < 700; B += 29) {
                pi::newton_inv_sqrt_limbs(y, a, B);
                mpz_set_ui(t, 0);
                mpz_setbit(t, 128 * B);
                mpz_tdiv_q_ui(t, t, a);
                mpz_sqrt(want, t);
                mpz_sub(t, y, want);
                mpz_abs(t, t);
                ok = ok && mpz_cmp_ui(t, 4) <= 0;
            }
        check(ok, "NewtonInverseSqrtLimbsWithinFew");
        ok = true;
        for (unsigned long e : {0ul, 1ul, 2ul, 77ul, 1000ul, 123457ul}) {
            pi::pow_ui(y, 10, e, gmp);
            mpz_ui_pow_ui(want, 10, e);
            ok = ok && mpz_cmp(y, want) == 0;
        }
        check(ok, "PowerBySquaring");
        // 2^K - p near 2^K on both sides (read from the low L bits) and far
        // from it (the full difference), at and around the bound 2^(L-1).
        ok = true;
        mpz_t p, e, delta;
        mpz_inits(p, e, delta, NULL);
        for (size_t K : {64ul, 100ul, 1000ul, 5001ul})
            for (size_t L : {0ul, 1ul, 2ul, 17ul, K / 2, K - 1, K, K + 5}) {
                std::vector<long> shifts = {-1, 0, 1};
                for (long sh : shifts)
                    for (int off : {-1, 0, 1})
                        for (int side : {-1, 1}) {
                            // delta = 2^(L - 1 + sh) + off (clamped at 0)
                            mpz_set_ui(delta, 0);
                            const long ex = (long)L - 1 + sh;
                            if (ex >= 0) mpz_setbit(delta, (mp_bitcnt_t)ex);
                            if (off < 0 && mpz_sgn(delta) > 0) mpz_sub_ui(delta, delta, 1);
                            if (off > 0) mpz_add_ui(delta, delta, 1);
                            mpz_set_ui(p, 0);
                            mpz_setbit(p, K);
                            if (side < 0) {
                                if (mpz_cmp(delta, p) > 0) continue;
                                mpz_sub(p, p, delta);
                            } else {
                                mpz_add(p, p, delta);
                            }
                            pi::pow2_minus(e, p, K, L);
                            mpz_set_ui(want, 0);
                            mpz_setbit(want, K);
                            mpz_sub(want, want, p);
                            ok = ok && mpz_cmp(e, want) == 0;
                        }
                for (size_t nb : {1ul, K / 3, K, K + 1, 2 * K}) {
                    pi::draw_operand(p, g, nb);
                    pi::pow2_minus(e, p, K, L);
                    mpz_set_ui(want, 0);
                    mpz_setbit(want, K);
                    mpz_sub(want, want, p);
                    ok = ok && mpz_cmp(e, want) == 0;
                }
                mpz_set_ui(p, 0);
                pi::pow2_minus(e, p, K, L);
                ok = ok && mpz_sizeinbase(e, 2) == K + 1 && mpz_scan1(e, 0) == K;
            }
        mpz_clears(p, e, delta, NULL);
        check(ok, "PowerOfTwoMinusProduct");
        mpz_clears(d, y, want, t, NULL);
    }

    // Decimal conversion with Newton quotients (small leaves force many levels).
    {
        pi::SplitMix64 g(77);
        bool ok = true;
        mpz_t x;
        mpz_init(x);
        // Every quotient estimate is the quotient or one less (the device
        // levels rely on it).
        bool est_ok = true;
        auto est = [&est_ok](const pi::DecimalStats& st) {
            est_ok = est_ok && st.min_correction >= 0 && st.max_correction <= 1;
        };
        pi::DecimalStats st;
        for (size_t nb : {20000ul, 100000ul, 654321ul, 2000000ul}) {
            pi::draw_operand(x, g, nb);
            char* s = mpz_get_str(nullptr, 10, x);
            ok = ok && pi::to_decimal(x, gmp, 4, 4096, &st) == s;
            est(st);
            free(s);
        }
        for (unsigned long k : {8192ul, 65536ul, 100000ul, 131072ul}) {
            for (int d : {-1, 0, 1}) {
                mpz_ui_pow_ui(x, 10, k);
                if (d < 0) mpz_sub_ui(x, x, 1);
                if (d > 0) mpz_add_ui(x, x, 1);
                char* s = mpz_get_str(nullptr, 10, x);
                ok = ok && pi::to_decimal(x, gmp, 3, 4096, &st) == s;
                est(st);
                free(s);
            }
        }
        // Multiples of a split power, one below and above them, and runs of
        // zeros and nines inside: remainders at the ends of their range.
        for (unsigned long k : {16384ul, 50000ul}) {
            mpz_t p;
            mpz_init(p);
            mpz_ui_pow_ui(p, 10, 8192);
            for (unsigned long mlt : {1ul, 7ul, 99999ul}) {
                for (int d : {-1, 0, 1}) {
                    mpz_ui_pow_ui(x, 10, k);
                    mpz_sub_ui(x, x, 1);
                    mpz_tdiv_q(x, x, p);
                    mpz_mul_ui(x, x, mlt);
                    mpz_tdiv_r(x, x, p);  // below 10^8192
                    mpz_mul(x, x, p);
                    mpz_mul(x, x, p);
                    if (d < 0) mpz_sub_ui(x, x, 1);
                    if (d > 0) mpz_add_ui(x, x, 1);
                    char* s = mpz_get_str(nullptr, 10, x);
                    ok = ok && pi::to_decimal(x, gmp, 3, 4096, &st) == s;
                    est(st);
                    free(s);
                }
            }
            mpz_clear(p);
        }
        check(est_ok, "DecimalQuotientEstimateWithinOne");
        // Powers built beforehand (for a width with the same k, or another k,
        // which is then ignored) give the same digits.
        bool pre_ok = true;
        for (size_t nb : {100000ul, 654321ul}) {
            pi::draw_operand(x, g, nb);
            char* s = mpz_get_str(nullptr, 10, x);
            const size_t w = std::strlen(s);
            for (size_t pw : {w, w - 1, 2 * w, w / 3}) {
                pi::DecimalPowers pre;
                pi::decimal_powers(pre, pw, gmp, 4096);
                pre_ok = pre_ok && pi::to_decimal(x, gmp, 4, 4096, nullptr, &pre) == s;
            }
            free(s);
        }
        mpz_clear(x);
        check(ok, "DecimalWithNewtonQuotients");
        check(pre_ok, "DecimalWithPowersBuiltBeforehand");
    }

    // Pi digits.
    {
        pi::PiEngine engine(gmp);
        check(engine.compute_digits(50) == "3.14159265358979323846264338327950288419716939937510", "First50Digits");
        check(engine.compute_digits(100) ==
                  "3.14159265358979323846264338327950288419716939937510"
                  "58209749445923078164062862089986280348253421170679",
              "First100Digits");
        if (argc > 1) {
            std::string want = slurp(argv[1]);
            pi::PiConfig c;
            c.digits = 1000;
            c.level_batches = true;
            std::string lv = engine.compute(c).digits;
            c.level_batches = false;
            std::string rc = engine.compute(c).digits;
            check(want.size() == 1002 && lv == want && rc == want, "First1000DigitsMatchReferenceFile");
            c.final_step = pi::PiConfig::FinalStep::Newton;
            c.conversion = pi::PiConfig::Conversion::Multiplier;
            check(engine.compute(c).digits == want, "First1000DigitsNewtonFinalStep");
        }
        for (size_t digits : {1ul, 2ul, 17ul, 5000ul, 250000ul}) {
            pi::PiConfig a, b;
            a.digits = b.digits = digits;
            b.final_step = pi::PiConfig::FinalStep::Newton;
            b.conversion = pi::PiConfig::Conversion::Multiplier;
            const std::string mpf = engine.compute(a).digits;
            check(mpf == engine.compute(b).digits, "NewtonFinalStepMatchesMpf D=" + std::to_string(digits));
            // sqrt(10005) and 10^D on another thread beside binary splitting, or after it
            b.overlap_constants = !b.overlap_constants;
            check(mpf == engine.compute(b).digits, "NewtonConstantsEitherOrder D=" + std::to_string(digits));
        }
    }

    // Dispatch rules.
    {
        pi::ThresholdDispatcher t(10000);
        check(t.choose({9999, 50000, 1, false}) == pi::Path::Cpu && t.choose({10000, 10000, 1, true}) == pi::Path::Gpu &&
                  t.choose({50000, 9999, 8, false}) == pi::Path::Cpu,
              "ThresholdRule");
        pi::MlpModel m;
        m.outputs = 2;
        m.mean = {10, 9, 3, 0.3f, 0.5f, 0.4f, 0.3f, 0.2f};
        m.scale = {4, 4, 2, 0.5f, 0.3f, 0.6f, 0.7f, 0.8f};
        m.lo = {3, 3, 0, 0, -1, -1, -1, -1};
        m.hi = {25, 25, 12, 1, 11, 11, 11, 11};
        m.params.resize(gpu_mlp_param_count(2));
        pi::SplitMix64 g(5);
        for (auto& v : m.params) v = (float)(g.uniform() - 0.5);
        const char* path = "test_host_model.tmp";
        pi::MlpModel back;
        bool ok = m.save(path) && back.load(path) && back.params == m.params && back.mean == m.mean &&
                  back.scale == m.scale && back.lo == m.lo && back.hi == m.hi;
        std::remove(path);
        check(ok, "ModelFileRoundTrip");
        // Out-of-range shapes go to the fallback.
        pi::FixedDispatcher cpu(pi::Path::Cpu);
        pi::MlpModel big = m;
        const int H = PINTT_MLP_HIDDEN, oB3 = H * PINTT_MLP_INPUTS + H + H * H + H + H * 2;
        for (auto& v : big.params) v = 0;
        big.params[oB3 + 0] = 5.f;  // host predicted slow, GPU fast
        big.params[oB3 + 1] = 0.f;
        pi::MlpDispatcher d(big, &cpu, 16);
        check(d.choose({1000, 1000, 1, false}) == pi::Path::Gpu &&
                  d.choose({1u << 27, 1000, 1, false}) == pi::Path::Cpu,
              "MlpFallbackOutsideTrainedRange");
    }

    // The batch split is within one product of the best max(device sum,
    // host sum / threads, longest host product) over all subsets (brute force
    // over 2^12 assignments, host-side scoring), and never worse than all-host
    // or all-device.
    {
        pi::MlpModel m;
        m.outputs = 2;
        m.mean = {12, 10, 4, 0.3f, 0.5f, 0.4f, 0.3f, 0.2f};
        m.scale = {4, 4, 3, 0.5f, 0.3f, 0.6f, 0.7f, 0.8f};
        m.lo = {0, 0, 0, 0, -1, -1, -1, -1};
        m.hi = {30, 30, 20, 1, 11, 11, 11, 11};
        m.params.resize(gpu_mlp_param_count(2));
        pi::SplitMix64 g(17);
        for (auto& v : m.params) v = (float)(0.6 * (g.uniform() - 0.5));
        pi::FixedDispatcher cpu(pi::Path::Cpu);
        const unsigned threads = 3;
        pi::MlpDispatcher d(m, &cpu, threads);
        bool ok = true;
        for (int trial = 0; trial < 20 && ok; ++trial) {
            const size_t n = 12;
            std::vector<pi::ProductShape> sh(n);
            std::vector<double> c(n), gg(n);
            for (size_t i = 0; i < n; ++i) {
                size_t an = 16 + (size_t)(g.uniform() * 200000), bn = 16 + (size_t)(g.uniform() * 200000);
                sh[i] = pi::ProductShape{an, bn, n, g.uniform() < 0.3};
                auto z = m.standardise(pi::MlpModel::features(sh[i]));
                float y[2];
                m.forward(z.data(), y);
                c[i] = std::exp((double)y[0]);
                gg[i] = std::exp((double)y[1]);
            }
            std::vector<pi::Path> out(n);
            d.choose_batch(sh.data(), n, out.data());
            auto cost = [&](unsigned mask) {
                double dv = 0, hs = 0, lg = 0;
                for (size_t i = 0; i < n; ++i) {
                    if (mask >> i & 1) {
                        dv += gg[i];
                    } else {
                        hs += c[i];
                        lg = std::max(lg, c[i]);
                    }
                }
                return std::max({dv, hs / threads, lg});
            };
            unsigned chosen = 0;
            for (size_t i = 0; i < n; ++i) chosen |= (out[i] == pi::Path::Gpu ? 1u : 0u) << i;
            double best = 1e300, slack = 0;
            for (unsigned mask = 0; mask < (1u << n); ++mask) best = std::min(best, cost(mask));
            for (size_t i = 0; i < n; ++i) slack = std::max(slack, std::max(gg[i], c[i] / threads));
            ok = cost(chosen) <= (best + slack) * (1 + 1e-6) && cost(chosen) <= cost(0) * (1 + 1e-6) &&
                 cost(chosen) <= cost((1u << n) - 1) * (1 + 1e-6);
        }
        check(ok, "BatchSplitWithinOneProductOfBest");
    }

    // parallel_for visits every index once.
    {
        std::vector<std::atomic<int>> hits(1000);
        std::vector<double> cost(1000);
        for (size_t i = 0; i < cost.size(); ++i) cost[i] = (double)((i * 7919) % 1000);
        pi::parallel_for(
            hits.size(), 7, [](void* p, size_t i) { (*static_cast<std::vector<std::atomic<int>>*>(p))[i]++; }, &hits,
            cost.data());
        bool ok = true;
        for (auto& h : hits) ok = ok && h.load() == 1;
        check(ok, "ParallelForCoversEachIndexOnce");
    }

    std::printf("%d failure(s)\n", failures);
    return failures;
}
