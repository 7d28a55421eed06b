// pi_engine.cpp - see pi_engine.h.
#include "pi_engine.h"

#include <chrono>
#include <cstdio>
#include <future>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <thread>

#include "decimal.h"
#include "newton.h"
#include "pintt_gpu.h"

namespace pi {

using Clock = std::chrono::steady_clock;
static double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

PiResult PiEngine::compute(const PiConfig& cfg) {
    if (cfg.digits == 0) throw std::invalid_argument("digits must be > 0");
    const auto t0 = Clock::now();
    const size_t precision = cfg.digits + 100;
    const unsigned long terms = BinarySplitting::terms_needed(precision);
    if (cfg.verbose) std::cout << "  terms: " << terms << std::endl;

    // Working precision of the Newton final step (bits + 64 below).
    const mp_bitcnt_t bits = static_cast<mp_bitcnt_t>(precision * 3.3219281) + 64;
    const bool newton = cfg.final_step == PiConfig::FinalStep::Newton;
    const bool conv_mult = cfg.conversion == PiConfig::Conversion::Multiplier;
    // With device-resident levels the Newton final step also runs on device
    // numbers, cut to whole limbs: its precision is rounded up to a limb.
    const bool final_device = newton && cfg.level_batches && cfg.device_levels && cfg.device_final;
    const size_t fin_w = final_device ? (bits + 64 + 63) / 64 * 64 : bits + 64;
    // Device-resident conversion levels end in device leaves, which take at most
    // PINTT_DECIMAL_LEAF_MAX_DIGITS digits: the powers are built for that leaf.
    const bool conv_device = conv_mult && cfg.level_batches && cfg.device_levels && cfg.device_conversion;
    const size_t conv_leaf = conv_device ? std::min<size_t>(cfg.conversion_leaf_digits, PINTT_DECIMAL_LEAF_MAX_DIGITS)
                                         : cfg.conversion_leaf_digits;
    // What does not depend on P, Q, R, built beside binary splitting on two
    // more threads: the final step's constants, and the conversion's powers of
    // ten and their reciprocals (pi * 10^D has D + 1 digits).
    std::unique_ptr<FinalConstants> pre;
    std::unique_ptr<DecimalPowers> dpow;
    std::promise<void> pre_set, dpow_set;
    std::future<void> pre_done, dpow_done;
    std::thread beside, beside2;
    if (cfg.overlap_constants && (newton || conv_mult)) {
        if (newton) {
            pre = std::make_unique<FinalConstants>();
            pre_done = pre_set.get_future();
        }
        if (conv_mult) {
            dpow = std::make_unique<DecimalPowers>();
            dpow_done = dpow_set.get_future();
        }
        if (pre)
            beside = std::thread([this, &cfg, fin_w, final_device, &pre, &pre_set]() {
                try {
                    try {
                        final_constants(*pre, cfg.digits, fin_w, mult_, final_device);
                    } catch (const std::runtime_error& e) {
                        if (!final_device) throw;
                        std::cerr << "  device inverse square root failed (" << e.what() << "); host instead"
                                  << std::endl;
                        pre = std::make_unique<FinalConstants>();
                        final_constants(*pre, cfg.digits, fin_w, mult_, false);
                    }
                    pre_set.set_value();
                } catch (...) {
                    pre_set.set_exception(std::current_exception());
                }
            });
        beside2 = std::thread([this, &cfg, conv_leaf, conv_device, &dpow, &dpow_set]() {
            if (dpow) {
                try {
                    try {
                        decimal_powers(*dpow, cfg.digits + 1, mult_, conv_leaf, conv_device);
                    } catch (const std::runtime_error& e) {
                        if (!conv_device) throw;
                        std::cerr << "  device reciprocals failed (" << e.what() << "); host reciprocals instead"
                                  << std::endl;
                        dpow = std::make_unique<DecimalPowers>();
                        decimal_powers(*dpow, cfg.digits + 1, mult_, conv_leaf, false);
                    }
                    dpow_set.set_value();
                } catch (...) {
                    dpow_set.set_exception(std::current_exception());
                }
            }
        });
    }
    // The threads are joined on every way out of this function.
    struct Joiner {
        std::thread& t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    } joiner{beside}, joiner2{beside2};

    BinarySplitting bs(mult_, cfg.num_threads);
    bs.set_switch_dispatcher(cfg.device_switch);
    bs.set_segments(cfg.bs_segments);
    bs.set_leaf_span(cfg.bs_leaf_span ? cfg.bs_leaf_span : (cfg.level_batches && cfg.device_levels ? 128 : 256));
    bs.set_device_leaves(cfg.level_batches && cfg.device_levels && cfg.device_leaves);
    BSResult r;
    std::unique_ptr<DeviceQR> qr;  // Q and R left on the device for the device final step
    if (cfg.level_batches && cfg.device_levels) {
        try {
            if (final_device)
                qr = bs.compute_device_qr(0, terms, cfg.device_from_limbs);
            else
                r = bs.compute_device(0, terms, false, cfg.device_from_limbs);
        } catch (const std::exception& e) {
            // Device trouble: the same tree on the host path, exact either way.
            std::cerr << "  device-resident levels failed (" << e.what() << "); host levels instead" << std::endl;
            qr.reset();
            r = bs.compute(0, terms, false);
        }
    } else {
        r = cfg.level_batches ? bs.compute(0, terms, false) : bs.compute_recursive(0, terms);
    }
    const double bs_s = since(t0);
    if (cfg.verbose) {
        std::cout << "  binary splitting: " << bs_s << " s (host subtrees " << bs.leaf_seconds() << " s";
        if (cfg.level_batches && cfg.device_levels) std::cout << ", waited for " << bs.leaf_wait_seconds() << " s";
        std::cout << ", upload " << bs.upload_seconds() << " s)" << std::endl;
        for (size_t i = 0; i < bs.levels().size(); ++i) {
            const auto& L = bs.levels()[i];
            std::cout << "    level " << i << ": " << L.products << " products, max " << L.max_limbs
                      << " limbs, " << L.seconds << " s" << std::endl;
        }
    }

    const auto t1 = Clock::now();
    mpz_realloc2(r.P, 0);
    if (pre_done.valid()) pre_done.get();  // rethrows what the constants' thread threw
    const double pre_wait_s = since(t1);
    mpz_t pi_int;
    mpz_init(pi_int);
    NewtonStats ns;
    bool fin_on_device = false;
    std::unique_ptr<DeviceInt> pi_dev;  // the final step's result left on the device for the conversion
    if (newton && (qr ? qr->r_sign > 0 : mpz_sgn(r.R) > 0)) {
        if (final_device) {
            try {
                const FinalConstants* c = pre.get();
                FinalConstants own;
                if (!c) {
                    final_constants(own, cfg.digits, fin_w, mult_);
                    c = &own;
                }
                if (qr && conv_device) {
                    pi_dev = std::make_unique<DeviceInt>();
                    pi_scaled_newton_device(*pi_dev, *qr, fin_w, *c);
                } else if (qr) {
                    pi_scaled_newton_device(pi_int, *qr, fin_w, *c);
                } else {
                    pi_scaled_newton_device(pi_int, r.Q, r.R, fin_w, *c);
                }
                fin_on_device = true;
            } catch (const std::exception& e) {
                pi_dev.reset();
                std::cerr << "  device final step failed (" << e.what() << "); host final step instead" << std::endl;
            }
        }
        if (!fin_on_device) {
            if (qr) qr->download(r);
            pi_scaled_newton(pi_int, r.Q, r.R, cfg.digits, fin_w, mult_, &ns, pre.get());
        }
        qr.reset();
        mpz_realloc2(r.Q, 0);
        mpz_realloc2(r.R, 0);
    } else {
        if (qr) qr->download(r);
        qr.reset();
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
        mpz_t scale;
        mpz_init(scale);
        mpz_ui_pow_ui(scale, 10, cfg.digits);
        mpf_set_z(num, scale);
        mpz_clear(scale);
        mpf_mul(x, x, num);
        mpz_set_f(pi_int, x);
        mpf_clear(s);
        mpf_clear(num);
        mpf_clear(den);
        mpf_clear(x);
    }
    const double fin_s = since(t1);

    const auto t2 = Clock::now();
    if (dpow_done.valid()) dpow_done.get();
    const double dpow_wait_s = since(t2);
    DecimalStats ds;
    std::string d;
    if (conv_device) {
        try {
            if (pi_dev) {
                DecimalPowers own;
                const DecimalPowers* dp = dpow.get();
                if (!dp) {
                    decimal_powers(own, cfg.digits + 1, mult_, conv_leaf, true);
                    dp = &own;
                }
                d = to_decimal_device(pi_dev->x.d, pi_dev->x.n, cfg.num_threads, conv_leaf, &ds, *dp);
            } else {
                d = to_decimal(pi_int, mult_, cfg.num_threads, conv_leaf, &ds, dpow.get(), true);
            }
        } catch (const std::exception& e) {
            std::cerr << "  device conversion failed (" << e.what() << "); host levels instead" << std::endl;
            if (pi_dev) {  // the value back to the host first
                const size_t n = pi_dev->x.n;
                uint64_t* h = reinterpret_cast<uint64_t*>(mpz_limbs_write(pi_int, (mp_size_t)n));
                const uint64_t* src = pi_dev->x.d;
                size_t nn = n;
                if (gpu_ntt_download(1, &h, &src, &nn) != PINTT_OK) throw;
                mpz_limbs_finish(pi_int, (mp_size_t)n);
                pi_dev.reset();
            }
            ds = DecimalStats{};
            d = to_decimal(pi_int, mult_, cfg.num_threads, conv_leaf, &ds, dpow.get());
        }
        pi_dev.reset();
    } else {
        d = conv_mult ? to_decimal(pi_int, mult_, cfg.num_threads, conv_leaf, &ds, dpow.get())
                      : to_decimal(pi_int, cfg.num_threads);
    }
    mpz_clear(pi_int);
    // "3." and the first D decimals, formatted in the digit string itself (a
    // 100 MB string at 10^8 digits: no second copy).
    if (d.size() > cfg.digits + 1) d.resize(cfg.digits + 1);
    d.insert(1, 1, '.');
    std::string out = std::move(d);
    const double conv_s = since(t2);

    if (!cfg.output_file.empty()) {
        FILE* f = std::fopen(cfg.output_file.c_str(), "wb");
        if (!f) throw std::runtime_error("cannot open output file: " + cfg.output_file);
        size_t w = std::fwrite(out.data(), 1, out.size(), f);
        std::fclose(f);
        if (w != out.size()) throw std::runtime_error("short write: " + cfg.output_file);
    }
    if (cfg.verbose) {
        std::cout << "  final step: " << fin_s << " s, decimal conversion: " << conv_s << " s" << std::endl;
        if (fin_on_device) {
            std::cout << "    final step on device-resident numbers" << std::endl;
        } else if (newton) {
            std::cout << "    final step: inverse square root " << ns.inv_sqrt_s << " s, 10^D " << ns.pow10_s
                      << " s, reciprocal " << ns.reciprocal_s << " s, products " << ns.products_s << " s";
            if (pre) std::cout << " (square root and 10^D beside binary splitting, waited " << pre_wait_s << " s)";
            std::cout << std::endl;
        }
        if (conv_mult) {
            std::cout << "    conversion" << (conv_device ? " (device-resident levels)" : "") << ": powers "
                      << ds.powers_s << " s, reciprocals " << ds.reciprocals_s
                      << " s, quotient products " << ds.products_s << " s, corrections " << ds.corrections_s
                      << " s, leaves " << ds.leaves_s << " s";
            if (dpow) std::cout << " (powers and reciprocals beside binary splitting, waited " << dpow_wait_s << " s)";
            std::cout << std::endl;
        }
    }
    return PiResult{std::move(out), since(t0), terms, bs_s, fin_s, conv_s};
}

FinalConstants::FinalConstants() { mpz_inits(inv_sqrt, pow10, NULL); }
FinalConstants::~FinalConstants() { mpz_clears(inv_sqrt, pow10, NULL); }

void final_constants(FinalConstants& c, size_t digits, size_t w, Multiplier& m, bool device) {
    auto t = Clock::now();
    if (device && w % 64 == 0)
        device_inv_sqrt_limbs(c.inv_sqrt, 10005, w / 64);  // the same 2^w / sqrt(10005), within 4
    else
        newton_inv_sqrt_ui(c.inv_sqrt, 10005, w, m);  // sqrt(10005) = 10005 * y / 2^w
    c.inv_sqrt_s = since(t);
    t = Clock::now();
    pow_ui(c.pow10, 10, digits, m);
    c.pow10_s = since(t);
}

void pi_scaled_newton(mpz_t out, mpz_srcptr Q, mpz_srcptr R, size_t digits, size_t w, Multiplier& m,
                      NewtonStats* stats, const FinalConstants* pre) {
    NewtonStats st;
    auto clock = Clock::now();
    auto lap = [&clock]() {
        const auto now = Clock::now();
        const double s = std::chrono::duration<double>(now - clock).count();
        clock = now;
        return s;
    };
    // 426880 * sqrt(10005) * Q * 10^D / R
    //   = 10005 * (426880 Q) * (2^w / sqrt(10005)) * 10^D * (2^(nr+w) / Rt) / 2^(...)
    // with every operand cut to its top w bits and the exponents of 2 tracked.
    // The reciprocal of R runs on another thread while the numerator's
    // products run here; they meet in the last product.
    mpz_t a, y, t, inv, rt;
    mpz_inits(a, y, t, inv, rt, NULL);
    size_t sr = 0, nr = 0;
    double recip_s = 0;
    auto recip = std::async(std::launch::async, [&]() {
        const auto t0 = Clock::now();
        sr = truncate_top(rt, R, w);
        nr = mpz_sizeinbase(rt, 2);
        newton_reciprocal(inv, rt, w, m);  // 1/R ~ inv / 2^(nr + w + sr)
        recip_s = since(t0);
    });
    long ex = 0;  // value = integer * 2^ex
    mpz_mul_ui(a, Q, 426880);
    ex += (long)truncate_top(a, a, w);
    if (pre) {
        st.inv_sqrt_s = pre->inv_sqrt_s;
    } else {
        newton_inv_sqrt_ui(y, 10005, w, m);  // sqrt(10005) = 10005 * y / 2^w
        st.inv_sqrt_s = lap();
    }
    m.multiply(t, a, pre ? pre->inv_sqrt : y);
    ex += (long)truncate_top(t, t, w) - (long)w;
    st.products_s += lap();
    if (pre) {
        m.multiply(y, t, pre->pow10);
        st.pow10_s = pre->pow10_s;
    } else {
        pow_ui(a, 10, digits, m);
        st.pow10_s = lap();
        m.multiply(y, t, a);
    }
    ex += (long)truncate_top(y, y, w);
    st.products_s += lap();
    recip.get();  // rethrows what the reciprocal's thread threw
    st.reciprocal_s = recip_s;
    lap();
    ex -= (long)(nr + w + sr);
    m.multiply(a, y, inv);
    mpz_mul_ui(a, a, 10005);
    if (ex >= 0)
        mpz_mul_2exp(out, a, (mp_bitcnt_t)ex);
    else
        mpz_fdiv_q_2exp(out, a, (mp_bitcnt_t)(-ex));
    mpz_clears(a, y, t, inv, rt, NULL);
    st.products_s += lap();
    if (stats) *stats = st;
}

std::string PiEngine::compute_digits(size_t digits) {
    PiConfig c;
    c.digits = digits;
    return compute(c).digits;
}

}  // namespace pi
