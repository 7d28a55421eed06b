// test_hip_multiplier.cpp - HipNttMultiplier against GmpMultiplier (mpz_cmp),
// the cases of the reference's GPU suite (tests/test_gpu_ntt_multiplier.cpp:65-145)
// plus signs, aliasing, seeded sizes and the fallback paths.  Needs a gfx950
// device; run by tests/test_gpu_host.py.  Exit status = number of failures.
#include <gmp.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "binary_splitting.h"
#include "dispatch.h"
#include "gmp_multiplier.h"
#include "decimal.h"
#include "hip_ntt_multiplier.h"
#include "operands.h"
#include "pi_engine.h"
#include "newton.h"

namespace {

int failures = 0;

void check(bool ok, const std::string& name) {
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", name.c_str());
    std::fflush(stdout);
    if (!ok) ++failures;
}

bool same_product(pi::Multiplier& m, mpz_t a, mpz_t b) {
    mpz_t got, want;
    mpz_inits(got, want, NULL);
    m.multiply(got, a, b);
    mpz_mul(want, a, b);
    bool ok = mpz_cmp(got, want) == 0;
    mpz_clears(got, want, NULL);
    return ok;
}

void pow_minus_one(mpz_t r, unsigned long base, unsigned long e) {
    mpz_ui_pow_ui(r, base, e);
    mpz_sub_ui(r, r, 1);
}

// splitmix64-seeded operand of exactly `bits` bits
void seeded(mpz_t r, uint64_t seed, size_t bits) {
    size_t n = (bits + 63) / 64;
    std::vector<uint64_t> limbs(n);
    uint64_t s = seed;
    for (size_t i = 0; i < n; ++i) {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        limbs[i] = z ^ (z >> 31);
    }
    size_t top = bits - 64 * (n - 1);
    if (top < 64) limbs[n - 1] &= (1ull << top) - 1;
    limbs[n - 1] |= 1ull << (top - 1);
    mpz_import(r, n, -1, 8, 0, 0, limbs.data());
}

}  // namespace

int main() {
    pi::HipNttMultiplier gpu(1);  // threshold 1: every non-trivial product on the GPU
    mpz_t a, b, r, e;
    mpz_inits(a, b, r, e, NULL);

    mpz_set_ui(a, 12345);
    mpz_set_ui(b, 67890);
    check(same_product(gpu, a, b), "MultiplySmallNumbers");
    mpz_set_ui(a, 999999);
    mpz_set_ui(b, 0);
    check(same_product(gpu, a, b), "MultiplyByZero");
    mpz_set_str(a, "123456789012345678901234567890", 10);
    mpz_set_ui(b, 1);
    check(same_product(gpu, a, b), "MultiplyByOne");
    mpz_set_si(a, -12345);
    mpz_set_si(b, 67890);
    check(same_product(gpu, a, b), "MultiplyNegative");
    mpz_set_si(a, -7);
    mpz_set_si(b, -13);
    check(same_product(gpu, a, b), "MultiplyBothNegative");
    pow_minus_one(a, 2, 3321);
    pow_minus_one(b, 3, 2095);
    check(same_product(gpu, a, b), "Multiply1000DigitNumbers");
    pow_minus_one(a, 2, 33219);
    pow_minus_one(b, 3, 20959);
    check(same_product(gpu, a, b), "Multiply10000DigitNumbers");
    pow_minus_one(a, 2, 332192);
    pow_minus_one(b, 3, 209590);
    check(same_product(gpu, a, b), "Multiply100000DigitNumbers");
    mpz_neg(b, b);
    check(same_product(gpu, a, b), "Multiply100000DigitMixedSign");

    pow_minus_one(a, 2, 33219);
    gpu.square(r, a);
    mpz_mul(e, a, a);
    check(mpz_cmp(r, e) == 0, "SquareMatchesCpu");
    mpz_neg(a, a);
    gpu.square(r, a);
    check(mpz_cmp(r, e) == 0, "SquareNegative");

    {
        pi::HipNttMultiplier high(999999);
        mpz_set_ui(a, 12345);
        mpz_set_ui(b, 67890);
        high.multiply(r, a, b);
        check(mpz_cmp_ui(r, 838102050ul) == 0 && high.stats().cpu_calls.load() == 1 &&
                  high.stats().gpu_calls.load() == 0,
              "ThresholdFallback");
    }

    // Aliasing: result is one of the operands (allowed by mpz semantics).
    seeded(a, 1, 1 << 20);
    seeded(b, 2, 700000);
    mpz_mul(e, a, b);
    mpz_set(r, a);
    gpu.multiply(r, r, b);
    check(mpz_cmp(r, e) == 0, "AliasResultIsA");
    mpz_set(r, b);
    gpu.multiply(r, a, r);
    check(mpz_cmp(r, e) == 0, "AliasResultIsB");
    mpz_mul(e, a, a);
    mpz_set(r, a);
    gpu.multiply(r, r, r);
    check(mpz_cmp(r, e) == 0, "AliasSquareInPlace");
    mpz_set(r, a);
    gpu.square(r, r);
    check(mpz_cmp(r, e) == 0, "AliasSquareCall");

    // Seeded sweep over sizes (bits of a, bits of b).
    const size_t sizes[][2] = {{64, 64},         {4096, 100},        {1 << 16, 1 << 16},
                               {1 << 20, 1 << 18}, {1 << 22, 1 << 22}, {(1 << 24) + 17, 999983}};
    for (size_t i = 0; i < sizeof(sizes) / sizeof(sizes[0]); ++i) {
        seeded(a, 10 + i, sizes[i][0]);
        seeded(b, 20 + i, sizes[i][1]);
        check(same_product(gpu, a, b), "Seeded" + std::to_string(sizes[i][0]) + "x" +
                                           std::to_string(sizes[i][1]));
    }

    // Oversize: above the engine's limit the wrapper falls back to mpz_mul.
    {
        pi::HipNttMultiplier m(1);
        seeded(a, 77, (size_t)64 << 23);
        seeded(b, 78, ((size_t)64 << 23) + 64);
        m.multiply(r, a, b);
        mpz_mul(e, a, b);
        check(mpz_cmp(r, e) == 0 && m.stats().fallback_calls.load() == 1, "OversizeFallsBackExact");
    }

    // multiply_batch: mixed sizes, signs, zeros and squares in one call, batched
    // and per-call, with host-path products running beside the device batch.
    {
        const size_t n = 40;
        std::vector<__mpz_struct> x(n), y(n), z(n);
        for (size_t i = 0; i < n; ++i) {
            mpz_init(&x[i]);
            mpz_init(&y[i]);
            mpz_init(&z[i]);
            seeded(&x[i], 100 + i, 64 + ((i * 7919) % 40) * 20000);
            seeded(&y[i], 200 + i, 64 + ((i * 104729) % 37) * 9000);
            if (i % 5 == 1) mpz_neg(&x[i], &x[i]);
            if (i % 7 == 3) mpz_neg(&y[i], &y[i]);
            if (i == 11) mpz_set_ui(&y[i], 0);
        }
        std::vector<pi::MulJob> jobs(n);
        for (size_t i = 0; i < n; ++i) jobs[i] = pi::MulJob{&z[i], &x[i], (i % 3 == 2) ? &x[i] : &y[i]};
        for (int mode = 0; mode < 3; ++mode) {
            pi::HipNttMultiplier m(mode == 2 ? 5000 : 1);  // mode 2: part of the batch on the host
            m.set_batching(mode != 1);
            for (size_t i = 0; i < n; ++i) mpz_set_ui(&z[i], 12345);
            m.multiply_batch(jobs.data(), n, 4);
            bool ok = true;
            for (size_t i = 0; i < n; ++i) {
                mpz_mul(e, jobs[i].a, jobs[i].b);
                ok = ok && mpz_cmp(&z[i], e) == 0;
            }
            const char* names[3] = {"BatchMixedSignsZerosSquares", "BatchPerCallMode", "BatchSplitHostAndDevice"};
            bool split_ok = mode != 2 || (m.stats().cpu_calls.load() > 0 && m.stats().gpu_calls.load() > 0);
            bool batch_ok = mode != 0 || m.stats().gpu_batches.load() == 1;
            check(ok && split_ok && batch_ok, names[mode]);
        }
        for (size_t i = 0; i < n; ++i) {
            mpz_clear(&x[i]);
            mpz_clear(&y[i]);
            mpz_clear(&z[i]);
        }
    }

    // Binary splitting through the GPU multiplier equals the host recursion.
    {
        pi::HipNttMultiplier m(64);
        pi::BinarySplitting bs(m, 8);
        bs.set_leaf_span(64);
        pi::BSResult want;
        pi::BinarySplitting::compute_small(0, 30000, want);
        pi::BSResult lv = bs.compute(0, 30000);
        m.set_batching(false);
        pi::BSResult pc = bs.compute(0, 30000);
        pi::BSResult rc = bs.compute_recursive(0, 30000);
        auto eq = [](const pi::BSResult& u, const pi::BSResult& v) {
            return mpz_cmp(u.P, v.P) == 0 && mpz_cmp(u.Q, v.Q) == 0 && mpz_cmp(u.R, v.R) == 0;
        };
        check(eq(lv, want) && eq(pc, want) && eq(rc, want) && m.stats().gpu_calls.load() > 0,
              "BinarySplittingOnGpuMatchesHost");
        for (size_t from : {(size_t)0, (size_t)500, (size_t)5000}) {
            for (unsigned long b : {30000ul, 4097ul, 3ul}) {
                pi::BSResult w;
                pi::BinarySplitting::compute_small(0, b, w);
                pi::BSResult dv = bs.compute_device(0, b, true, from);
                pi::BSResult dq = bs.compute_device(0, b, false, from);
                check(eq(dv, w) && mpz_cmp(dq.Q, w.Q) == 0 && mpz_cmp(dq.R, w.R) == 0,
                      "DeviceResidentLevels b=" + std::to_string(b) + " from=" + std::to_string(from));
                // Q and R left on the device, then read back.
                pi::BSResult kq;
                bs.compute_device_qr(0, b, from)->download(kq);
                check(mpz_cmp(kq.Q, w.Q) == 0 && mpz_cmp(kq.R, w.R) == 0,
                      "DeviceResidentQR b=" + std::to_string(b) + " from=" + std::to_string(from));
            }
        }
    }

    // Decimal conversion with device-resident levels (decimal.h) against
    // mpz_get_str: random sizes, 10^k and its neighbours, and numbers whose
    // remainders sit at the ends of their range.
    {
        bool ok = true, est_ok = true;
        std::string err;
        pi::SplitMix64 g(91);
        mpz_t x, p;
        mpz_inits(x, p, NULL);
        auto conv = [&](unsigned threads) {
            char* s = mpz_get_str(nullptr, 10, x);
            pi::DecimalStats st;
            try {
                ok = ok && pi::to_decimal(x, gpu, threads, 4096, &st, nullptr, true) == s;
                est_ok = est_ok && st.min_correction == 0 && st.max_correction <= 1;
            } catch (const std::exception& ex) {
                ok = false;
                err = ex.what();
            }
            std::free(s);
        };
        for (size_t nb : {30000ul, 100000ul, 654321ul, 2000000ul, 20000000ul}) {
            pi::draw_operand(x, g, nb);
            conv(4);
        }
        for (unsigned long k : {8193ul, 65536ul, 100000ul, 131072ul})
            for (int d : {-1, 0, 1}) {
                mpz_ui_pow_ui(x, 10, k);
                if (d < 0) mpz_sub_ui(x, x, 1);
                if (d > 0) mpz_add_ui(x, x, 1);
                conv(3);
            }
        mpz_ui_pow_ui(p, 10, 8192);
        for (unsigned long mlt : {1ul, 7ul, 99999ul})
            for (int d : {-1, 0, 1}) {
                mpz_ui_pow_ui(x, 10, 50000);
                mpz_sub_ui(x, x, 1);
                mpz_tdiv_q(x, x, p);
                mpz_mul_ui(x, x, mlt);
                mpz_tdiv_r(x, x, p);
                mpz_mul(x, x, p);
                mpz_mul(x, x, p);
                if (d < 0) mpz_sub_ui(x, x, 1);
                if (d > 0) mpz_add_ui(x, x, 1);
                conv(2);
            }
        // From a number already in device memory (zero top limbs included),
        // with powers built beforehand for its width, and below the level
        // threshold (converted on the host).
        bool dev_ok = true;
        for (size_t nb : {3000ul, 100000ul, 654321ul, 3000000ul}) {
            pi::draw_operand(x, g, nb);
            char* s = mpz_get_str(nullptr, 10, x);
            const size_t w = std::strlen(s);
            try {
                pi::DecimalPowers pw;
                pi::decimal_powers(pw, w, gpu, 4096, true);
                const size_t n = mpz_size(x) + 5;
                uint64_t* d = nullptr;
                if (gpu_ntt_dev_alloc(n, &d) != 0) throw std::runtime_error("alloc");
                std::vector<uint64_t> h(n, 0);
                std::memcpy(h.data(), mpz_limbs_read(x), mpz_size(x) * 8);
                const uint64_t* hp = h.data();
                size_t nn = n;
                if (gpu_ntt_upload(1, &d, &hp, &nn) != 0) throw std::runtime_error("upload");
                const std::string got = pi::to_decimal_device(d, n, 4, 4096, nullptr, pw);
                (void)gpu_ntt_dev_free(d);
                dev_ok = dev_ok && got == s;
            } catch (const std::exception& ex) {
                dev_ok = false;
                err = ex.what();
            }
            std::free(s);
        }
        mpz_clears(x, p, NULL);
        check(ok, "DecimalDeviceLevelsMatchMpzGetStr" + (err.empty() ? std::string() : " (" + err + ")"));
        check(dev_ok, "DecimalFromDeviceNumberMatchesMpzGetStr" + (err.empty() ? std::string() : " (" + err + ")"));
        check(est_ok, "DecimalDeviceLevelsEstimateWithinOne");
    }

    // Subtrees of 64 terms on the device against the host recursion, at the
    // start of the series (leaf 0), mid-range, near 2^24 and with partial last
    // subtrees (identity terms), then whole device-resident trees built on them.
    {
        bool ok = true;
        std::string err;
        const unsigned long starts[] = {0ul, 1ul, 1000ul, 7000000ul, (1ul << 24) - 300};
        const unsigned long spans[] = {1ul, 63ul, 64ul, 65ul, 200ul, 4096ul};
        for (unsigned long a0 : starts)
            for (unsigned long n : spans) {
                const unsigned long b0 = std::min(a0 + n, 1ul << 24);
                const size_t count = (b0 - a0 + 63) / 64;
                uint64_t* d = nullptr;
                if (gpu_ntt_dev_alloc(count * 3 * PINTT_BS_LEAF_LIMBS, &d) != 0) {
                    ok = false;
                    err = "alloc";
                    continue;
                }
                std::vector<uint32_t> len(3 * count);
                std::vector<int32_t> sg(2 * count);
                std::vector<uint64_t> h(count * 3 * PINTT_BS_LEAF_LIMBS);
                const uint64_t* src = d;
                uint64_t* hp = h.data();
                size_t hn = h.size();
                if (gpu_ntt_bs_leaves(a0, b0, count, d, len.data(), sg.data()) != 0 ||
                    gpu_ntt_download(1, &hp, &src, &hn) != 0) {
                    ok = false;
                    err = gpu_ntt_last_error();
                }
                (void)gpu_ntt_dev_free(d);
                for (size_t i = 0; i < count && ok; ++i) {
                    pi::BSResult w;
                    pi::BinarySplitting::compute_small(a0 + 64 * i, std::min(a0 + 64 * (i + 1), b0), w);
                    mpz_srcptr want[3] = {w.P, w.Q, w.R};
                    for (int v = 0; v < 3; ++v) {
                        mpz_t got;
                        mpz_roinit_n(got, reinterpret_cast<const mp_limb_t*>(h.data() + (3 * i + v) * PINTT_BS_LEAF_LIMBS),
                                     (mp_size_t)len[3 * i + v]);
                        const int sign = v == 1 ? 1 : sg[2 * i + (v == 0 ? 0 : 1)];
                        mpz_t signed_got;
                        mpz_init_set(signed_got, got);
                        if (sign < 0) mpz_neg(signed_got, signed_got);
                        ok = ok && mpz_cmp(signed_got, want[v]) == 0;
                        mpz_clear(signed_got);
                    }
                    if (!ok) err = "subtree " + std::to_string(a0 + 64 * i) + " differs";
                }
            }
        check(ok, "DeviceSubtreesMatchHost" + (err.empty() ? std::string() : " (" + err + ")"));
        ok = true;
        pi::HipNttMultiplier m(1);
        pi::BinarySplitting bs(m, 4);
        bs.set_device_leaves(true);
        for (unsigned long b0 : {3ul, 65ul, 4097ul, 30000ul, 200001ul}) {
            pi::BSResult w, got, kq;
            pi::BinarySplitting::compute_small(0, b0, w);
            try {
                got = bs.compute_device(0, b0, true, 0);
                bs.compute_device_qr(0, b0, 0)->download(kq);
                ok = ok && mpz_cmp(got.P, w.P) == 0 && mpz_cmp(got.Q, w.Q) == 0 && mpz_cmp(got.R, w.R) == 0 &&
                     mpz_cmp(kq.Q, w.Q) == 0 && mpz_cmp(kq.R, w.R) == 0;
            } catch (const std::exception& ex) {
                ok = false;
                err = ex.what();
            }
        }
        check(ok, "DeviceResidentTreeOnDeviceSubtrees" + (err.empty() ? std::string() : " (" + err + ")"));
    }

    // Whole-limb reciprocal on the device: the same integer as its host
    // statement, step for step, for divisors shorter and longer than the
    // precision, 2^(64 L - 1) and all-ones divisors.
    {
        bool ok = true;
        std::string err;
        pi::SplitMix64 g(5);
        mpz_t d, yd, yh;
        mpz_inits(d, yd, yh, NULL);
        const size_t cases[][2] = {{1, 40}, {3, 17}, {50, 400}, {700, 300}, {2000, 5000}, {40000, 40000},
                                   {100000, 60000}};
        for (int kind = 0; kind < 3; ++kind)
            for (const auto& cs : cases) {
                const size_t Ld = cs[0], B = cs[1];
                if (kind == 0) {
                    pi::draw_operand(d, g, 64 * Ld);
                    mpz_setbit(d, 64 * Ld - 1);
                } else if (kind == 1) {
                    mpz_set_ui(d, 0);
                    mpz_setbit(d, 64 * Ld - 1);
                } else {
                    mpz_set_ui(d, 0);
                    mpz_setbit(d, 64 * Ld);
                    mpz_sub_ui(d, d, 1);
                }
                try {
                    pi::device_reciprocal_limbs(yd, d, B);
                    pi::newton_reciprocal_limbs(yh, d, B);
                    ok = ok && mpz_cmp(yd, yh) == 0;
                } catch (const std::exception& ex) {
                    ok = false;
                    err = ex.what();
                }
            }
        for (unsigned long a : {2ul, 10005ul, 999999937ul})
            for (size_t B : {10ul, 300ul, 20000ul, 70000ul}) {
                try {
                    pi::device_inv_sqrt_limbs(yd, a, B);
                    pi::newton_inv_sqrt_limbs(yh, a, B);
                    ok = ok && mpz_cmp(yd, yh) == 0;
                } catch (const std::exception& ex) {
                    ok = false;
                    err = ex.what();
                }
            }
        mpz_clears(d, yd, yh, NULL);
        check(ok, "DeviceReciprocalMatchesHostStatement" + (err.empty() ? std::string() : " (" + err + ")"));
    }
    // The final step on device numbers against the host final step.
    {
        bool ok = true;
        std::string err;
        for (size_t digits : {20000ul, 150001ul}) {
            const size_t prec = digits + 100;
            const unsigned long terms = pi::BinarySplitting::terms_needed(prec);
            pi::BSResult bs;
            pi::BinarySplitting::compute_small(0, terms, bs);
            const size_t bits = (size_t)(prec * 3.3219281) + 64;
            const size_t w = (bits + 64 + 63) / 64 * 64;
            pi::FinalConstants c;
            pi::final_constants(c, digits, w, gpu);
            mpz_t hd, hh, hq;
            mpz_inits(hd, hh, hq, NULL);
            try {
                pi::pi_scaled_newton_device(hd, bs.Q, bs.R, w, c);
                pi::pi_scaled_newton(hh, bs.Q, bs.R, digits, w, gpu, nullptr, &c);
                pi::BinarySplitting bsd(gpu, 4);
                pi::pi_scaled_newton_device(hq, *bsd.compute_device_qr(0, terms), w, c);
                ok = ok && mpz_cmp(hd, hh) == 0 && mpz_cmp(hq, hh) == 0;
            } catch (const std::exception& ex) {
                ok = false;
                err = ex.what();
            }
            mpz_clears(hd, hh, hq, NULL);
        }
        check(ok, "DeviceFinalStepMatchesHost" + (err.empty() ? std::string() : " (" + err + ")"));
    }

    mpz_clears(a, b, r, e, NULL);
    gpu.print_stats();
    std::printf("%d failure(s)\n", failures);
    return failures;
}
