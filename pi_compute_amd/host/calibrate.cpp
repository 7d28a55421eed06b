// calibrate.cpp - pi_calibrate: logs per-product timings of both paths for
// training the learned dispatcher.  Every line comes from HipNttMultiplier's
// own timing log (open_log), with the path forced by a FixedDispatcher.
//
//   size sweep      |a| = 2^k bits, |b| = 2^k/3 bits and a^2, k = kmin..kmax,
//                   single calls, reps per size (the crossover region)
//   uniform batches B products of ~2^lg limbs (b from 1/4 to 1 of a, every
//                   third a square), B = 2..4096, one multiply_batch each,
//                   and once more with device-resident operands
//   tree level      4096 products, lognormal bit lengths (median 2^18,
//                   sigma 0.6, clipped to [2^12, 2^24]), every third a square
//   chunk cuts      single device-resident products of 2^10 .. 2^22 limbs by
//                   1/1 .. 1/16 as long, with the longer operand cut into each
//                   allowed m = 1 .. 4 chunks (forced, path "gpu_dev_m<m>")
//                   and as the engine picks it (path "gpu_dev")
//
//   pi_calibrate --out FILE [--kmin 8] [--kmax 28] [--reps 10] [--limit 48]
//                [--threads T] [--cpu-budget SECONDS] [--quick]
//                [--cut-step OCTAVES] [--cut-reps N] [--cuts-only]
#include <gmp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "dispatch.h"
#include "hip_ntt_multiplier.h"
#include "operands.h"
#include "pintt_gpu.h"

using namespace pi;
using Clock = std::chrono::steady_clock;

namespace {

struct Batch {
    std::vector<__mpz_struct> a, b, r;
    std::vector<bool> sq;
    explicit Batch(size_t n) : a(n), b(n), r(n), sq(n) {
        for (size_t i = 0; i < n; ++i) {
            mpz_init(&a[i]);
            mpz_init(&b[i]);
            mpz_init(&r[i]);
        }
    }
    ~Batch() {
        for (size_t i = 0; i < a.size(); ++i) {
            mpz_clear(&a[i]);
            mpz_clear(&b[i]);
            mpz_clear(&r[i]);
        }
    }
    std::vector<MulJob> jobs() {
        std::vector<MulJob> j(a.size());
        for (size_t i = 0; i < a.size(); ++i) j[i] = MulJob{&r[i], &a[i], sq[i] ? &a[i] : &b[i]};
        return j;
    }
    double cpu_estimate_s() const {  // single-thread GMP, rough n^1.5 model in limbs
        double t = 0;
        for (size_t i = 0; i < a.size(); ++i) {
            double n = (double)mpz_size(&a[i]) + (double)(sq[i] ? mpz_size(&a[i]) : mpz_size(&b[i]));
            t += 2e-9 * std::pow(n, 1.45) + 2e-7;
        }
        return t;
    }
};

double seconds(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

}  // namespace

int main(int argc, char** argv) {
    std::string out;
    int kmin = 8, kmax = 28, reps = 10;
    size_t limit = 48;
    unsigned threads = 0;
    double cpu_budget = 3.0;  // seconds of estimated host work per configuration
    bool quick = false, cuts_only = false;
    double cut_step = 0.125;
    int cut_reps = 3;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() { return std::string(i + 1 < argc ? argv[++i] : ""); };
        if (a == "--out") out = val();
        else if (a == "--kmin") kmin = std::stoi(val());
        else if (a == "--kmax") kmax = std::stoi(val());
        else if (a == "--reps") reps = std::stoi(val());
        else if (a == "--limit") limit = std::stoul(val());
        else if (a == "--threads") threads = std::stoul(val());
        else if (a == "--cpu-budget") cpu_budget = std::stod(val());
        else if (a == "--quick") quick = true;
        else if (a == "--cut-step") cut_step = std::stod(val());
        else if (a == "--cut-reps") cut_reps = std::stoi(val());
        else if (a == "--cuts-only") cuts_only = true;
        else {
            std::cerr << "unknown option " << a << std::endl;
            return 2;
        }
    }
    if (out.empty()) {
        std::cerr << "--out FILE is required" << std::endl;
        return 2;
    }
    HipNttMultiplier m(0);
    FixedDispatcher cpu(Path::Cpu), gpu(Path::Gpu);
    if (!m.open_log(out)) {
        std::cerr << "cannot open " << out << std::endl;
        return 1;
    }
    m.set_log_limit(limit);
    const Dispatcher* paths[2] = {&cpu, &gpu};

    // Size sweep, single calls.
    auto t0 = Clock::now();
    if (cuts_only) kmax = kmin - 1;
    for (int k = kmin; k <= kmax; ++k) {
        SplitMix64 g(k);
        mpz_t a, b, r;
        mpz_inits(a, b, r, NULL);
        draw_operand(a, g, (size_t)1 << k);
        draw_operand(b, g, ((size_t)1 << k) / 3);
        for (const Dispatcher* d : paths) {
            m.set_dispatcher(d);
            int n = reps;
            if (d == &cpu) n = k >= 27 ? 1 : k >= 25 ? 2 : k >= 23 ? 4 : reps;
            if (quick) n = std::min(n, 2);
            for (int rep = 0; rep < n; ++rep) {
                m.multiply(r, a, b);
                m.square(r, a);
            }
        }
        mpz_clears(a, b, r, NULL);
        std::printf("size sweep k=%d done (%.1f s)\n", k, seconds(t0));
        std::fflush(stdout);
    }

    // Uniform batches.
    for (int lg = 4; lg <= (cuts_only ? 3 : 20); ++lg) {
        for (size_t B = 2; B <= 4096; B *= (quick ? 8 : 2)) {
            if ((B << lg) > ((size_t)1 << 24)) break;
            SplitMix64 g(1000 * lg + B);
            Batch bt(B);
            for (size_t i = 0; i < B; ++i) {
                double la = lg + g.uniform() - 0.5;
                size_t abits = std::max<size_t>(64, (size_t)(64.0 * std::exp2(la)));
                size_t bbits = std::max<size_t>(64, (size_t)(abits * (0.25 + 0.75 * g.uniform())));
                bt.sq[i] = (i % 3) == 2;
                draw_operand(&bt.a[i], g, abits);
                if (!bt.sq[i]) draw_operand(&bt.b[i], g, bbits);
            }
            auto jobs = bt.jobs();
            m.set_dispatcher(&gpu);
            m.multiply_batch(jobs.data(), jobs.size(), threads);
            m.multiply_batch_resident(jobs.data(), jobs.size());
            if (bt.cpu_estimate_s() / std::max(1u, threads ? threads : 16u) < cpu_budget) {
                m.set_dispatcher(&cpu);
                m.multiply_batch(jobs.data(), jobs.size(), threads);
            }
        }
        std::printf("uniform batches lg=%d done (%.1f s)\n", lg, seconds(t0));
        std::fflush(stdout);
    }

    // Tree-level replay.
    if (!cuts_only) {
        const size_t n = 4096;
        SplitMix64 g(7);
        std::vector<double> u(4 * n);
        for (auto& x : u) x = g.uniform();
        Batch bt(n);
        for (size_t i = 0; i < n; ++i) {
            auto bits = [&](double u1, double u2) {
                double z = std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(2.0 * M_PI * u2);
                double v = std::nearbyint(262144.0 * std::exp(0.6 * z));
                return (size_t)std::min(16777216.0, std::max(4096.0, v));
            };
            size_t ba = bits(u[4 * i], u[4 * i + 1]), bb = bits(u[4 * i + 2], u[4 * i + 3]);
            bt.sq[i] = (i % 3) == 2;
            draw_operand(&bt.a[i], g, ba);
            if (!bt.sq[i]) draw_operand(&bt.b[i], g, bb);
        }
        auto jobs = bt.jobs();
        m.set_log_limit(limit * 8);
        for (const Dispatcher* d : paths) {
            m.set_dispatcher(d);
            m.multiply_batch(jobs.data(), jobs.size(), threads);
        }
        m.multiply_batch_resident(jobs.data(), jobs.size());
        std::printf("tree level done (%.1f s)\n", seconds(t0));
    }
    // Chunk cuts: every allowed cut of single device-resident products.
    {
        m.set_log_limit(0);
        m.set_dispatcher(&gpu);
        const double ratios[] = {1, 1.5, 2, 3, 4, 6, 8, 12, 16};
        int shapes = 0;
        for (double la = 10; la <= 22.51; la += quick ? 2.0 : cut_step) {
            for (double ratio : ratios) {
                const size_t an = (size_t)std::llround(std::exp2(la));
                const size_t bn = std::max<size_t>(1, (size_t)std::llround((double)an / ratio));
                if (an + bn > gpu_ntt_max_limbs()) continue;
                SplitMix64 g(7000 + 64 * shapes);
                Batch bt(1);
                draw_operand(&bt.a[0], g, 64 * an);
                draw_operand(&bt.b[0], g, 64 * bn);
                auto jobs = bt.jobs();
                m.multiply_batch_resident(jobs.data(), 1);  // warm-up: plans and buffers for this shape
                for (int c = 1; c <= 4; ++c) {
                    int log2n = 0;
                    if (!gpu_ntt_chunk_plan(an, bn, c, nullptr, &log2n)) continue;
                    gpu_ntt_set_split(c == 1 ? 0 : c);
                    m.set_log_tag("_m" + std::to_string(c));
                    for (int rep = 0; rep < (quick ? 1 : cut_reps); ++rep) m.multiply_batch_resident(jobs.data(), 1);
                }
                gpu_ntt_set_split(1);
                m.set_log_tag("");
                for (int rep = 0; rep < (quick ? 1 : cut_reps); ++rep) m.multiply_batch_resident(jobs.data(), 1);
                ++shapes;
            }
            std::printf("chunk cuts up to 2^%.1f limbs done (%.1f s)\n", la, seconds(t0));
            std::fflush(stdout);
        }
    }
    m.close_log();
    m.print_stats();
    return 0;
}
