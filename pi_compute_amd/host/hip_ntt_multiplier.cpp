// hip_ntt_multiplier.cpp - see hip_ntt_multiplier.h.
#include "hip_ntt_multiplier.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <thread>
#include <vector>

#include "pintt_gpu.h"

namespace pi {

namespace {
using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}
// Transform work of a product, for sharing a batch's time out among members.
double work_of(size_t an, size_t bn) {
    double n = (double)(2 * (an + bn));
    return n * std::log2(n + 2.0);
}
}  // namespace

HipNttMultiplier::HipNttMultiplier(size_t threshold, int device)
    : heuristic_(threshold), dispatcher_(&heuristic_), device_(device) {
    if (gpu_ntt_init(device) != PINTT_OK)
        throw std::runtime_error(std::string("HipNttMultiplier: ") + gpu_ntt_last_error());
    // The engine asks the dispatcher for the cut of each product (a learned
    // model prices it; the threshold rule leaves it to the engine).
    set_dispatcher_cuts(true);
}

HipNttMultiplier::~HipNttMultiplier() {
    gpu_ntt_set_chunk_cost(nullptr, nullptr);
    close_log();
}

void HipNttMultiplier::set_dispatcher_cuts(bool on) {
    if (on) gpu_ntt_set_chunk_cost(&HipNttMultiplier::chunk_cost, this);
    else gpu_ntt_set_chunk_cost(nullptr, nullptr);
}

int HipNttMultiplier::chunk_cost(size_t an, size_t bn, double* cost, void* self) {
    return static_cast<HipNttMultiplier*>(self)->dispatcher_->chunk_costs(an, bn, cost) ? 1 : 0;
}

std::string HipNttMultiplier::device_name() const { return "gfx950 device " + std::to_string(device_); }

bool HipNttMultiplier::open_log(const std::string& path) {
    std::lock_guard<std::mutex> lock(log_mu_);
    if (log_) std::fclose(log_);
    log_ = std::fopen(path.c_str(), "a");
    return log_ != nullptr;
}

void HipNttMultiplier::close_log() {
    std::lock_guard<std::mutex> lock(log_mu_);
    if (log_) std::fclose(log_);
    log_ = nullptr;
}

void HipNttMultiplier::log_line(size_t an, size_t bn, size_t batch, bool square, const char* path, double ms) {
    std::lock_guard<std::mutex> lock(log_mu_);
    if (!log_) return;
    std::fprintf(log_, "{\"an\":%zu,\"bn\":%zu,\"batch\":%zu,\"square\":%d,\"path\":\"%s%s\",\"ms\":%.6f}\n", an,
                 bn, batch, square ? 1 : 0, path, log_tag_.c_str(), ms);
}

bool HipNttMultiplier::log_pick(size_t k, size_t n) const {
    if (!log_ || log_limit_ == 0 || n <= log_limit_) return log_ != nullptr;
    // evenly spaced members: k hits a multiple of n / log_limit
    return (k * log_limit_) / n != ((k + 1) * log_limit_) / n;
}

void HipNttMultiplier::host_product(mpz_ptr result, mpz_srcptr a, mpz_srcptr b, size_t batch, bool log) {
    auto t0 = Clock::now();
    mpz_mul(result, a, b);
    double ms = ms_since(t0);
    stats_.cpu_calls.fetch_add(1);
    stats_.cpu_ns.fetch_add((uint64_t)(ms * 1e6));
    if (log) log_line(mpz_size(a), mpz_size(b), batch, a == b, "cpu", ms);
}

// |result| = |a| * |b| (or |a|^2) through the C ABI, then the sign.  The ABI
// forbids output/operand overlap, so an aliased result is built in a temporary.
void HipNttMultiplier::gpu_product(mpz_t result, const mpz_t a, const mpz_t b, bool square) {
    const int sign = mpz_sgn(a) * mpz_sgn(b);
    if (sign == 0) {
        mpz_set_ui(result, 0);
        return;
    }
    const size_t an = mpz_size(a), bn = mpz_size(b);
    if (an + bn > gpu_ntt_max_limbs()) {
        stats_.fallback_calls.fetch_add(1);
        mpz_mul(result, a, b);
        return;
    }
    const bool alias = (result == a) || (result == b);
    mpz_t tmp;
    mpz_ptr dst = result;
    if (alias) {
        mpz_init(tmp);
        dst = tmp;
    }
    auto t0 = Clock::now();
    const mp_limb_t* ap = mpz_limbs_read(a);
    const mp_limb_t* bp = square ? ap : mpz_limbs_read(b);
    mp_limb_t* out = mpz_limbs_write(dst, an + bn);
    static_assert(sizeof(mp_limb_t) == sizeof(uint64_t), "64-bit limbs expected");
    int rc = square ? gpu_ntt_sqr(reinterpret_cast<const uint64_t*>(ap), an, reinterpret_cast<uint64_t*>(out))
                    : gpu_ntt_mul(reinterpret_cast<const uint64_t*>(ap), an,
                                  reinterpret_cast<const uint64_t*>(bp), bn, reinterpret_cast<uint64_t*>(out));
    if (rc != PINTT_OK) {
        stats_.fallback_calls.fetch_add(1);
        mpz_limbs_finish(dst, 0);
        mpz_mul(dst, a, b);
    } else {
        size_t n = an + bn;
        while (n > 0 && out[n - 1] == 0) --n;
        mpz_limbs_finish(dst, static_cast<mp_size_t>(n));
        if (sign < 0) mpz_neg(dst, dst);
        double ms = ms_since(t0);
        stats_.gpu_calls.fetch_add(1);
        stats_.gpu_batches.fetch_add(1);
        stats_.gpu_ns.fetch_add((uint64_t)(ms * 1e6));
        log_line(an, bn, 1, square, "gpu", ms);
    }
    if (alias) {
        mpz_swap(result, tmp);
        mpz_clear(tmp);
    }
}

void HipNttMultiplier::multiply(mpz_t result, const mpz_t a, const mpz_t b) {
    ProductShape sh{mpz_size(a), mpz_size(b), 1, a == b};
    if (dispatcher_->choose(sh) == Path::Cpu) {
        host_product(result, a, b, 1, true);
        return;
    }
    gpu_product(result, a, b, a == b);
}

void HipNttMultiplier::square(mpz_t result, const mpz_t a) {
    ProductShape sh{mpz_size(a), mpz_size(a), 1, true};
    if (dispatcher_->choose(sh) == Path::Cpu) {
        host_product(result, a, a, 1, true);
        return;
    }
    gpu_product(result, a, a, true);
}

void HipNttMultiplier::multiply_batch(const MulJob* jobs, size_t count, unsigned threads) {
    if (count == 0) return;
    std::vector<size_t> dev, host, live;
    std::vector<ProductShape> shapes;
    for (size_t i = 0; i < count; ++i) {
        const MulJob& j = jobs[i];
        if (mpz_sgn(j.a) == 0 || mpz_sgn(j.b) == 0) {
            mpz_set_ui(j.result, 0);
            continue;
        }
        live.push_back(i);
        shapes.push_back(ProductShape{mpz_size(j.a), mpz_size(j.b), count, j.a == j.b});
    }
    std::vector<Path> paths(live.size());
    dispatcher_->choose_batch(shapes.data(), shapes.size(), paths.data());
    for (size_t k = 0; k < live.size(); ++k) {
        bool gpu = paths[k] == Path::Gpu && shapes[k].an + shapes[k].bn <= gpu_ntt_max_limbs();
        (gpu ? dev : host).push_back(live[k]);
    }
    // Host-path products run on host threads while the device batch runs.
    std::thread host_thread([&]() {
        if (host.empty()) return;
        struct Ctx {
            HipNttMultiplier* self;
            const MulJob* jobs;
            const std::vector<size_t>* idx;
            size_t batch;
        } ctx{this, jobs, &host, count};
        std::vector<double> cost(host.size());
        for (size_t k = 0; k < host.size(); ++k)
            cost[k] = (double)mpz_size(jobs[host[k]].a) + (double)mpz_size(jobs[host[k]].b);
        unsigned nt = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
        if (!dev.empty() && nt > 1) --nt;  // one core feeds the device
        parallel_for(
            host.size(), nt,
            [](void* p, size_t k) {
                auto* c = static_cast<Ctx*>(p);
                const MulJob& j = c->jobs[(*c->idx)[k]];
                c->self->host_product(j.result, j.a, j.b, c->batch, c->self->log_pick(k, c->idx->size()));
            },
            &ctx, cost.data());
    });
    if (!dev.empty()) {
        if (!batching_) {
            for (size_t i : dev) gpu_product(jobs[i].result, jobs[i].a, jobs[i].b, jobs[i].a == jobs[i].b);
        } else {
            const size_t n = dev.size();
            std::vector<const uint64_t*> ap(n), bp(n);
            std::vector<uint64_t*> op(n);
            std::vector<size_t> an(n), bn(n);
            for (size_t k = 0; k < n; ++k) {
                const MulJob& j = jobs[dev[k]];
                an[k] = mpz_size(j.a);
                bn[k] = mpz_size(j.b);
                ap[k] = reinterpret_cast<const uint64_t*>(mpz_limbs_read(j.a));
                bp[k] = j.a == j.b ? ap[k] : reinterpret_cast<const uint64_t*>(mpz_limbs_read(j.b));
                op[k] = reinterpret_cast<uint64_t*>(mpz_limbs_write(j.result, an[k] + bn[k]));
            }
            auto t0 = Clock::now();
            int rc = gpu_ntt_mul_batch(n, ap.data(), an.data(), bp.data(), bn.data(), op.data());
            double ms = ms_since(t0);
            if (rc != PINTT_OK) {
                for (size_t k = 0; k < n; ++k) {
                    const MulJob& j = jobs[dev[k]];
                    mpz_limbs_finish(j.result, 0);
                    mpz_mul(j.result, j.a, j.b);
                }
                stats_.fallback_calls.fetch_add(n);
            } else {
                double total_work = 0;
                for (size_t k = 0; k < n; ++k) total_work += work_of(an[k], bn[k]);
                for (size_t k = 0; k < n; ++k) {
                    const MulJob& j = jobs[dev[k]];
                    size_t len = an[k] + bn[k];
                    while (len > 0 && op[k][len - 1] == 0) --len;
                    mpz_limbs_finish(j.result, static_cast<mp_size_t>(len));
                    if (mpz_sgn(j.a) * mpz_sgn(j.b) < 0) mpz_neg(j.result, j.result);
                    if (log_pick(k, n))
                        log_line(an[k], bn[k], n, j.a == j.b, "gpu", ms * work_of(an[k], bn[k]) / total_work);
                }
                stats_.gpu_calls.fetch_add(n);
                stats_.gpu_batches.fetch_add(1);
                stats_.gpu_ns.fetch_add((uint64_t)(ms * 1e6));
            }
        }
    }
    host_thread.join();
}

void HipNttMultiplier::multiply_batch_resident(const MulJob* jobs, size_t count) {
    std::vector<size_t> live;
    for (size_t i = 0; i < count; ++i) {
        if (mpz_sgn(jobs[i].a) == 0 || mpz_sgn(jobs[i].b) == 0) mpz_set_ui(jobs[i].result, 0);
        else live.push_back(i);
    }
    const size_t n = live.size();
    if (n == 0) return;
    size_t total = 0;
    std::vector<size_t> an(n), bn(n);
    for (size_t k = 0; k < n; ++k) {
        const MulJob& j = jobs[live[k]];
        an[k] = mpz_size(j.a);
        bn[k] = mpz_size(j.b);
        total += an[k] + (j.a == j.b ? 0 : bn[k]) + an[k] + bn[k];
    }
    uint64_t* base = nullptr;
    bool ok = an.size() && total && gpu_ntt_dev_alloc(total, &base) == PINTT_OK;
    std::vector<const uint64_t*> da(n), db(n), hs;
    std::vector<uint64_t*> dout(n), dd;
    std::vector<size_t> len;
    if (ok) {
        size_t off = 0;
        for (size_t k = 0; k < n; ++k) {
            const MulJob& j = jobs[live[k]];
            uint64_t* pa = base + off;
            off += an[k];
            dd.push_back(pa);
            hs.push_back(reinterpret_cast<const uint64_t*>(mpz_limbs_read(j.a)));
            len.push_back(an[k]);
            uint64_t* pb = pa;
            if (j.a != j.b) {
                pb = base + off;
                off += bn[k];
                dd.push_back(pb);
                hs.push_back(reinterpret_cast<const uint64_t*>(mpz_limbs_read(j.b)));
                len.push_back(bn[k]);
            }
            da[k] = pa;
            db[k] = pb;
            dout[k] = base + off;
            off += an[k] + bn[k];
        }
        ok = gpu_ntt_upload(dd.size(), dd.data(), hs.data(), len.data()) == PINTT_OK;
    }
    double ms = 0;
    if (ok) {
        auto t0 = Clock::now();
        ok = gpu_ntt_mul_batch_device(n, da.data(), an.data(), db.data(), bn.data(), dout.data(), nullptr) == PINTT_OK &&
             gpu_ntt_synchronize() == PINTT_OK;
        ms = ms_since(t0);
    }
    if (ok) {
        std::vector<uint64_t*> hd(n);
        std::vector<const uint64_t*> sd(n);
        std::vector<size_t> ln(n);
        for (size_t k = 0; k < n; ++k) {
            hd[k] = reinterpret_cast<uint64_t*>(mpz_limbs_write(jobs[live[k]].result, an[k] + bn[k]));
            sd[k] = dout[k];
            ln[k] = an[k] + bn[k];
        }
        ok = gpu_ntt_download(n, hd.data(), sd.data(), ln.data()) == PINTT_OK;
        for (size_t k = 0; k < n; ++k) {
            const MulJob& j = jobs[live[k]];
            size_t l = ok ? ln[k] : 0;
            while (l && !hd[k][l - 1]) --l;
            mpz_limbs_finish(j.result, (mp_size_t)l);
            if (!ok) mpz_mul(j.result, j.a, j.b);
            else if (mpz_sgn(j.a) * mpz_sgn(j.b) < 0) mpz_neg(j.result, j.result);
        }
    } else {
        for (size_t k = 0; k < n; ++k) mpz_mul(jobs[live[k]].result, jobs[live[k]].a, jobs[live[k]].b);
    }
    if (base) (void)gpu_ntt_dev_free(base);
    if (!ok) {
        stats_.fallback_calls.fetch_add(n);
        return;
    }
    double total_work = 0;
    for (size_t k = 0; k < n; ++k) total_work += work_of(an[k], bn[k]);
    for (size_t k = 0; k < n; ++k)
        if (log_pick(k, n)) log_line(an[k], bn[k], n, jobs[live[k]].a == jobs[live[k]].b, "gpu_dev",
                                     ms * work_of(an[k], bn[k]) / total_work);
    stats_.gpu_calls.fetch_add(n);
    stats_.gpu_batches.fetch_add(1);
}

void HipNttMultiplier::reset_stats() {
    stats_.gpu_calls = 0;
    stats_.gpu_batches = 0;
    stats_.cpu_calls = 0;
    stats_.fallback_calls = 0;
    stats_.gpu_ns = 0;
    stats_.cpu_ns = 0;
}

void HipNttMultiplier::print_stats() const {
    const uint64_t g = stats_.gpu_calls.load();
    std::printf("  HIP NTT multiplier (%s dispatch): %llu GPU products in %llu engine calls (%.3f s), "
                "%llu host products (%.3f s summed over threads), %llu fallbacks\n",
                dispatcher_->name(), (unsigned long long)g, (unsigned long long)stats_.gpu_batches.load(),
                stats_.gpu_ns.load() / 1e9, (unsigned long long)stats_.cpu_calls.load(),
                stats_.cpu_ns.load() / 1e9, (unsigned long long)stats_.fallback_calls.load());
}

}  // namespace pi
