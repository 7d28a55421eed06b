// hip_ntt_multiplier.h - Multiplier backed by the gfx950 NTT engine.
//
// Drop-in beside GmpMultiplier.  Semantics follow the reference GPU wrappers
// (src/arithmetic/int_ntt_multiplier.cpp:115-159,
//  src/arithmetic/gpu_ntt_multiplier.cpp:144-210):
//   * the dispatcher (default: smaller operand below `threshold` limbs) sends
//     a product to mpz_mul on the host;
//   * sign = sgn(a) * sgn(b), zero operand -> result 0;
//   * engine refusal (too large, device error) -> mpz_mul, product still exact.
// Differences: operands cross as raw GMP limbs (no digit vectors), squaring
// takes one forward transform, result may alias a or b, and multiply_batch
// sends a whole set of products to the device in one engine call while the
// host-path products of the set run on host threads at the same time.
#pragma once

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

#include "dispatch.h"
#include "multiplier.h"

namespace pi {

class HipNttMultiplier : public Multiplier {
public:
    // threshold: minimum limbs of the smaller operand for the GPU path
    // (reference default 10000, src/main.cpp:63).  device: HIP device index.
    // Throws std::runtime_error when no usable device exists, like the
    // reference constructors (src/arithmetic/int_ntt_multiplier.cpp:26-31).
    explicit HipNttMultiplier(size_t threshold = 10000, int device = 0);
    ~HipNttMultiplier() override;

    void multiply(mpz_t result, const mpz_t a, const mpz_t b) override;
    void square(mpz_t result, const mpz_t a) override;
    void multiply_batch(const MulJob* jobs, size_t count, unsigned threads = 0) override;

    size_t threshold() const { return heuristic_.threshold(); }
    void set_threshold(size_t t) { heuristic_.set_threshold(t); }
    // Replace the path choice (not owned; nullptr restores the threshold rule).
    void set_dispatcher(const Dispatcher* d) { dispatcher_ = d ? d : &heuristic_; }
    const Dispatcher& dispatcher() const { return *dispatcher_; }
    // When true (the default) the engine asks the dispatcher for the cost of
    // each chunk count of a product (a learned model prices the cuts); when
    // false the engine cuts by its own tuned per-length costs.
    void set_dispatcher_cuts(bool on);
    // Every product on the GPU with operands uploaded first and results
    // downloaded after; only the device batch itself is timed and logged (path
    // "gpu_dev"): the cost of products whose operands already live on the
    // device (binary splitting's device-resident levels).
    void multiply_batch_resident(const MulJob* jobs, size_t count);
    // When false, multiply_batch issues its device products one engine call
    // each (the per-call mode the batching is measured against).
    void set_batching(bool on) { batching_ = on; }
    std::string device_name() const;

    // Per-product timing log, one JSON object per line:
    //   {"an":..,"bn":..,"batch":..,"square":0|1,"path":"gpu"|"cpu","ms":..}
    // For a device batch, "ms" is the batch wall time shared out in proportion
    // to each product's transform work.  These lines train the learned dispatcher.
    bool open_log(const std::string& path);
    void close_log();
    // At most this many evenly spaced lines per path and batch (0: all).
    void set_log_limit(size_t n) { log_limit_ = n; }
    // Appended to the path name of logged lines (calibration of forced cuts).
    void set_log_tag(const std::string& tag) { log_tag_ = tag; }

    struct Stats {
        std::atomic<uint64_t> gpu_calls{0};       // products on the device
        std::atomic<uint64_t> gpu_batches{0};     // engine calls
        std::atomic<uint64_t> cpu_calls{0};       // products sent to the host by the dispatcher
        std::atomic<uint64_t> fallback_calls{0};  // engine refused or failed
        std::atomic<uint64_t> gpu_ns{0};
        std::atomic<uint64_t> cpu_ns{0};
    };
    const Stats& stats() const { return stats_; }
    void reset_stats();
    void print_stats() const;

private:
    void gpu_product(mpz_t result, const mpz_t a, const mpz_t b, bool square);
    void host_product(mpz_ptr result, mpz_srcptr a, mpz_srcptr b, size_t batch, bool log);
    bool log_pick(size_t k, size_t n) const;
    void log_line(size_t an, size_t bn, size_t batch, bool square, const char* path, double ms);
    static int chunk_cost(size_t an, size_t bn, double* cost, void* self);

    ThresholdDispatcher heuristic_;
    const Dispatcher* dispatcher_;
    int device_;
    bool batching_ = true;
    std::FILE* log_ = nullptr;
    size_t log_limit_ = 0;
    std::string log_tag_;
    std::mutex log_mu_;
    mutable Stats stats_;
};

}  // namespace pi
