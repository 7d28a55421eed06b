// mlp_dispatch.h - learned CPU/GPU dispatch: a small FP32 network predicts the
// log run time of a product on each path; the cheaper path wins.
//
// Replaces the reference's fixed crossover (src/arithmetic/int_ntt_multiplier.cpp:56-59)
// while keeping it as the fallback: shapes outside the range the model was
// trained on, and any model that fails to load, go to the threshold rule.
// Either path returns the exact product; only wall time depends on the choice.
//
// Inputs (standardised with the model's mean/scale):
//   log2(larger operand limbs), log2(smaller operand limbs), log2(batch), square,
//   and for m = 1 .. 4 the padding of the transform the product uses when its
//   longer operand is cut into m chunks: log2(2^k / (2 h + 2 bn - 1)) for
//   chunks of h limbs and transform length 2^k (gpu_ntt_chunk_plan), -1 where
//   that cut is not allowed (squares are never cut).  Run time steps up each
//   time a product crosses a power-of-two transform length; the padding
//   inputs let a smooth network place those steps exactly.
// Outputs: predicted ln(milliseconds) per product, output 0 = host GMP,
// output 1 = GPU engine with host operands (a device batch's time, staging
// included, shared out by transform work), output 2 (models with three) =
// GPU engine on device-resident operands (no staging); a resident shape is
// priced with output 2 when the model has it.  Outputs 3 .. 6 (models with
// seven) = one device-resident product with the longer operand cut into
// m = 1 .. 4 chunks: the engine asks for them (chunk_costs) and takes the
// cheapest cut it allows, in place of its tuned per-length costs.
// A single product goes to the path with the lower prediction.  A batch is
// split so that its device share and its host share (run on `cpu_threads`
// threads at the same time) finish together as early as the predictions allow.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "dispatch.h"
#include "pintt_mlp.h"

namespace pi {

using MlpFeatures = std::array<float, PINTT_MLP_INPUTS>;

struct MlpModel {
    int outputs = 2;
    MlpFeatures mean{}, scale{}, lo{}, hi{};  // feature standardisation and trained range
    std::vector<float> params;                // layout of include/pintt_mlp.h

    static MlpFeatures features(const ProductShape& s);
    bool in_range(const MlpFeatures& f) const;
    MlpFeatures standardise(const MlpFeatures& f) const;
    // Host forward pass, same arithmetic order as the device kernel.
    void forward(const float* x, float* y) const;

    bool load(const std::string& path, std::string* err = nullptr);
    bool save(const std::string& path) const;
};

class MlpDispatcher : public Dispatcher {
public:
    // Throws std::runtime_error if the model cannot be loaded.
    MlpDispatcher(const std::string& model_path, const Dispatcher* fallback, unsigned cpu_threads = 0);
    MlpDispatcher(MlpModel model, const Dispatcher* fallback, unsigned cpu_threads = 0);

    Path choose(const ProductShape& s) const override;
    void choose_batch(const ProductShape* s, size_t n, Path* out) const override;
    bool chunk_costs(size_t an, size_t bn, double* cost) const override;
    const char* name() const override { return "mlp"; }

    const MlpModel& model() const { return model_; }
    // Batches of at least this many products are scored by the device kernel.
    void set_device_batch(size_t n) { device_batch_ = n; }

private:
    Path decide(const ProductShape& s, const float* y) const;

    MlpModel model_;
    const Dispatcher* fallback_;
    float log_threads_;
    size_t device_batch_ = 64;
};

}  // namespace pi
