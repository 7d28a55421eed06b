// dispatch.h - which path (host GMP or the GPU engine) runs a product.
//
// The reference decides with a fixed limb threshold on the smaller operand
// (src/arithmetic/int_ntt_multiplier.cpp:56-59, default 10000 limbs,
// src/main.cpp:63); ThresholdDispatcher keeps that rule.  A learned cost model
// (mlp_dispatch.h) implements the same interface.  Any choice yields the exact
// product; only wall time differs.
#pragma once

#include <cstddef>

namespace pi {

enum class Path { Cpu = 0, Gpu = 1 };

struct ProductShape {
    size_t an, bn;  // limbs
    size_t batch;   // products issued together with this one
    bool square;
    bool resident = false;  // operands and result can stay in device memory (no host staging)
};

class Dispatcher {
public:
    virtual ~Dispatcher() = default;
    virtual Path choose(const ProductShape& s) const = 0;
    // All products of one batch at once (a learned model evaluates them in
    // one kernel launch); the default asks choose() for each.
    virtual void choose_batch(const ProductShape* s, size_t n, Path* out) const {
        for (size_t i = 0; i < n; ++i) out[i] = choose(s[i]);
    }
    // Predicted costs (any monotone scale) of a device product of an x bn
    // limbs (an >= bn) with the longer operand cut into m = 1 .. 4 chunks, in
    // cost[m - 1]; false leaves the cut to the engine's own tuned costs.
    virtual bool chunk_costs(size_t an, size_t bn, double* cost) const {
        (void)an, (void)bn, (void)cost;
        return false;
    }
    virtual const char* name() const = 0;
};

class ThresholdDispatcher : public Dispatcher {
public:
    explicit ThresholdDispatcher(size_t threshold = 10000) : threshold_(threshold) {}
    Path choose(const ProductShape& s) const override {
        size_t m = s.an < s.bn ? s.an : s.bn;
        return m >= threshold_ ? Path::Gpu : Path::Cpu;
    }
    const char* name() const override { return "threshold"; }
    size_t threshold() const { return threshold_; }
    void set_threshold(size_t t) { threshold_ = t; }

private:
    size_t threshold_;
};

// Always one path (the GMP-only and GPU-only runs).
class FixedDispatcher : public Dispatcher {
public:
    explicit FixedDispatcher(Path p) : p_(p) {}
    Path choose(const ProductShape&) const override { return p_; }
    const char* name() const override { return p_ == Path::Gpu ? "gpu" : "cpu"; }

private:
    Path p_;
};

}  // namespace pi
