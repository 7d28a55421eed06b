// binary_splitting.h - Chudnovsky P/Q/R binary splitting over a Multiplier.
//
// Same recursion as the reference (src/engine/binary_splitting.cpp:137-293):
//   leaf k:  P = -(6k-5)(2k-1)(6k-1), Q = k^3 * 10939058860032000,
//            R = P * (13591409 + 545140134 k);  leaf 0: P = Q = 1, R = 13591409
//   merge:   P = Pl*Pr, Q = Ql*Qr, R = Qr*Rl + Pl*Rr.
// P(a,b), Q(a,b) and R(a,b) do not depend on where the range is split (R is
// the sum over k of P(a,k+1)*(13591409 + 545140134 k)*Q(k+1,b)), so any tree
// shape yields identical integers and identical digits.
//
// compute() evaluates the tree level by level: small subtrees of `leaf_span`
// terms on host threads, then every merge of a level issues its products as
// one Multiplier::multiply_batch (one engine call per level on the GPU).
// compute_recursive() is the depth-first order of the reference, one
// Multiplier::multiply per product.
#pragma once

#include <gmp.h>

#include <cstddef>
#include <memory>
#include <vector>

#include "device_ops.h"
#include "dispatch.h"
#include "multiplier.h"

namespace pi {

struct BSResult {
    mpz_t P, Q, R;
    BSResult();
    ~BSResult();
    BSResult(const BSResult&) = delete;
    BSResult& operator=(const BSResult&) = delete;
    BSResult(BSResult&& o) noexcept;
    BSResult& operator=(BSResult&& o) noexcept;
};

struct DeviceQR;

class BinarySplitting {
public:
    explicit BinarySplitting(Multiplier& m, unsigned threads = 0);

    // [a, b); need_top_p = false skips the final P product (unused by pi).
    BSResult compute(unsigned long a, unsigned long b, bool need_top_p = true);
    BSResult compute_recursive(unsigned long a, unsigned long b);

    // Level-synchronous like compute(), but from the first level whose smallest
    // operand has at least `device_from_limbs` limbs upward, P, Q and R stay in
    // device memory: each level is one device product batch (gpu_ntt_mul_batch_device)
    // and one device add batch for R, host copies only at the switch and at
    // the top.  Every such product runs on the GPU engine.  Throws
    // std::runtime_error if the device fails (the caller can rerun on the host).
    BSResult compute_device(unsigned long a, unsigned long b, bool need_top_p = true, size_t device_from_limbs = 0);

    // The same without the top P, leaving Q and R in device memory (no copy to
    // the host): for a final step that also runs on the device.
    std::unique_ptr<DeviceQR> compute_device_qr(unsigned long a, unsigned long b, size_t device_from_limbs = 0);

    // compute_device cuts [a, b) into this many segments of whole leaf spans
    // (at least 64 leaf spans each): segment i+1 is prepared (host subtrees on
    // threads - 1 host threads, any host levels, upload) while the device
    // reduces segment i, and the segment roots merge on the device.  1: one tree.
    void set_segments(unsigned s) { segments_ = s ? s : 1; }

    // With a dispatcher set, compute_device switches to device-resident values
    // at the first level whose products the dispatcher sends to the GPU (by
    // transform work, at least half), instead of by device_from_limbs.
    void set_switch_dispatcher(const Dispatcher* d) { switch_by_ = d; }

    static unsigned long terms_needed(size_t digits);
    static void base_case(unsigned long k, BSResult& r);
    // Host-only recursion with mpz_mul (the bottom of the tree).
    static void compute_small(unsigned long a, unsigned long b, BSResult& r);

    void set_leaf_span(unsigned long s) { leaf_span_ = s ? s : 1; }

    // compute_device / compute_device_qr: the subtrees of the bottom of the
    // tree computed on the device (gpu_ntt_bs_leaves, 64 terms each; ranges up
    // to 2^24 terms) instead of on host threads.
    void set_device_leaves(bool on) { device_leaves_ = on; }

    struct Level {
        size_t merges, products;
        size_t max_limbs;
        double seconds;
    };
    const std::vector<Level>& levels() const { return levels_; }
    double leaf_seconds() const { return leaf_s_; }      // host subtrees of the last compute* (summed)
    double leaf_wait_seconds() const { return leaf_wait_s_; }  // compute_device: waited for prepared segments
    double upload_seconds() const { return upload_s_; }  // compute_device: host -> device switch

private:
    template <class F>
    BSResult run_segments(unsigned long a, unsigned long b, bool need_top_p, size_t device_from_limbs, F&& at_root);
    void merge(BSResult& left, BSResult& right, BSResult& out, bool need_p);
    std::vector<BSResult> leaves(unsigned long a, unsigned long b);
    std::vector<BSResult> leaves(unsigned long a, unsigned long b, unsigned threads, double* seconds) const;
    // One level on the host (Multiplier::multiply_batch); its record goes to `levels`.
    void host_level(std::vector<BSResult>& nodes, bool need_top_p, std::vector<Level>& levels);
    BSResult recurse(unsigned long a, unsigned long b, int depth);

    Multiplier& mult_;
    unsigned threads_;
    unsigned long leaf_span_ = 256;
    unsigned segments_ = 8;
    bool device_leaves_ = false;
    std::vector<Level> levels_;
    double leaf_s_ = 0, leaf_wait_s_ = 0, upload_s_ = 0;
    const Dispatcher* switch_by_ = nullptr;
    bool level_prefers_device(const std::vector<BSResult>& nodes, bool need_top_p) const;
};

// Q and R of a range kept in device memory (compute_device_qr): magnitudes,
// exact lengths, the sign of R apart (Q > 0).
struct DeviceQR {
    dev::Buffers mem;
    dev::Num Q, R;
    int r_sign = 1;
    void download(BSResult& r) const;  // Q and R into host integers
};

}  // namespace pi
