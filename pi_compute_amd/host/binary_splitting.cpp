// binary_splitting.cpp - see binary_splitting.h.
#include "binary_splitting.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <future>
#include <thread>

namespace pi {

BSResult::BSResult() {
    mpz_init(P);
    mpz_init(Q);
    mpz_init(R);
}
BSResult::~BSResult() {
    mpz_clear(P);
    mpz_clear(Q);
    mpz_clear(R);
}
BSResult::BSResult(BSResult&& o) noexcept {
    mpz_init(P);
    mpz_init(Q);
    mpz_init(R);
    mpz_swap(P, o.P);
    mpz_swap(Q, o.Q);
    mpz_swap(R, o.R);
}
BSResult& BSResult::operator=(BSResult&& o) noexcept {
    if (this != &o) {
        mpz_swap(P, o.P);
        mpz_swap(Q, o.Q);
        mpz_swap(R, o.R);
    }
    return *this;
}

BinarySplitting::BinarySplitting(Multiplier& m, unsigned threads)
    : mult_(m), threads_(threads ? threads : std::max(1u, std::thread::hardware_concurrency())) {}

// 14.1816474627... decimal digits per term: log10(640320^3 / 1728) (reference
// src/engine/binary_splitting.cpp:100-102).
unsigned long BinarySplitting::terms_needed(size_t digits) {
    return static_cast<unsigned long>(std::ceil(static_cast<double>(digits) / 14.1816474627)) + 1;
}

namespace {
static_assert(GMP_LIMB_BITS == 64, "64-bit limbs");
void set_u128(mpz_ptr x, unsigned __int128 v) {
    mp_limb_t* d = mpz_limbs_write(x, 2);
    d[0] = (mp_limb_t)v;
    d[1] = (mp_limb_t)(v >> 64);
    mpz_limbs_finish(x, d[1] ? 2 : (d[0] ? 1 : 0));
}
}  // namespace

void BinarySplitting::base_case(unsigned long k, BSResult& r) {
    if (k == 0) {
        mpz_set_ui(r.P, 1);
        mpz_set_ui(r.Q, 1);
        mpz_set_ui(r.R, 13591409);
        return;
    }
    if (k < (1ul << 24)) {
        // (6k-5)(2k-1) < 2^52, P < 72 k^3 < 2^79, Q < 2^72 * 2^54: all fit 128 bits.
        set_u128(r.P, (unsigned __int128)((6 * k - 5) * (2 * k - 1)) * (6 * k - 1));
        mpz_neg(r.P, r.P);
        set_u128(r.Q, (unsigned __int128)(k * k) * k * 10939058860032000ul);
        mpz_mul_ui(r.R, r.P, 545140134ul * k + 13591409ul);
        return;
    }
    mpz_set_ui(r.P, 6 * k - 5);
    mpz_mul_ui(r.P, r.P, 2 * k - 1);
    mpz_mul_ui(r.P, r.P, 6 * k - 1);
    mpz_neg(r.P, r.P);
    mpz_set_ui(r.Q, k);
    mpz_mul_ui(r.Q, r.Q, k);
    mpz_mul_ui(r.Q, r.Q, k);
    mpz_mul_ui(r.Q, r.Q, 10939058860032000ul);  // 640320^3 / 24
    mpz_set_ui(r.R, 545140134);
    mpz_mul_ui(r.R, r.R, k);
    mpz_add_ui(r.R, r.R, 13591409);
    mpz_mul(r.R, r.R, r.P);
}

namespace {
// l covers the terms just before r's: P = Pl Pr, Q = Ql Qr, R = Rl Qr + Pl Rr.
void merge_into(BSResult& l, const BSResult& r, mpz_ptr t) {
    mpz_mul(t, l.R, r.Q);
    mpz_mul(l.R, l.P, r.R);
    mpz_add(l.R, l.R, t);
    mpz_mul(l.P, l.P, r.P);
    mpz_mul(l.Q, l.Q, r.Q);
}

// Per-thread scratch of compute_small, kept across calls so that its GMP
// buffers are allocated once per thread rather than once per merge.
struct SmallScratch {
    std::vector<BSResult> stack;
    std::vector<unsigned long> count;
    mpz_t t;
    SmallScratch() { mpz_init(t); }
    ~SmallScratch() { mpz_clear(t); }
};
}  // namespace

// Bottom-up over [a, b): a stack of partial results whose term counts are
// decreasing powers of two (a binary counter); two equal neighbours merge, and
// the rest merges right to left at the end.  The tree differs from the
// recursive halving only in shape, which P, Q and R do not depend on.
void BinarySplitting::compute_small(unsigned long a, unsigned long b, BSResult& r) {
    if (b <= a) {
        mpz_set_ui(r.P, 1);
        mpz_set_ui(r.Q, 1);
        mpz_set_ui(r.R, 0);
        return;
    }
    thread_local SmallScratch sc;
    size_t top = 0;
    for (unsigned long k = a; k < b; ++k) {
        if (sc.stack.size() <= top) {
            sc.stack.emplace_back();
            sc.count.push_back(0);
        }
        base_case(k, sc.stack[top]);
        sc.count[top] = 1;
        ++top;
        while (top >= 2 && sc.count[top - 2] == sc.count[top - 1]) {
            merge_into(sc.stack[top - 2], sc.stack[top - 1], sc.t);
            sc.count[top - 2] *= 2;
            --top;
        }
    }
    while (top >= 2) {
        merge_into(sc.stack[top - 2], sc.stack[top - 1], sc.t);
        sc.count[top - 2] += sc.count[top - 1];
        --top;
    }
    mpz_swap(r.P, sc.stack[0].P);
    mpz_swap(r.Q, sc.stack[0].Q);
    mpz_swap(r.R, sc.stack[0].R);
}

void BinarySplitting::merge(BSResult& l, BSResult& r, BSResult& out, bool need_p) {
    mpz_t t;
    mpz_init(t);
    mult_.multiply(t, r.Q, l.R);
    mult_.multiply(out.R, l.P, r.R);
    mpz_add(out.R, out.R, t);
    mpz_clear(t);
    if (need_p) mult_.multiply(out.P, l.P, r.P);
    mult_.multiply(out.Q, l.Q, r.Q);
}

std::vector<BSResult> BinarySplitting::leaves(unsigned long a, unsigned long b) {
    return leaves(a, b, threads_, &leaf_s_);
}

// Bottom: subtrees of leaf_span terms on host threads.
std::vector<BSResult> BinarySplitting::leaves(unsigned long a, unsigned long b, unsigned threads,
                                              double* seconds) const {
    const auto t0 = std::chrono::steady_clock::now();
    const unsigned long span = leaf_span_;
    const size_t nleaf = (size_t)((b - a + span - 1) / span);
    std::vector<BSResult> nodes(nleaf);
    struct LeafCtx {
        unsigned long a, b, span;
        std::vector<BSResult>* nodes;
    } lctx{a, b, span, &nodes};
    parallel_for(
        nleaf, threads,
        [](void* p, size_t i) {
            auto* c = static_cast<LeafCtx*>(p);
            unsigned long lo = c->a + i * c->span, hi = std::min(c->b, lo + c->span);
            compute_small(lo, hi, (*c->nodes)[i]);
        },
        &lctx);
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return nodes;
}

// One level: adjacent pairs merge (all products as one multiply_batch); an odd
// last node moves up unchanged.
void BinarySplitting::host_level(std::vector<BSResult>& nodes, bool need_top_p, std::vector<Level>& levels) {
    using Clock = std::chrono::steady_clock;
    auto t0 = Clock::now();
    const size_t pairs = nodes.size() / 2;
    const bool top = nodes.size() == 2;
    std::vector<BSResult> next(pairs + (nodes.size() & 1));
    std::vector<__mpz_struct> tmp(pairs);
    std::vector<MulJob> jobs;
    jobs.reserve(4 * pairs);
    size_t max_limbs = 0;
    for (size_t k = 0; k < pairs; ++k) {
        BSResult& l = nodes[2 * k];
        BSResult& r = nodes[2 * k + 1];
        BSResult& o = next[k];
        mpz_init(&tmp[k]);
        jobs.push_back(MulJob{&tmp[k], r.Q, l.R});
        jobs.push_back(MulJob{o.R, l.P, r.R});
        if (!(top && !need_top_p)) jobs.push_back(MulJob{o.P, l.P, r.P});
        jobs.push_back(MulJob{o.Q, l.Q, r.Q});
        max_limbs = std::max({max_limbs, mpz_size(l.Q), mpz_size(r.Q), mpz_size(l.R), mpz_size(r.R)});
    }
    mult_.multiply_batch(jobs.data(), jobs.size(), threads_);
    for (size_t k = 0; k < pairs; ++k) {
        mpz_add(next[k].R, next[k].R, &tmp[k]);
        mpz_clear(&tmp[k]);
    }
    if (nodes.size() & 1) next.back() = std::move(nodes.back());
    levels.push_back(Level{pairs, jobs.size(), max_limbs,
                            std::chrono::duration<double>(Clock::now() - t0).count()});
    nodes.swap(next);
}

BSResult BinarySplitting::compute(unsigned long a, unsigned long b, bool need_top_p) {
    levels_.clear();
    std::vector<BSResult> nodes = leaves(a, b);
    while (nodes.size() > 1) host_level(nodes, need_top_p, levels_);
    return std::move(nodes[0]);
}

BSResult BinarySplitting::recurse(unsigned long a, unsigned long b, int depth) {
    if (b - a <= leaf_span_) {
        BSResult r;
        compute_small(a, b, r);
        return r;
    }
    unsigned long m = a + (b - a) / 2;
    BSResult l, r;
    if (depth > 0) {
        auto fut = std::async(std::launch::async, [&]() { return recurse(m, b, depth - 1); });
        l = recurse(a, m, depth - 1);
        r = fut.get();
    } else {
        l = recurse(a, m, 0);
        r = recurse(m, b, 0);
    }
    BSResult out;
    merge(l, r, out, true);
    return out;
}

BSResult BinarySplitting::compute_recursive(unsigned long a, unsigned long b) {
    int depth = 0;
    for (unsigned t = threads_; t > 1; t >>= 1) ++depth;
    return recurse(a, b, depth);
}

}  // namespace pi
