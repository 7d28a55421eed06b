// decimal.cpp - see decimal.h.
#include "decimal.h"

#include <algorithm>
#include <chrono>

#include <cstring>
#include <future>
#include <memory>
#include <thread>
#include <vector>

#include "newton.h"

namespace pi {
namespace {

constexpr size_t kLeafDigits = 4096;

struct Powers {
    std::vector<__mpz_struct> p;  // p[i] = 10^(2^i)
    explicit Powers(size_t max_digits) {
        size_t d = 1;
        while (d < max_digits) {
            p.emplace_back();
            mpz_init(&p.back());
            if (p.size() == 1)
                mpz_set_ui(&p.back(), 10);
            else
                mpz_mul(&p.back(), &p[p.size() - 2], &p[p.size() - 2]);
            d <<= 1;
        }
    }
    ~Powers() {
        for (auto& x : p) mpz_clear(&x);
    }
};

void leaf(mpz_srcptr n, char* slot, size_t width) {
    size_t len = mpz_sizeinbase(n, 10) + 2;
    std::vector<char> buf(len);
    mpz_get_str(buf.data(), 10, n);
    len = strlen(buf.data());
    if (len == 1 && buf[0] == '0') len = 0;
    memset(slot, '0', width - len);
    memcpy(slot + width - len, buf.data(), len);
}

// n < 10^width; writes exactly `width` characters.
void convert(mpz_srcptr n, char* slot, size_t width, const Powers& pw, int par_depth) {
    if (width <= kLeafDigits) {
        leaf(n, slot, width);
        return;
    }
    int lvl = 0;
    while (lvl + 1 < (int)pw.p.size() && ((size_t)2 << lvl) < width) ++lvl;
    const size_t low = (size_t)1 << lvl;  // low < width
    mpz_t q, r;
    mpz_init(q);
    mpz_init(r);
    mpz_tdiv_qr(q, r, n, &pw.p[lvl]);
    if (par_depth > 0) {
        auto fut = std::async(std::launch::async, [&]() { convert(r, slot + width - low, low, pw, par_depth - 1); });
        convert(q, slot, width - low, pw, par_depth - 1);
        fut.get();
    } else {
        convert(q, slot, width - low, pw, 0);
        convert(r, slot + width - low, low, pw, 0);
    }
    mpz_clear(q);
    mpz_clear(r);
}

}  // namespace

std::string to_decimal(mpz_srcptr n, unsigned threads) {
    if (mpz_sgn(n) == 0) return "0";
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    // sizeinbase may overshoot by one; convert at that width and strip.
    const size_t width = mpz_sizeinbase(n, 10);
    std::string out(width, '0');
    mpz_t a;
    mpz_init(a);
    mpz_abs(a, n);
    if (width <= kLeafDigits) {
        leaf(a, &out[0], width);
    } else {
        Powers pw(width);
        int depth = 0;
        for (unsigned t = threads; t > 1; t >>= 1) ++depth;
        convert(a, &out[0], width, pw, depth + 1);
    }
    mpz_clear(a);
    size_t z = out.find_first_not_of('0');
    if (z != 0) out.erase(0, z);
    return out;
}

namespace {

// A node of the level-by-level conversion: the value is the `size` limbs at
// `limbs`, which `own` holds (or the caller's number, for the root), so a
// quotient taken as the top limbs of a product needs no copy.
struct Node {
    mpz_t own;
    const mp_limb_t* limbs = nullptr;
    size_t size = 0;
    __mpz_struct view;
    size_t offset, width;  // digits [offset, offset + width) of the output; value < 10^width
    Node() { mpz_init(own); }
    ~Node() { mpz_clear(own); }
    Node(const Node&) = delete;
    Node& operator=(const Node&) = delete;
    void take(mpz_t x) {  // the value becomes x (x is left with the old storage)
        mpz_swap(own, x);
        limbs = mpz_limbs_read(own);
        size = mpz_size(own);
    }
    void take_top(mpz_t x, size_t skip) {  // the value becomes x >> (GMP_NUMB_BITS * skip), x >= 0
        mpz_swap(own, x);
        const size_t n = mpz_size(own);
        limbs = mpz_limbs_read(own) + std::min(skip, n);
        size = n > skip ? n - skip : 0;
    }
    mpz_srcptr x() { return mpz_roinit_n(&view, limbs, (mp_size_t)size); }
};

// One split by P = 10^(2^j) of bit length np: the quotient estimate is
// ((x >> sx) * inv) >> (e2 - sx) with inv ~ 2^e2 / P.  Both shifts are whole
// limbs, so the shifted operands are views of the numbers they come from.
// With x < P^2, |inv - 2^e2/P| <= 4 and 2^sx < P / 2^63 the estimate is
// within 6 of floor(x / P).
struct SplitShifts {
    size_t sx, e2;
};
SplitShifts split_shifts(size_t np) {
    const size_t L = GMP_NUMB_BITS;
    return SplitShifts{(np > 64 ? np - 64 : 0) / L * L, (2 * np + L - 1) / L * L};
}

}  // namespace

DecimalPowers::~DecimalPowers() {
    for (auto& x : pw) mpz_clear(&x);
    for (auto& x : inv) mpz_clear(&x);
}

void decimal_powers(DecimalPowers& d, size_t width, Multiplier& m, size_t leaf_digits) {
    using Clock = std::chrono::steady_clock;
    auto t = Clock::now();
    leaf_digits = std::max<size_t>(leaf_digits, kLeafDigits);
    d.width = width;
    d.leaf_digits = leaf_digits;
    int k = 0;
    while (((size_t)2 << k) < width) ++k;  // 2^k < width <= 2^(k+1)
    // Powers 10^(2^j), j <= k, by squaring through m.
    d.pw.resize(k + 1);
    d.inv.resize(k + 1);
    for (int j = 0; j <= k; ++j) {
        mpz_init(&d.pw[j]);
        mpz_init(&d.inv[j]);
        if (j == 0)
            mpz_set_ui(&d.pw[j], 10);
        else
            m.square(&d.pw[j], &d.pw[j - 1]);
    }
    d.powers_s = std::chrono::duration<double>(Clock::now() - t).count();
    t = Clock::now();
    for (int j = k; j >= 0 && ((size_t)1 << j) >= leaf_digits; --j) {
        const size_t np = mpz_sizeinbase(&d.pw[j], 2);
        newton_reciprocal(&d.inv[j], &d.pw[j], split_shifts(np).e2 - np, m);  // ~ 2^e2 / P, within 4
    }
    d.reciprocals_s = std::chrono::duration<double>(Clock::now() - t).count();
}

std::string to_decimal(mpz_srcptr n, Multiplier& m, unsigned threads, size_t leaf_digits, DecimalStats* stats,
                       const DecimalPowers* pre) {
    using Clock = std::chrono::steady_clock;
    DecimalStats st;
    auto lap = [](Clock::time_point& t) {
        const auto now = Clock::now();
        const double s = std::chrono::duration<double>(now - t).count();
        t = now;
        return s;
    };
    auto t = Clock::now();
    if (mpz_sgn(n) == 0) return "0";
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    leaf_digits = std::max<size_t>(leaf_digits, kLeafDigits);
    const size_t width = mpz_sizeinbase(n, 10);
    if (width <= 2 * leaf_digits) return to_decimal(n, threads);
    std::string out(width, '0');

    // The powers depend on the width only through k (2^k < width <= 2^(k+1)).
    int kw = 0;
    while (((size_t)2 << kw) < width) ++kw;
    DecimalPowers own;
    if (!pre || (int)pre->pw.size() != kw + 1 || pre->leaf_digits != leaf_digits) {
        decimal_powers(own, width, m, leaf_digits);
        pre = &own;
    }
    const int k = (int)pre->pw.size() - 1;
    st.powers_s = pre->powers_s;
    st.reciprocals_s = pre->reciprocals_s;
    lap(t);
    std::vector<std::unique_ptr<Node>> nodes;
    nodes.emplace_back(new Node);
    nodes[0]->limbs = mpz_limbs_read(n);  // |n|
    nodes[0]->size = mpz_size(n);
    nodes[0]->offset = 0;
    nodes[0]->width = width;

    for (int j = k; j >= 0 && ((size_t)1 << j) >= leaf_digits; --j) {
        const size_t low = (size_t)1 << j;
        std::vector<size_t> split;
        for (size_t i = 0; i < nodes.size(); ++i)
            if (nodes[i]->width > low) split.push_back(i);
        if (split.empty()) continue;
        mpz_srcptr P = &pre->pw[j];
        mpz_srcptr inv = &pre->inv[j];
        const SplitShifts sh = split_shifts(mpz_sizeinbase(P, 2));
        const size_t qskip = (sh.e2 - sh.sx) / GMP_NUMB_BITS;
        lap(t);
        const size_t ns = split.size();
        std::vector<__mpz_struct> hi(ns), qv(ns), qfull(ns), qp(ns);
        std::vector<MulJob> jobs(ns);
        for (size_t s = 0; s < ns; ++s) {
            const Node& nd = *nodes[split[s]];
            const size_t skip = std::min(sh.sx / GMP_NUMB_BITS, nd.size);
            mpz_roinit_n(&hi[s], nd.limbs + skip, (mp_size_t)(nd.size - skip));  // x >> sx
            mpz_init(&qfull[s]);
            mpz_init(&qp[s]);
            jobs[s] = MulJob{&qfull[s], &hi[s], inv};
        }
        m.multiply_batch(jobs.data(), ns, threads);
        for (size_t s = 0; s < ns; ++s) {
            const size_t n = mpz_size(&qfull[s]);
            const size_t skip = std::min(qskip, n);
            mpz_roinit_n(&qv[s], mpz_limbs_read(&qfull[s]) + skip, (mp_size_t)(n - skip));  // the estimate
            jobs[s] = MulJob{&qp[s], &qv[s], P};
        }
        m.multiply_batch(jobs.data(), ns, threads);
        st.products_s += lap(t);
        // Corrections, then the node becomes its high part and a low node joins.
        std::vector<std::unique_ptr<Node>> lows(ns);
        struct Ctx {
            std::vector<std::unique_ptr<Node>>* nodes;
            const std::vector<size_t>* split;
            std::vector<__mpz_struct>* qv;
            std::vector<__mpz_struct>* qfull;
            std::vector<__mpz_struct>* qp;
            std::vector<std::unique_ptr<Node>>* lows;
            mpz_srcptr P;
            size_t low, qskip;
        } ctx{&nodes, &split, &qv, &qfull, &qp, &lows, P, low, qskip};
        parallel_for(
            ns, threads,
            [](void* p, size_t s) {
                auto* c = static_cast<Ctx*>(p);
                Node& nd = *(*c->nodes)[(*c->split)[s]];
                auto* lo = new Node;
                mpz_t r;
                mpz_init(r);
                mpz_sub(r, nd.x(), &(*c->qp)[s]);  // remainder of the estimate
                long delta = 0;                    // quotient = estimate + delta
                while (mpz_sgn(r) < 0) {
                    mpz_add(r, r, c->P);
                    --delta;
                }
                while (mpz_cmp(r, c->P) >= 0) {
                    mpz_sub(r, r, c->P);
                    ++delta;
                }
                lo->take(r);
                mpz_clear(r);
                lo->width = c->low;
                lo->offset = nd.offset + nd.width - c->low;
                nd.width -= c->low;
                if (delta == 0) {
                    nd.take_top(&(*c->qfull)[s], c->qskip);
                } else {
                    mpz_t q;
                    mpz_init(q);
                    if (delta > 0)
                        mpz_add_ui(q, &(*c->qv)[s], (unsigned long)delta);
                    else
                        mpz_sub_ui(q, &(*c->qv)[s], (unsigned long)-delta);
                    nd.take(q);
                    mpz_clear(q);
                }
                (*c->lows)[s].reset(lo);
            },
            &ctx);
        for (auto& lo : lows) nodes.push_back(std::move(lo));
        st.corrections_s += lap(t);
        for (size_t s = 0; s < ns; ++s) {
            mpz_clear(&qfull[s]);
            mpz_clear(&qp[s]);
        }
    }

    // Leaves: host conversion, one node per task (largest first).
    Powers leaf_pw(2 * leaf_digits);  // shared read-only by every task
    struct LeafCtx {
        std::vector<std::unique_ptr<Node>>* nodes;
        std::string* out;
        const Powers* pw;
    } lctx{&nodes, &out, &leaf_pw};
    std::vector<double> cost(nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i) cost[i] = (double)nodes[i]->width;
    parallel_for(
        nodes.size(), threads,
        [](void* p, size_t i) {
            auto* c = static_cast<LeafCtx*>(p);
            Node& nd = *(*c->nodes)[i];
            if (nd.width <= kLeafDigits) {
                leaf(nd.x(), &(*c->out)[nd.offset], nd.width);
            } else {
                convert(nd.x(), &(*c->out)[nd.offset], nd.width, *c->pw, 0);
            }
        },
        &lctx, cost.data());
    size_t z = out.find_first_not_of('0');
    if (z != 0) out.erase(0, z);
    st.leaves_s += lap(t);
    if (stats) *stats = st;
    return out;
}

}  // namespace pi
