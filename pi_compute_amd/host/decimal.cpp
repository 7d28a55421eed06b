// decimal.cpp - see decimal.h.
#include "decimal.h"

#include <algorithm>
#include <chrono>

#include <cstring>
#include <future>
#include <memory>
#include <thread>
#include <vector>

#include "device_ops.h"
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
// ((x >> sx) * inv) >> (e2 - sx) with 2^e2/P - 9 <= inv <= 2^e2/P - 1 (the
// Newton reciprocal, within 4, minus 5).  Both shifts are whole limbs, so the
// shifted operands are views of the numbers they come from.  With x < P^2,
// 2^sx <= P / 2^63 and 2^e2 >= 2^64 P^2 the estimate is floor(x / P) or one
// less: it never exceeds x / P, and it falls short by at most 2^-63 +
// 9 x / 2^e2 + 1 < 2.
struct SplitShifts {
    size_t sx, e2;
};
SplitShifts split_shifts(size_t np) {
    const size_t L = GMP_NUMB_BITS;
    return SplitShifts{(np > 64 ? np - 64 : 0) / L * L, (2 * np + 64 + L - 1) / L * L};
}

// The conversion's levels on the host: both product batches of a level
// through m, corrections on host threads.
void split_on_host(mpz_srcptr n, size_t width, const DecimalPowers& pre, size_t leaf_digits, Multiplier& m,
                   unsigned threads, std::vector<std::unique_ptr<Node>>& nodes, DecimalStats& st) {
    using Clock = std::chrono::steady_clock;
    auto t = Clock::now();
    auto lap = [](Clock::time_point& t) {
        const auto now = Clock::now();
        const double s = std::chrono::duration<double>(now - t).count();
        t = now;
        return s;
    };
    const int k = (int)pre.pw.size() - 1;
    nodes.clear();
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
        mpz_srcptr P = &pre.pw[j];
        mpz_srcptr inv = &pre.inv[j];
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
        std::vector<long> deltas(ns, 0);
        struct Ctx {
            std::vector<std::unique_ptr<Node>>* nodes;
            const std::vector<size_t>* split;
            std::vector<__mpz_struct>* qv;
            std::vector<__mpz_struct>* qfull;
            std::vector<__mpz_struct>* qp;
            std::vector<std::unique_ptr<Node>>* lows;
            std::vector<long>* delta;
            mpz_srcptr P;
            size_t low, qskip;
        } ctx{&nodes, &split, &qv, &qfull, &qp, &lows, &deltas, P, low, qskip};
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
                (*c->delta)[s] = delta;
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
        for (long d : deltas) {
            st.min_correction = std::min(st.min_correction, d);
            st.max_correction = std::max(st.max_correction, d);
        }
        st.corrections_s += lap(t);
        for (size_t s = 0; s < ns; ++s) {
            mpz_clear(&qfull[s]);
            mpz_clear(&qp[s]);
        }
    }

}

void dcheck(int rc, const char* what) { dev::check(rc, what); }
using DevBuffers = dev::Buffers;
using DevNum = dev::Num;

// Limbs that hold every value below 10^digits.
size_t limbs_for_digits(size_t digits) { return (size_t)((double)digits * 3.3219280948873623 / 64.0) + 2; }

// Each nonzero src[i] into dst[i] (mpz_size(src[i]) limbs).
void upload(const std::vector<mpz_srcptr>& src, const std::vector<DevNum>& dst) {
    std::vector<uint64_t*> d;
    std::vector<const uint64_t*> h;
    std::vector<size_t> n;
    for (size_t i = 0; i < src.size(); ++i) {
        const size_t sz = mpz_size(src[i]);
        if (!sz) throw std::runtime_error("device conversion: zero operand");
        d.push_back(dst[i].d);
        h.push_back(reinterpret_cast<const uint64_t*>(mpz_limbs_read(src[i])));
        n.push_back(sz);
    }
    if (!d.empty()) dcheck(gpu_ntt_upload(d.size(), d.data(), h.data(), n.data()), "gpu_ntt_upload");
}

struct DevSplitNode {
    DevNum x;
    size_t offset, width;
};

// The whole conversion on device-resident numbers, into out (width
// characters).  Each level is two product batches (estimate, estimate * P), a
// flagged subtraction for the remainder x - q*P (never negative: the estimate
// never exceeds x / P), a flagged subtraction of P (the estimate is
// floor(x / P) or one less) and one that confirms the remainder is below P,
// and an increment of the quotients that were one short.  The leaves (at most
// PINTT_DECIMAL_LEAF_MAX_DIGITS digits: leaf_digits must not be larger) are
// written as digits by the device and come back in one copy.
void device_convert(const DevNum* xdev, mpz_srcptr n, size_t width, const DecimalPowers& pre, size_t leaf_digits,
                    std::string& out, DecimalStats& st) {
    using Clock = std::chrono::steady_clock;
    auto t = Clock::now();
    auto lap = [&t]() {
        const auto now = Clock::now();
        const double s = std::chrono::duration<double>(now - t).count();
        t = now;
        return s;
    };
    static_assert(sizeof(mp_limb_t) == sizeof(uint64_t), "64-bit limbs");
    DevBuffers mem;
    const int k = (int)pre.pw.size() - 1;
    int jmin = k + 1;
    while (jmin > 0 && ((size_t)1 << (jmin - 1)) >= leaf_digits) --jmin;
    // Powers and reciprocals (uploaded here unless the caller's copies are on
    // the device already), the constant 1 (an all-zero limb follows it) and n.
    std::vector<DevNum> P(k + 1), inv(k + 1);
    const bool have = pre.dev && pre.dev->jmin <= jmin && (int)pre.dev->P.size() == k + 1;
    size_t total = 2;
    if (!have)
        for (int j = jmin; j <= k; ++j) total += mpz_size(&pre.pw[j]) + mpz_size(&pre.inv[j]);
    const size_t nn = xdev ? 0 : std::max<size_t>(mpz_size(n), 1);
    total += nn;
    uint64_t* base = mem.alloc(total);
    std::vector<mpz_srcptr> src;
    std::vector<DevNum> dst;
    size_t used = 0;
    auto carve = [&](size_t limbs) {
        DevNum v{base + used, std::max<size_t>(limbs, 1)};
        used += v.n;
        return v;
    };
    for (int j = jmin; j <= k; ++j) {
        if (have) {
            P[j] = pre.dev->P[j];
            inv[j] = pre.dev->inv[j];
            continue;
        }
        P[j] = carve(mpz_size(&pre.pw[j]));
        inv[j] = carve(mpz_size(&pre.inv[j]));
        src.push_back(&pre.pw[j]);
        dst.push_back(P[j]);
        src.push_back(&pre.inv[j]);
        dst.push_back(inv[j]);
    }
    DevNum one = carve(2);
    DevNum root = xdev ? *xdev : carve(nn);
    {
        const uint64_t o[2] = {1, 0};
        const uint64_t* op = o;
        size_t on = 2;
        dcheck(gpu_ntt_upload(1, &one.d, &op, &on), "gpu_ntt_upload");
    }
    if (!xdev) {
        src.push_back(n);
        dst.push_back(root);
    }
    if (!src.empty()) upload(src, dst);
    // The root's length: at most the limbs its width allows.
    root.n = std::min(root.n, limbs_for_digits(width));
    DevNum zero{one.d + 1, 1};

    std::vector<DevSplitNode> nodes{{root, 0, width}};
    for (int j = k; j >= jmin; --j) {
        const size_t low = (size_t)1 << j;
        std::vector<size_t> split;
        for (size_t i = 0; i < nodes.size(); ++i)
            if (nodes[i].width > low) split.push_back(i);
        if (split.empty()) continue;
        const size_t ns = split.size();
        const SplitShifts sh = split_shifts(mpz_sizeinbase(&pre.pw[j], 2));
        const size_t skip = sh.sx / GMP_NUMB_BITS, qskip = (sh.e2 - sh.sx) / GMP_NUMB_BITS;
        const DevNum Pj = P[j], Ij = inv[j];
        // Operands and room for the level: estimates, products, remainders (two
        // candidates each) and incremented quotients.
        std::vector<DevNum> hi(ns), qv(ns);
        std::vector<size_t> qf_n(ns), qp_n(ns), r_n(ns), qq_n(ns);
        size_t room = 0;
        for (size_t s = 0; s < ns; ++s) {
            const DevNum x = nodes[split[s]].x;
            hi[s] = x.n > skip ? DevNum{x.d + skip, x.n - skip} : zero;
            qf_n[s] = hi[s].n + Ij.n;
            const size_t qvn = qf_n[s] > qskip ? qf_n[s] - qskip : 1;
            qp_n[s] = qvn + Pj.n;
            r_n[s] = std::max(x.n, qp_n[s]);
            qq_n[s] = qvn + 1;
            room += qf_n[s] + qp_n[s] + 2 * r_n[s] + qq_n[s];
        }
        uint64_t* lv = mem.alloc(room);
        size_t at = 0;
        auto take = [&](size_t limbs) {
            uint64_t* p = lv + at;
            at += limbs;
            return p;
        };
        std::vector<DevNum> qf(ns), qp(ns), r(ns), r1(ns), qq(ns);
        for (size_t s = 0; s < ns; ++s) {
            qf[s] = {take(qf_n[s]), qf_n[s]};
            qp[s] = {take(qp_n[s]), qp_n[s]};
            r[s] = {take(r_n[s]), r_n[s]};
            r1[s] = {take(r_n[s]), r_n[s]};
            qq[s] = {take(qq_n[s]), qq_n[s]};
            qv[s] = qf[s].n > qskip ? DevNum{qf[s].d + qskip, qf[s].n - qskip} : zero;
        }
        // q = (hi * inv) >> (e2 - sx), then q * P.
        std::vector<const uint64_t*> A(ns), B(ns), X(ns), Y(ns);
        std::vector<uint64_t*> O(ns);
        std::vector<size_t> AN(ns), BN(ns), XN(ns), YN(ns);
        std::vector<int> SUB(ns, 1), NEG(ns);
        for (size_t s = 0; s < ns; ++s) {
            A[s] = hi[s].d, AN[s] = hi[s].n, B[s] = Ij.d, BN[s] = Ij.n, O[s] = qf[s].d;
        }
        dcheck(gpu_ntt_mul_batch_device(ns, A.data(), AN.data(), B.data(), BN.data(), O.data(), nullptr),
               "gpu_ntt_mul_batch_device");
        for (size_t s = 0; s < ns; ++s) {
            A[s] = qv[s].d, AN[s] = qv[s].n, B[s] = Pj.d, BN[s] = Pj.n, O[s] = qp[s].d;
        }
        dcheck(gpu_ntt_mul_batch_device(ns, A.data(), AN.data(), B.data(), BN.data(), O.data(), nullptr),
               "gpu_ntt_mul_batch_device");
        // r = x - q P >= 0.
        for (size_t s = 0; s < ns; ++s) {
            const DevNum x = nodes[split[s]].x;
            X[s] = x.d, XN[s] = x.n, Y[s] = qp[s].d, YN[s] = qp[s].n, O[s] = r[s].d;
        }
        dcheck(gpu_ntt_add_batch_device_flags(ns, X.data(), XN.data(), Y.data(), YN.data(), SUB.data(), O.data(),
                                              NEG.data()),
               "gpu_ntt_add_batch_device_flags");
        st.products_s += lap();
        for (size_t s = 0; s < ns; ++s)
            if (NEG[s]) throw std::runtime_error("device conversion: quotient estimate above x / P");
        // r1 = r - P: where not negative the quotient was one short.
        for (size_t s = 0; s < ns; ++s) {
            X[s] = r[s].d, XN[s] = r[s].n, Y[s] = Pj.d, YN[s] = Pj.n, O[s] = r1[s].d;
        }
        dcheck(gpu_ntt_add_batch_device_flags(ns, X.data(), XN.data(), Y.data(), YN.data(), SUB.data(), O.data(),
                                              NEG.data()),
               "gpu_ntt_add_batch_device_flags");
        std::vector<size_t> short_by_one;
        for (size_t s = 0; s < ns; ++s)
            if (!NEG[s]) short_by_one.push_back(s);
        if (!short_by_one.empty()) st.max_correction = 1;
        if (!short_by_one.empty()) {
            // Those remainders must now be below P (r1 - P negative; the
            // difference lands in r, no longer needed), and q + 1 replaces q.
            const size_t m = short_by_one.size();
            std::vector<const uint64_t*> X2(m), Y2(m);
            std::vector<uint64_t*> O2(m);
            std::vector<size_t> XN2(m), YN2(m);
            std::vector<int> SUB2(m, 1), NEG2(m);
            for (size_t i = 0; i < m; ++i) {
                const size_t s = short_by_one[i];
                X2[i] = r1[s].d, XN2[i] = r1[s].n, Y2[i] = Pj.d, YN2[i] = Pj.n, O2[i] = r[s].d;
            }
            dcheck(gpu_ntt_add_batch_device_flags(m, X2.data(), XN2.data(), Y2.data(), YN2.data(), SUB2.data(),
                                                  O2.data(), NEG2.data()),
                   "gpu_ntt_add_batch_device_flags");
            for (size_t i = 0; i < m; ++i)
                if (!NEG2[i]) throw std::runtime_error("device conversion: quotient estimate two or more short");
            std::vector<int> ADD(m, 0);
            for (size_t i = 0; i < m; ++i) {
                const size_t s = short_by_one[i];
                X2[i] = qv[s].d, XN2[i] = qv[s].n, Y2[i] = one.d, YN2[i] = 1, O2[i] = qq[s].d;
            }
            dcheck(gpu_ntt_add_batch_device(m, X2.data(), XN2.data(), Y2.data(), YN2.data(), ADD.data(), O2.data()),
                   "gpu_ntt_add_batch_device");
        }
        // The node keeps the quotient (width - low digits), a new node takes the
        // remainder (below P: its limbs).
        std::vector<bool> incremented(ns, false);
        for (size_t s : short_by_one) incremented[s] = true;
        for (size_t s = 0; s < ns; ++s) {
            DevSplitNode& nd = nodes[split[s]];
            DevNum rem = incremented[s] ? r1[s] : r[s];
            rem.n = std::min(rem.n, Pj.n);
            DevNum q = incremented[s] ? qq[s] : qv[s];
            nd.width -= low;
            q.n = std::min(q.n, limbs_for_digits(nd.width));
            nd.x = q;
            nodes.push_back(DevSplitNode{rem, nd.offset + nd.width, low});
        }
        st.corrections_s += lap();
    }
    // Leaves: digits written on the device, then one copy into out.
    const size_t cnt = nodes.size();
    std::vector<const uint64_t*> LX(cnt);
    std::vector<size_t> LN(cnt), LW(cnt), LO(cnt);
    for (size_t i = 0; i < cnt; ++i) {
        if (nodes[i].width > (size_t)PINTT_DECIMAL_LEAF_MAX_DIGITS)
            throw std::runtime_error("device conversion: leaf wider than the device leaves take");
        LX[i] = nodes[i].x.d;
        LN[i] = nodes[i].x.n;
        LW[i] = nodes[i].width;
        LO[i] = nodes[i].offset;
    }
    const size_t words = (width + 7) / 8;
    uint64_t* digits = mem.alloc(words);
    dcheck(gpu_ntt_decimal_leaves_device(cnt, LX.data(), LN.data(), LW.data(), LO.data(),
                                         reinterpret_cast<uint8_t*>(digits), width),
           "gpu_ntt_decimal_leaves_device");
    out.resize(words * 8);
    uint64_t* h = reinterpret_cast<uint64_t*>(&out[0]);
    const uint64_t* dd = digits;
    dcheck(gpu_ntt_download(1, &h, &dd, &words), "gpu_ntt_download");
    out.resize(width);
    st.leaves_s += lap();
}

}  // namespace

DecimalPowers::~DecimalPowers() {
    for (auto& x : pw) mpz_clear(&x);
    for (auto& x : inv) mpz_clear(&x);
}

DevicePowers::DevicePowers(const DecimalPowers& d, int jmin_) : jmin(jmin_) {
    const int k = (int)d.pw.size() - 1;
    P.resize(k + 1);
    inv.resize(k + 1);
    size_t total = 0;
    for (int j = jmin; j <= k; ++j) total += mpz_size(&d.pw[j]) + mpz_size(&d.inv[j]);
    uint64_t* base = mem.alloc(total);
    std::vector<mpz_srcptr> src;
    std::vector<DevNum> dst;
    size_t used = 0;
    for (int j = jmin; j <= k; ++j) {
        P[j] = DevNum{base + used, mpz_size(&d.pw[j])};
        used += P[j].n;
        inv[j] = DevNum{base + used, mpz_size(&d.inv[j])};
        used += inv[j].n;
        src.push_back(&d.pw[j]);
        dst.push_back(P[j]);
        src.push_back(&d.inv[j]);
        dst.push_back(inv[j]);
    }
    upload(src, dst);
}

// The conversion's reciprocals take their host steps up to 2^16 limbs: those
// run for every reciprocal at once on threads of their own, while each device
// step waits on the library stream behind binary splitting.
constexpr size_t kConvHostLimbs = 65536;

void decimal_powers(DecimalPowers& d, size_t width, Multiplier& m, size_t leaf_digits, bool device) {
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
    int jlo = k + 1;  // the reciprocals made: j = jlo .. k
    while (jlo > 0 && ((size_t)1 << (jlo - 1)) >= leaf_digits) --jlo;
    auto shift_limbs = [&d](int j) {  // whole limbs: e2 / 64 = limbs(P) + B
        return split_shifts(mpz_sizeinbase(&d.pw[j], 2)).e2 / 64 - mpz_size(&d.pw[j]);
    };
    // On the device, the host steps that start every reciprocal (GMP alone,
    // up to 2^14 limbs) run first, all at once on threads of their own; the
    // device steps then run one reciprocal after another.
    struct Start {
        mpz_t v;
        Start() { mpz_init(v); }
        ~Start() { mpz_clear(v); }
    };
    std::unique_ptr<Start[]> starts(device && jlo <= k ? new Start[k + 1 - jlo] : nullptr);
    if (starts) {
        std::vector<std::future<void>> pending;
        for (int j = k; j >= jlo; --j)
            pending.push_back(std::async(std::launch::async, [&d, &starts, &shift_limbs, j, jlo] {
                device_reciprocal_start(starts[j - jlo].v, &d.pw[j], shift_limbs(j), kConvHostLimbs);
            }));
        for (auto& f : pending) f.get();
    }
    for (int j = k; j >= jlo; --j) {
        if (device) {
            device_reciprocal_limbs(&d.inv[j], &d.pw[j], shift_limbs(j), starts[j - jlo].v, kConvHostLimbs);
        } else {
            const size_t np = mpz_sizeinbase(&d.pw[j], 2);
            newton_reciprocal(&d.inv[j], &d.pw[j], split_shifts(np).e2 - np, m);  // ~ 2^e2 / P, within 4
        }
        mpz_sub_ui(&d.inv[j], &d.inv[j], 5);  // below 2^e2 / P
    }
    if (device) {
        int jmin = k + 1;
        while (jmin > 0 && ((size_t)1 << (jmin - 1)) >= leaf_digits) --jmin;
        d.dev = std::make_unique<DevicePowers>(d, jmin);
    }
    d.reciprocals_s = std::chrono::duration<double>(Clock::now() - t).count();
}

std::string to_decimal(mpz_srcptr n, Multiplier& m, unsigned threads, size_t leaf_digits, DecimalStats* stats,
                       const DecimalPowers* pre, bool device) {
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
    if (device) leaf_digits = std::min<size_t>(leaf_digits, PINTT_DECIMAL_LEAF_MAX_DIGITS);
    const size_t width = mpz_sizeinbase(n, 10);
    if (width <= 2 * leaf_digits) return to_decimal(n, threads);
    std::string out;

    // The powers depend on the width only through k (2^k < width <= 2^(k+1)).
    int kw = 0;
    while (((size_t)2 << kw) < width) ++kw;
    DecimalPowers own;
    if (!pre || (int)pre->pw.size() != kw + 1 || pre->leaf_digits != leaf_digits) {
        decimal_powers(own, width, m, leaf_digits);
        pre = &own;
    }
    st.powers_s = pre->powers_s;
    st.reciprocals_s = pre->reciprocals_s;
    lap(t);
    if (device) {
        device_convert(nullptr, n, width, *pre, leaf_digits, out, st);
        size_t z = out.find_first_not_of('0');
        if (z != 0) out.erase(0, z);
        if (stats) *stats = st;
        return out;
    }
    out.assign(width, '0');
    std::vector<std::unique_ptr<Node>> nodes;
    split_on_host(n, width, *pre, leaf_digits, m, threads, nodes, st);
    lap(t);

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

std::string to_decimal_device(const uint64_t* x, size_t xn, unsigned threads, size_t leaf_digits,
                              DecimalStats* stats, const DecimalPowers& pre) {
    DecimalStats st;
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    leaf_digits = std::min<size_t>(std::max<size_t>(leaf_digits, kLeafDigits), PINTT_DECIMAL_LEAF_MAX_DIGITS);
    // The true length and top limb, from the top in windows of 64 limbs that
    // double while they hold only zeros.
    size_t n = xn;
    uint64_t top = 0;
    for (size_t w = 64; n > 0 && !top; w = std::min<size_t>(2 * w, (size_t)1 << 20)) {
        const size_t k = std::min<size_t>(n, w);
        std::vector<uint64_t> h(k);
        uint64_t* hp = h.data();
        const uint64_t* src = x + (n - k);
        size_t kk = k;
        dcheck(gpu_ntt_download(1, &hp, &src, &kk), "gpu_ntt_download");
        size_t i = k;
        while (i > 0 && h[i - 1] == 0) --i;
        if (i) {
            n -= k - i;
            top = h[i - 1];
        } else {
            n -= k;
        }
    }
    if (n == 0) return "0";
    size_t bits = 64 * (n - 1);
    for (uint64_t t = top; t; t >>= 1) ++bits;
    // x < 2^bits, so its digits are at most floor(bits log10(2)) + 1.
    const size_t width = (size_t)((double)bits * 0.30102999566398120) + 1;
    if (width <= 2 * leaf_digits) {
        mpz_t h;
        mpz_init(h);
        uint64_t* hp = reinterpret_cast<uint64_t*>(mpz_limbs_write(h, (mp_size_t)n));
        size_t nn = n;
        dcheck(gpu_ntt_download(1, &hp, &x, &nn), "gpu_ntt_download");
        mpz_limbs_finish(h, (mp_size_t)n);
        std::string s = to_decimal(h, threads);
        mpz_clear(h);
        return s;
    }
    int kw = 0;
    while (((size_t)2 << kw) < width) ++kw;
    if ((int)pre.pw.size() != kw + 1 || pre.leaf_digits != leaf_digits)
        throw std::runtime_error("device conversion: powers built for another width or leaf");
    st.powers_s = pre.powers_s;
    st.reciprocals_s = pre.reciprocals_s;
    std::string out;
    const DevNum root{const_cast<uint64_t*>(x), n};
    device_convert(&root, nullptr, width, pre, leaf_digits, out, st);
    size_t z = out.find_first_not_of('0');
    if (z == std::string::npos) return "0";
    if (z != 0) out.erase(0, z);
    if (stats) *stats = st;
    return out;
}

}  // namespace pi
