// device_splitting.cpp - BinarySplitting::compute_device: the upper tree
// levels with P, Q and R resident in device memory (see binary_splitting.h).
//
// The term range is cut into segments of whole leaf spans.  While the device
// reduces segment i (the calling thread drives it), segment i+1 is prepared on
// the other host threads (host subtrees, host levels, upload); the segment
// roots then merge on the device.  P, Q and R of a range do not depend on the tree
// shape, so the result is the same as one tree's.
//
// Per level: the level's results (P, Q, R of every merge, odd node copied up)
// go to one buffer and the temporaries t = Qr*Rl, u = Pl*Rr to another; the
// products run as one gpu_ntt_mul_batch_device, then R = t +- u as one
// gpu_ntt_add_batch_device.  The buffers are reused level after level.
// Magnitudes live on the device, signs on the host.  Lengths are upper
// bounds: a product has an + bn limbs, a sum max + 1, so a top limb may be 0.
//
// Signs: Q > 0; P(k) < 0 for every leaf k >= 1; R(a, b) has the sign of its
// first term (the series alternates with terms shrinking ~1.5e14-fold), so
// when t and u differ in sign |t| > |u| and R = sign(t) * (|t| - |u|).  The
// levels are enqueued without waits; a subtraction that would go negative
// sets a device flag, read once a tree is reduced, and then the caller
// computes the range on the host instead.
#include <gmp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <future>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "binary_splitting.h"
#include "device_ops.h"
#include "pintt_gpu.h"

namespace pi {
namespace {

struct DevVal {
    uint64_t* d = nullptr;
    size_t n = 0;  // limbs (upper bound)
    int sign = 1;
};
struct DevNode {
    DevVal P, Q, R;
};

void check(int rc, const char* what) {
    if (rc != PINTT_OK) throw std::runtime_error(std::string(what) + ": " + gpu_ntt_last_error());
}

// A grow-only device buffer carved into consecutive values; reset() makes it
// reusable for the next level (its previous contents are no longer needed).
struct Arena {
    uint64_t* base = nullptr;
    size_t used = 0, cap = 0;
    Arena() = default;
    ~Arena() {
        if (base) (void)gpu_ntt_dev_free(base);
    }
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    void reset(size_t limbs) {
        used = 0;
        if (limbs <= cap) return;
        if (base) check(gpu_ntt_dev_free(base), "gpu_ntt_dev_free");
        base = nullptr;
        cap = 0;
        const size_t want = limbs + limbs / 4;  // some room: later levels grow slowly
        check(gpu_ntt_dev_alloc(want, &base), "gpu_ntt_dev_alloc");
        cap = want;
    }
    uint64_t* take(size_t limbs) {
        if (used + limbs > cap) throw std::runtime_error("device arena overflow");
        uint64_t* p = base + used;
        used += limbs;
        return p;
    }
};

size_t bytes_of(mpz_srcptr x) { return std::max<size_t>(mpz_size(x), 1); }

// The device-resident part of one tree: the nodes of its current level and
// three arenas (this level's values, the next level's, the temporaries t and
// u), reused level after level.
class DeviceTree {
public:
    // Magnitudes of the host nodes into one arena, signs kept here.
    void upload(std::vector<BSResult>& nodes) {
        size_t total = 0;
        for (const auto& n : nodes) total += bytes_of(n.P) + bytes_of(n.Q) + bytes_of(n.R);
        cur_->reset(total);
        dn_.assign(nodes.size(), DevNode{});
        std::vector<uint64_t*> dst;
        std::vector<const uint64_t*> src;
        std::vector<size_t> len;
        std::vector<uint64_t> zero(1, 0);
        auto put = [&](mpz_srcptr x, DevVal& v) {
            v.n = bytes_of(x);
            v.d = cur_->take(v.n);
            v.sign = mpz_sgn(x) < 0 ? -1 : 1;
            dst.push_back(v.d);
            src.push_back(mpz_size(x) ? reinterpret_cast<const uint64_t*>(mpz_limbs_read(x)) : zero.data());
            len.push_back(v.n);
        };
        for (size_t i = 0; i < nodes.size(); ++i) {
            put(nodes[i].P, dn_[i].P);
            put(nodes[i].Q, dn_[i].Q);
            put(nodes[i].R, dn_[i].R);
        }
        check(gpu_ntt_upload(dst.size(), dst.data(), src.data(), len.data()), "gpu_ntt_upload");
        nodes.clear();
    }

    // Subtrees of PINTT_BS_LEAF_TERMS terms computed on the device
    // (gpu_ntt_bs_leaves) as the first level, a < b <= 2^24.
    void from_device_leaves(unsigned long a, unsigned long b) {
        const size_t count = (b - a + PINTT_BS_LEAF_TERMS - 1) / PINTT_BS_LEAF_TERMS;
        const size_t per = 3 * (size_t)PINTT_BS_LEAF_LIMBS;
        cur_->reset(count * per);
        uint64_t* d = cur_->take(count * per);
        std::vector<uint32_t> len(3 * count);
        std::vector<int32_t> sg(2 * count);
        check(gpu_ntt_bs_leaves(a, b, count, d, len.data(), sg.data()), "gpu_ntt_bs_leaves");
        dn_.assign(count, DevNode{});
        for (size_t i = 0; i < count; ++i) {
            uint64_t* base = d + i * per;
            dn_[i].P = DevVal{base, len[3 * i], sg[2 * i]};
            dn_[i].Q = DevVal{base + PINTT_BS_LEAF_LIMBS, len[3 * i + 1], 1};
            dn_[i].R = DevVal{base + 2 * PINTT_BS_LEAF_LIMBS, len[3 * i + 2], sg[2 * i + 1]};
        }
    }

    // Nodes that live in other trees' arenas (which must outlive this tree's
    // first level).
    void adopt(std::vector<DevNode> nodes) { dn_ = std::move(nodes); }

    // Levels until one node remains.
    void reduce(bool need_top_p, std::vector<BinarySplitting::Level>& levels);

    const DevNode& root() const { return dn_.front(); }

    // The root into host integers.
    BSResult download(bool need_p) const {
        BSResult res;
        auto get = [&](const DevVal& v, mpz_ptr x) {
            uint64_t* p = reinterpret_cast<uint64_t*>(mpz_limbs_write(x, (mp_size_t)v.n));
            const uint64_t* src = v.d;
            size_t n = v.n;
            check(gpu_ntt_download(1, &p, &src, &n), "gpu_ntt_download");
            while (n && !p[n - 1]) --n;
            mpz_limbs_finish(x, (mp_size_t)n);
            if (v.sign < 0) mpz_neg(x, x);
        };
        if (need_p && root().P.d) get(root().P, res.P);
        get(root().Q, res.Q);
        get(root().R, res.R);
        return res;
    }

    // Throws if a level's subtraction went negative (waits for the levels).
    void check_negative() const {
        if (!flag_) return;
        uint64_t v = 0;
        uint64_t* h = &v;
        const uint64_t* src = flag_;
        size_t one = 1;
        check(gpu_ntt_download(1, &h, &src, &one), "gpu_ntt_download");
        if (v) throw std::runtime_error("device-resident levels: a subtraction for R went negative");
    }

    DeviceTree() = default;
    ~DeviceTree() {
        if (flag_) (void)gpu_ntt_dev_free(flag_);
    }
    DeviceTree(const DeviceTree&) = delete;
    DeviceTree& operator=(const DeviceTree&) = delete;

private:
    // A device word the level adds set when a subtraction goes negative.
    uint32_t* negative_flag() {
        if (!flag_) {
            check(gpu_ntt_dev_alloc(1, &flag_), "gpu_ntt_dev_alloc");
            const uint64_t zero = 0;
            const uint64_t* src = &zero;
            size_t one = 1;
            check(gpu_ntt_upload(1, &flag_, &src, &one), "gpu_ntt_upload");
        }
        return reinterpret_cast<uint32_t*>(flag_);
    }

    Arena buf_[3];
    Arena* cur_ = &buf_[0];
    Arena* nxt_ = &buf_[1];
    Arena& tmp_ = buf_[2];
    std::vector<DevNode> dn_;
    uint64_t* flag_ = nullptr;
};

void DeviceTree::reduce(bool need_top_p, std::vector<BinarySplitting::Level>& levels) {
    using Clock = std::chrono::steady_clock;
    while (dn_.size() > 1) {
        auto t0 = Clock::now();
        const size_t pairs = dn_.size() / 2;
        const bool top = dn_.size() == 2;
        const bool want_p = !(top && !need_top_p);
        std::vector<DevNode> next(pairs + (dn_.size() & 1));
        size_t need = 0, tmp_need = 0, max_limbs = 0;
        for (size_t k = 0; k < pairs; ++k) {
            const DevNode &l = dn_[2 * k], &r = dn_[2 * k + 1];
            const size_t tn = r.Q.n + l.R.n, un = l.P.n + r.R.n;
            const bool same = l.R.sign == l.P.sign * r.R.sign;
            tmp_need += tn + un;
            need += (want_p ? l.P.n + r.P.n : 0) + l.Q.n + r.Q.n + std::max(tn, un) + (same ? 1 : 0);
            max_limbs = std::max({max_limbs, l.Q.n, r.Q.n, l.R.n, r.R.n});
        }
        if (dn_.size() & 1) need += dn_.back().P.n + dn_.back().Q.n + dn_.back().R.n;
        nxt_->reset(need);
        tmp_.reset(tmp_need);
        // Products.
        std::vector<const uint64_t*> A, B;
        std::vector<size_t> AN, BN;
        std::vector<uint64_t*> O;
        std::vector<DevVal> T(pairs), U(pairs);
        auto job = [&](const DevVal& x, const DevVal& y, DevVal& out, Arena& ar) {
            out.n = x.n + y.n;
            out.d = ar.take(out.n);
            out.sign = x.sign * y.sign;
            A.push_back(x.d);
            AN.push_back(x.n);
            B.push_back(y.d);
            BN.push_back(y.n);
            O.push_back(out.d);
        };
        for (size_t k = 0; k < pairs; ++k) {
            const DevNode &l = dn_[2 * k], &r = dn_[2 * k + 1];
            job(r.Q, l.R, T[k], tmp_);
            job(l.P, r.R, U[k], tmp_);
            if (want_p) job(l.P, r.P, next[k].P, *nxt_);
            job(l.Q, r.Q, next[k].Q, *nxt_);
        }
        check(gpu_ntt_mul_batch_device(O.size(), A.data(), AN.data(), B.data(), BN.data(), O.data(), nullptr),
              "gpu_ntt_mul_batch_device");
        // R = t +- u.
        std::vector<const uint64_t*> X, Y;
        std::vector<size_t> XN, YN;
        std::vector<int> SUB;
        std::vector<uint64_t*> RO;
        for (size_t k = 0; k < pairs; ++k) {
            const bool same = T[k].sign == U[k].sign;
            DevVal& R = next[k].R;
            R.n = std::max(T[k].n, U[k].n) + (same ? 1 : 0);
            R.d = nxt_->take(R.n);
            R.sign = T[k].sign;
            X.push_back(T[k].d);
            XN.push_back(T[k].n);
            Y.push_back(U[k].d);
            YN.push_back(U[k].n);
            SUB.push_back(same ? 0 : 1);
            RO.push_back(R.d);
        }
        // Enqueued without a wait: a subtraction that went negative (|t| < |u|,
        // which the sign rule above excludes) sets the tree's flag, read once
        // at the end.
        check(gpu_ntt_add_batch_device_async(pairs, X.data(), XN.data(), Y.data(), YN.data(), SUB.data(), RO.data(),
                                             negative_flag()),
              "gpu_ntt_add_batch_device_async");
        if (dn_.size() & 1) {
            const DevNode& o = dn_.back();
            DevNode& m = next.back();
            for (auto pr : {std::make_pair(&o.P, &m.P), std::make_pair(&o.Q, &m.Q), std::make_pair(&o.R, &m.R)}) {
                pr.second->n = pr.first->n;
                pr.second->sign = pr.first->sign;
                pr.second->d = nxt_->take(pr.first->n);
                check(gpu_ntt_dev_copy(pr.second->d, pr.first->d, pr.first->n), "gpu_ntt_dev_copy");
            }
        }
        std::swap(cur_, nxt_);  // the old level's buffer is reused two levels on (stream order)
        dn_.swap(next);
        levels.push_back(BinarySplitting::Level{pairs, O.size(), max_limbs,
                                                std::chrono::duration<double>(Clock::now() - t0).count()});
    }
}

}  // namespace

// The segment pipeline shared by compute_device and compute_device_qr:
// at_root(root node) runs while every device tree is alive and gives the
// result; a range reduced on the host alone is returned as it is.
template <class F>
BSResult BinarySplitting::run_segments(unsigned long a, unsigned long b, bool need_top_p, size_t device_from_limbs,
                                       F&& at_root) {
    using Clock = std::chrono::steady_clock;
    levels_.clear();
    leaf_s_ = leaf_wait_s_ = upload_s_ = 0;
    // Segments of whole leaf spans, at least 64 leaves each.  Device subtrees
    // (ranges below 2^24 terms) have a fixed span.
    const bool dev_leaves = device_leaves_ && b <= (1ul << 24) && b > a;
    const unsigned long span = dev_leaves ? (unsigned long)PINTT_BS_LEAF_TERMS : leaf_span_;
    const unsigned long nleaf = (b - a + span - 1) / span;
    const unsigned long nseg = std::max(1ul, std::min<unsigned long>(segments_, nleaf / 64));
    std::vector<unsigned long> cut(nseg + 1);
    for (unsigned long i = 0; i < nseg; ++i) cut[i] = a + (nleaf * i / nseg) * span;
    cut[nseg] = b;
    const unsigned bg_threads = threads_ > 1 ? threads_ - 1 : 1;

    // One segment from its terms to device-resident values, on any thread:
    // host subtrees, host levels while the operands are small (or the
    // dispatcher keeps them on the host), upload.
    struct Segment {
        std::unique_ptr<DeviceTree> tree;
        BSResult root;         // the whole range reduced on the host (one segment only)
        bool on_host = false;
        std::vector<Level> levels;
        double leaf_s = 0, upload_s = 0;
    };
    auto prepare = [this, &cut, nseg, need_top_p, device_from_limbs, dev_leaves](unsigned long i, unsigned threads) {
        Segment seg;
        if (dev_leaves) {
            const auto tu = std::chrono::steady_clock::now();
            seg.tree = std::make_unique<DeviceTree>();
            seg.tree->from_device_leaves(cut[i], cut[i + 1]);
            seg.leaf_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tu).count();
            return seg;
        }
        const bool top_p = nseg == 1 ? need_top_p : true;
        std::vector<BSResult> nodes = leaves(cut[i], cut[i + 1], threads, &seg.leaf_s);
        while (nodes.size() > 1) {
            if (switch_by_) {
                if (level_prefers_device(nodes, top_p)) break;
            } else {
                size_t smallest = SIZE_MAX;
                for (const auto& n : nodes)
                    smallest = std::min({smallest, mpz_size(n.P), mpz_size(n.Q), mpz_size(n.R)});
                if (smallest >= device_from_limbs) break;
            }
            host_level(nodes, top_p, seg.levels);
        }
        if (nodes.size() == 1 && nseg == 1) {
            seg.root = std::move(nodes[0]);
            seg.on_host = true;
            return seg;
        }
        const auto tu = Clock::now();
        seg.tree = std::make_unique<DeviceTree>();
        seg.tree->upload(nodes);
        seg.upload_s = std::chrono::duration<double>(Clock::now() - tu).count();
        return seg;
    };

    std::vector<std::unique_ptr<DeviceTree>> trees;
    std::vector<DevNode> roots;
    auto tw = Clock::now();
    Segment seg = prepare(0, threads_);
    leaf_wait_s_ += std::chrono::duration<double>(Clock::now() - tw).count();
    for (unsigned long i = 0; i < nseg; ++i) {
        // The next segment is prepared on the other host threads meanwhile.
        std::future<Segment> later;
        if (i + 1 < nseg) later = std::async(std::launch::async, prepare, i + 1, bg_threads);
        levels_.insert(levels_.end(), seg.levels.begin(), seg.levels.end());
        leaf_s_ += seg.leaf_s;
        upload_s_ += seg.upload_s;
        if (seg.on_host) return std::move(seg.root);
        seg.tree->reduce(nseg == 1 ? need_top_p : true, levels_);
        roots.push_back(seg.tree->root());
        trees.push_back(std::move(seg.tree));
        if (i + 1 < nseg) {
            tw = Clock::now();
            seg = later.get();
            leaf_wait_s_ += std::chrono::duration<double>(Clock::now() - tw).count();
        }
    }
    if (nseg == 1) {
        trees[0]->check_negative();
        return at_root(*trees[0]);
    }
    // The segment roots merge on the device.
    DeviceTree top;
    top.adopt(std::move(roots));
    top.reduce(need_top_p, levels_);
    for (const auto& t : trees) t->check_negative();
    top.check_negative();
    return at_root(top);
}

BSResult BinarySplitting::compute_device(unsigned long a, unsigned long b, bool need_top_p, size_t device_from_limbs) {
    return run_segments(a, b, need_top_p, device_from_limbs,
                        [need_top_p](const DeviceTree& t) { return t.download(need_top_p); });
}

void DeviceQR::download(BSResult& r) const {
    auto get = [](const dev::Num& v, int sign, mpz_ptr x) {
        uint64_t* p = reinterpret_cast<uint64_t*>(mpz_limbs_write(x, (mp_size_t)v.n));
        const uint64_t* src = v.d;
        size_t n = v.n;
        check(gpu_ntt_download(1, &p, &src, &n), "gpu_ntt_download");
        while (n && !p[n - 1]) --n;
        mpz_limbs_finish(x, (mp_size_t)n);
        if (sign < 0) mpz_neg(x, x);
    };
    get(Q, 1, r.Q);
    get(R, r_sign, r.R);
}

std::unique_ptr<DeviceQR> BinarySplitting::compute_device_qr(unsigned long a, unsigned long b,
                                                             size_t device_from_limbs) {
    auto out = std::make_unique<DeviceQR>();
    // Q and R of the root into buffers of their own (their exact lengths), so
    // the trees can go.
    auto keep = [&out](const DevVal& v, dev::Num& dst) {
        // The length is an upper bound: drop zero top limbs, reading the top
        // in windows of 64 limbs that double while they hold only zeros (the
        // root's bound can exceed its value by tens of thousands of limbs,
        // and every window read waits for the stream).
        size_t n = v.n;
        size_t w = 64;
        for (bool more = true; more && n > 1; w = std::min<size_t>(2 * w, (size_t)1 << 20)) {
            const size_t k = std::min<size_t>(n, w);
            std::vector<uint64_t> h(k);
            uint64_t* hp = h.data();
            const uint64_t* src = v.d + (n - k);
            size_t kk = k;
            check(gpu_ntt_download(1, &hp, &src, &kk), "gpu_ntt_download");
            size_t i = k;
            while (i > 0 && h[i - 1] == 0 && n > 1) {
                --i;
                --n;
            }
            more = i == 0;  // the whole window was zero
        }
        dst.n = n;
        dst.d = out->mem.alloc(n);
        check(gpu_ntt_dev_copy(dst.d, v.d, n), "gpu_ntt_dev_copy");
    };
    BSResult host = run_segments(a, b, false, device_from_limbs, [&](const DeviceTree& t) {
        keep(t.root().Q, out->Q);
        keep(t.root().R, out->R);
        out->r_sign = t.root().R.sign;
        check(gpu_ntt_synchronize(), "gpu_ntt_synchronize");
        return BSResult{};
    });
    if (!out->Q.d) {
        // The range was reduced on the host alone: upload its Q and R.
        auto put = [&out](mpz_srcptr x, dev::Num& dst) {
            dst.n = std::max<size_t>(mpz_size(x), 1);
            dst.d = out->mem.alloc(dst.n);
            const uint64_t zero = 0;
            const uint64_t* src = mpz_size(x) ? reinterpret_cast<const uint64_t*>(mpz_limbs_read(x)) : &zero;
            check(gpu_ntt_upload(1, &dst.d, &src, &dst.n), "gpu_ntt_upload");
        };
        put(host.Q, out->Q);
        put(host.R, out->R);
        out->r_sign = mpz_sgn(host.R) < 0 ? -1 : 1;
    }
    return out;
}

bool BinarySplitting::level_prefers_device(const std::vector<BSResult>& nodes, bool need_top_p) const {
    const size_t pairs = nodes.size() / 2;
    const bool top = nodes.size() == 2;
    std::vector<ProductShape> shapes;
    std::vector<double> work;
    auto add = [&](mpz_srcptr x, mpz_srcptr y) {
        const size_t xn = mpz_size(x), yn = mpz_size(y);
        shapes.push_back(ProductShape{xn, yn, 0, false, true});
        const double n = 2.0 * (double)(xn + yn);
        work.push_back(n * std::log2(n + 2.0));
    };
    for (size_t k = 0; k < pairs; ++k) {
        const BSResult &l = nodes[2 * k], &r = nodes[2 * k + 1];
        add(r.Q, l.R);
        add(l.P, r.R);
        if (!(top && !need_top_p)) add(l.P, r.P);
        add(l.Q, r.Q);
    }
    for (auto& sh : shapes) sh.batch = shapes.size();
    std::vector<Path> paths(shapes.size());
    switch_by_->choose_batch(shapes.data(), shapes.size(), paths.data());
    double gpu = 0, all = 0;
    for (size_t i = 0; i < shapes.size(); ++i) {
        all += work[i];
        if (paths[i] == Path::Gpu) gpu += work[i];
    }
    return gpu >= 0.5 * all;
}

}  // namespace pi
