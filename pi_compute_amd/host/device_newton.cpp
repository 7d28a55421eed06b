// device_newton.cpp - whole-limb Newton iterations on device numbers: the
// reciprocal and the inverse square root (newton.h), and the final step of pi
// (pi_engine.h, pi_scaled_newton_device).
//
// Every truncation keeps whole limbs (a pointer offset), products are
// gpu_ntt_mul_batch_device calls, additions and subtractions
// gpu_ntt_add_batch_device_flags, and the host only reads a few limbs to learn
// signs.  Steps up to 16384 limbs run on the host (GMP is faster there than
// device round trips); both sides compute the same integers.
//
// Reciprocal (beta = 2^64): Y ~ beta^(l + c) / D_l at precision c limbs,
// D_l the top l = min(Ld, c + G) limbs of D.  One step to c' <= 2c - 2:
//   D' = top l' = min(Ld, c' + G) limbs of D, K = l' + c, p = D' Y ~ beta^K,
//   E = beta^K - p (|E| < beta^(K - c + 2) = beta^m), E_t = the limbs of |E|
//   from m - (c' - c + G) up to m, term = (Y E_t) >> (K - (m - (c' - c + G))
//   - (c' - c)) limbs, Y' = Y beta^(c' - c) +- term.
// Inverse square root: the same step with p = a Y^2 ~ beta^(2c) (K = 2c) and
// the correction halved, as E_t times 2^63 = beta / 2 with one more limb
// dropped.
// With G = 3 guard limbs each truncation costs under one unit of Y', and
// the step's own error (the square of the relative error of Y) is below one
// as well: Y ends within a few units of the exact value.  The same steps on
// host integers (newton_reciprocal_limbs, newton_inv_sqrt_limbs) are checked
// against exact quotients and square roots by the host tests, and the device
// results against them, integer for integer, by the GPU tests.
#include <gmp.h>

#include <cstring>
#include <stdexcept>
#include <utility>
#include <vector>

#include "device_ops.h"
#include "newton.h"
#include "pi_engine.h"

namespace pi {
namespace {

using dev::Num;

class Dev {
public:
    explicit Dev(size_t limbs) : base_(mem_.alloc(limbs)), cap_(limbs) {
        const uint64_t c[4] = {0, 1, 1ull << 63, 0};
        zero = take(1);
        one = take(1);
        half = take(1);
        sink_ = reinterpret_cast<uint32_t*>(take(1).d);
        const uint64_t* src[4] = {&c[0], &c[1], &c[2], &c[3]};
        uint64_t* dst[4] = {zero.d, one.d, half.d, reinterpret_cast<uint64_t*>(sink_)};
        const size_t n[4] = {1, 1, 1, 1};
        dev::check(gpu_ntt_upload(4, dst, src, n), "gpu_ntt_upload");
    }
    Num take(size_t n) {
        n = std::max<size_t>(n, 1);
        if (used_ + n > cap_) throw std::runtime_error("device final step: arena overflow");
        Num r{base_ + used_, n};
        used_ += n;
        return r;
    }
    Num upload(mpz_srcptr x, size_t from_limb, size_t limbs) {
        Num r = take(limbs);
        const uint64_t* h = reinterpret_cast<const uint64_t*>(mpz_limbs_read(x)) + from_limb;
        dev::check(gpu_ntt_upload(1, &r.d, &h, &limbs), "gpu_ntt_upload");
        return r;
    }
    Num mul(Num a, Num b) {
        Num r = take(a.n + b.n);
        const uint64_t* A = a.d;
        const uint64_t* B = b.d;
        dev::check(gpu_ntt_mul_batch_device(1, &A, &a.n, &B, &b.n, &r.d, nullptr), "gpu_ntt_mul_batch_device");
        return r;
    }
    // Limbs [from, from + count) of x (zero past its length).
    std::vector<uint64_t> fetch(Num x, size_t from, size_t count) {
        std::vector<uint64_t> h(count, 0);
        if (from < x.n) {
            size_t k = std::min(count, x.n - from);
            uint64_t* hp = h.data();
            const uint64_t* s = x.d + from;
            dev::check(gpu_ntt_download(1, &hp, &s, &k), "gpu_ntt_download");
        }
        return h;
    }
    // out = x + y (max + 1 limbs) or x - y (max limbs); returns whether a
    // subtraction went negative (the result is then x - y + beta^max).
    bool addsub(uint64_t* out, Num x, Num y, bool sub) {
        const uint64_t* X = x.d;
        const uint64_t* Y = y.d;
        const int s = sub ? 1 : 0;
        int neg = 0;
        dev::check(gpu_ntt_add_batch_device_flags(1, &X, &x.n, &Y, &y.n, &s, &out, &neg),
                   "gpu_ntt_add_batch_device_flags");
        return neg != 0;
    }
    // out = x + y or x - y as addsub, enqueued without a wait: for a result
    // whose sign is known (x - y below zero is still x - y + beta^max; its
    // flag goes to a device word nobody reads).
    void addsub_async(uint64_t* out, Num x, Num y, bool sub) {
        const uint64_t* X = x.d;
        const uint64_t* Y = y.d;
        const int s = sub ? 1 : 0;
        dev::check(gpu_ntt_add_batch_device_async(1, &X, &x.n, &Y, &y.n, &s, &out, sink_),
                   "gpu_ntt_add_batch_device_async");
    }
    Num zero, one, half;  // 0, 1 and 2^63 (beta / 2)

private:
    uint32_t* sink_ = nullptr;  // the negative flag of addsub_async
    dev::Buffers mem_;
    uint64_t* base_;
    size_t cap_, used_ = 0;
};

constexpr size_t kGuard = 3;
constexpr size_t kExactLimbs = 16;   // the host mirror's exact start (newton.cpp)
constexpr size_t kHostLimbs = kReciprocalHostLimbs;

// One Newton step from precision c to cp limbs, given p ~ beta^K on the
// device (p.n >= K + 1): E = beta^K - p, E_t = the limbs of |E| from m - T to
// m (m = K - c + 2, T = cp - c + kGuard), Y' = Y beta^(cp - c) +- (Y E_t) >>
// (K - (m - T) - (cp - c)) limbs.  half: the correction is halved (inverse
// square root), as E_t times 2^63 with one more limb dropped.
Num newton_step(Dev& g, Num Y, Num p, size_t K, size_t c, size_t cp, bool half) {
    const size_t s = cp - c;
    if (p.n < K + 1) throw std::runtime_error("device Newton step: product shorter than expected");
    const std::vector<uint64_t> top = g.fetch(p, K - 1, p.n - K + 1);
    bool above = false;
    for (size_t i = 2; i < top.size(); ++i) above = above || top[i] != 0;
    Num E;  // |beta^K - p|, K limbs
    bool positive;
    if (!above && top[1] == 1) {  // p >= beta^K
        positive = false;
        E = Num{p.d, K};
    } else if (!above && top[1] == 0 && top[0] == ~0ull) {  // p < beta^K
        positive = true;
        E = g.take(K);
        g.addsub_async(E.d, g.zero, Num{p.d, K}, true);  // beta^K - p (0 - p + beta^K)
    } else {
        throw std::runtime_error("device Newton step: product far from beta^K");
    }
    const size_t m = K - c + 2;
    if (m < K) {
        const std::vector<uint64_t> z = g.fetch(E, m, std::min<size_t>(2, K - m));
        for (uint64_t v : z)
            if (v) throw std::runtime_error("device Newton step: correction above its bound");
    }
    const size_t T = s + kGuard;
    const size_t lo = m > T ? m - T : 0;
    Num Et{E.d + lo, std::min(m, K) - lo};
    if (half) Et = g.mul(Et, g.half);
    const Num p2 = g.mul(Y, Et);
    const size_t sh = K - lo - s + (half ? 1 : 0);
    if (p2.n <= sh + s) throw std::runtime_error("device Newton step: correction shorter than expected");
    const Num term{p2.d + sh, p2.n - sh};
    const Num hi{term.d + s, term.n - s};
    const size_t hn = std::max(Y.n, hi.n) + 1;
    Num Yn = g.take(s + hn);
    if (positive) {
        dev::check(gpu_ntt_dev_copy(Yn.d, term.d, s), "gpu_ntt_dev_copy");
        g.addsub_async(Yn.d + s, Y, hi, false);
    } else {
        // Y beta^s - term: the low limbs are -term mod beta^s, with a borrow
        // into the high part when they are not all zero.
        const bool borrow = g.addsub(Yn.d, g.zero, Num{term.d, s}, true);
        if (borrow) {
            Num t = g.take(std::max(Y.n, hi.n));
            if (g.addsub(t.d, Y, hi, true) || g.addsub(Yn.d + s, t, g.one, true))
                throw std::runtime_error("device Newton step: negative estimate");
        } else if (g.addsub(Yn.d + s, Y, hi, true)) {
            throw std::runtime_error("device Newton step: negative estimate");
        }
        Yn.n = s + std::max(Y.n, hi.n);
    }
    Yn.n = std::min(Yn.n, cp + 2);  // Y' <= beta^(cp + 1) plus a few units
    return Yn;
}

// The precisions of a reciprocal's device steps (descending) and the one its
// host start reaches: the steps up to kHostLimbs run on the host
// (newton_reciprocal_limbs to that precision is the same integer the device
// steps would reach; small products are cheaper there than device round
// trips).
std::vector<size_t> reciprocal_schedule(size_t B, size_t* c_start, size_t host_limbs = kHostLimbs) {
    std::vector<size_t> sched;
    for (size_t c = B; c > kExactLimbs; c = (c + 3) / 2) sched.push_back(c);
    size_t c = sched.empty() ? B : (sched.back() + 3) / 2;
    while (!sched.empty() && sched.back() <= host_limbs) {
        c = sched.back();
        sched.pop_back();
    }
    *c_start = c;
    return sched;
}

// The host start: beta^(Ld + c) / D to precision c, from D's top limbs Dh (at
// least min(Ld, c + kGuard) of them, or all).
void reciprocal_start(mpz_t t, mpz_srcptr Dh, size_t Ld, size_t c) {
    const size_t need = std::min(Ld, c + kGuard);
    const size_t have = mpz_size(Dh);
    if (have < need || have > Ld) throw std::runtime_error("device reciprocal: host top limbs do not fit");
    // The top `need` limbs (the same integer the whole divisor gives the host
    // steps: each reads at most c + kGuard top limbs).
    mpz_t top;
    mpz_roinit_n(top, mpz_limbs_read(Dh) + (have - need), (mp_size_t)need);
    newton_reciprocal_limbs(t, need == Ld ? Dh : top, c);
}

// Y ~ beta^(Ld + B) / D; D on the device (exactly Ld limbs, top limb nonzero)
// and its top limbs on the host (at least min(Ld, kHostLimbs + kGuard + 1)
// of them, or all): the steps up to kHostLimbs read no others.  start: the
// host start (reciprocal_start) computed beforehand, or null; host_limbs: the
// precision the host steps go to (the same integer whatever it is).
Num reciprocal(Dev& g, Num D, mpz_srcptr Dh, size_t B, mpz_srcptr start = nullptr,
               size_t host_limbs = kHostLimbs) {
    const size_t Ld = D.n;
    size_t c = 0;
    const std::vector<size_t> sched = reciprocal_schedule(B, &c, host_limbs);
    Num Y;
    if (start) {
        Y = g.upload(start, 0, mpz_size(start));
    } else {
        mpz_t t;
        mpz_init(t);
        reciprocal_start(t, Dh, Ld, c);
        Y = g.upload(t, 0, mpz_size(t));
        mpz_clear(t);
    }
    for (auto it = sched.rbegin(); it != sched.rend(); ++it) {
        const size_t cp = *it;
        const size_t l = std::min(Ld, cp + kGuard);
        const Num Dp{D.d + (Ld - l), l};
        Y = newton_step(g, Y, g.mul(Dp, Y), l + c, c, cp, false);  // D' Y ~ beta^K, K = l + c
        c = cp;
    }
    return Y;
}

// The top `limbs` limbs of x: (first limb, count).
std::pair<size_t, size_t> top_limbs(mpz_srcptr x, size_t limbs) {
    const size_t n = mpz_size(x);
    return n > limbs ? std::make_pair(n - limbs, limbs) : std::make_pair((size_t)0, n);
}

}  // namespace

void device_inv_sqrt_limbs(mpz_t y, unsigned long a, size_t B) {
    if (a == 0) throw std::runtime_error("device inverse square root of 0");
    std::vector<size_t> sched;
    for (size_t c = B; c > kExactLimbs; c = (c + 3) / 2) sched.push_back(c);
    size_t c = sched.empty() ? B : (sched.back() + 3) / 2;
    while (!sched.empty() && sched.back() <= kHostLimbs) {
        c = sched.back();
        sched.pop_back();
    }
    Dev g(16 * (B + 4) + 64);
    Num Y, ka = g.take(1);
    {
        const uint64_t av = a;
        const uint64_t* h = &av;
        size_t one = 1;
        dev::check(gpu_ntt_upload(1, &ka.d, &h, &one), "gpu_ntt_upload");
        mpz_t t;
        mpz_init(t);
        newton_inv_sqrt_limbs(t, a, c);  // the host steps up to kHostLimbs
        Y = g.upload(t, 0, mpz_size(t));
        mpz_clear(t);
    }
    for (auto it = sched.rbegin(); it != sched.rend(); ++it) {
        const size_t cp = *it;
        Y = newton_step(g, Y, g.mul(g.mul(Y, Y), ka), 2 * c, c, cp, true);  // a Y^2 ~ beta^(2c)
        c = cp;
    }
    uint64_t* h = reinterpret_cast<uint64_t*>(mpz_limbs_write(y, (mp_size_t)Y.n));
    const uint64_t* src = Y.d;
    size_t n = Y.n;
    dev::check(gpu_ntt_download(1, &h, &src, &n), "gpu_ntt_download");
    mpz_limbs_finish(y, (mp_size_t)Y.n);
}

void device_reciprocal_start(mpz_t t, mpz_srcptr d, size_t B, size_t host_limbs) {
    if (mpz_sgn(d) <= 0) throw std::runtime_error("device reciprocal: divisor must be positive");
    size_t c = 0;
    (void)reciprocal_schedule(B, &c, host_limbs);
    reciprocal_start(t, d, mpz_size(d), c);
}

void device_reciprocal_limbs(mpz_t y, mpz_srcptr d, size_t B, mpz_srcptr start, size_t host_limbs) {
    if (mpz_sgn(d) <= 0) throw std::runtime_error("device reciprocal: divisor must be positive");
    const size_t Ld = mpz_size(d);
    size_t c = 0;
    if (reciprocal_schedule(B, &c, host_limbs).empty() && start) {
        // Every step ran on the host: the start is the reciprocal.
        mpz_set(y, start);
        return;
    }
    Dev g(8 * (Ld + B) + 64);
    const Num D = g.upload(d, 0, Ld);
    const Num Y = reciprocal(g, D, d, B, start, host_limbs);
    uint64_t* h = reinterpret_cast<uint64_t*>(mpz_limbs_write(y, (mp_size_t)Y.n));
    const uint64_t* src = Y.d;
    size_t n = Y.n;
    dev::check(gpu_ntt_download(1, &h, &src, &n), "gpu_ntt_download");
    mpz_limbs_finish(y, (mp_size_t)Y.n);
}

namespace {

// The final step from Q (its top limbs on the device, q_first limbs below
// them dropped), R likewise, and the top limbs of that R on the host.
void final_on_device(Dev& g, Num Qd, size_t q_first, Num Rd, size_t r_first, mpz_srcptr r_top, size_t w,
                     const FinalConstants& pre, mpz_ptr out, DeviceInt* dout) {
    const size_t W = w / 64 + 2;  // working limbs
    long ex = 0;  // value = integer * beta^ex, before the 2^w of the square root's scale
    // 426880 * sqrt(10005) = 426880 * 10005 * y / 2^w, y = pre.inv_sqrt
    const auto y = top_limbs(pre.inv_sqrt, W), t10 = top_limbs(pre.pow10, W);
    const Num yd = g.upload(pre.inv_sqrt, y.first, y.second);
    const Num pd = g.upload(pre.pow10, t10.first, t10.second);
    ex += (long)(q_first + y.first + t10.first);
    auto cut = [&](Num x) {  // the top W limbs of x; the exponent follows
        if (x.n <= W) return x;
        ex += (long)(x.n - W);
        return Num{x.d + (x.n - W), W};
    };
    Num k = g.take(1);
    {
        const uint64_t v = 426880ull * 10005ull;
        const uint64_t* h = &v;
        size_t one = 1;
        dev::check(gpu_ntt_upload(1, &k.d, &h, &one), "gpu_ntt_upload");
    }
    Num t = cut(g.mul(Qd, yd));
    t = cut(g.mul(t, pd));
    t = cut(g.mul(t, k));
    const Num inv = reciprocal(g, Rd, r_top, W);  // 1/R ~ inv / beta^(Rd.n + W + r_first)
    const Num z = g.mul(t, inv);
    ex -= (long)(Rd.n + W + r_first) + (long)(w / 64);
    // floor(z * beta^ex): ex < 0 drops limbs.
    if (ex >= 0) throw std::runtime_error("device final step: unexpected scale");
    const size_t drop = (size_t)(-ex);
    if (drop >= z.n) throw std::runtime_error("device final step: empty result");
    const size_t n = z.n - drop;
    if (dout) {  // into a buffer of its own (the arena goes with g)
        dout->x.n = n;
        dout->x.d = dout->mem.alloc(n);
        dev::check(gpu_ntt_dev_copy(dout->x.d, z.d + drop, n), "gpu_ntt_dev_copy");
        dev::check(gpu_ntt_synchronize(), "gpu_ntt_synchronize");
        return;
    }
    uint64_t* h = reinterpret_cast<uint64_t*>(mpz_limbs_write(out, (mp_size_t)n));
    const uint64_t* src = z.d + drop;
    size_t nn = n;
    dev::check(gpu_ntt_download(1, &h, &src, &nn), "gpu_ntt_download");
    mpz_limbs_finish(out, (mp_size_t)n);
}

void check_final_args(size_t w, const FinalConstants& pre) {
    if (w % 64 != 0) throw std::runtime_error("device final step: precision not a whole number of limbs");
    if (mpz_sgn(pre.inv_sqrt) <= 0 || mpz_sgn(pre.pow10) <= 0)
        throw std::runtime_error("device final step: constants must be positive");
}

}  // namespace

void pi_scaled_newton_device(mpz_t out, mpz_srcptr Q, mpz_srcptr R, size_t w, const FinalConstants& pre) {
    check_final_args(w, pre);
    if (mpz_sgn(Q) <= 0 || mpz_sgn(R) <= 0) throw std::runtime_error("device final step: operands must be positive");
    const size_t W = w / 64 + 2;
    Dev g(40 * W + 64);
    const auto q = top_limbs(Q, W), r = top_limbs(R, W);
    const Num Qd = g.upload(Q, q.first, q.second);
    const Num Rd = g.upload(R, r.first, r.second);
    mpz_t rh;
    mpz_roinit_n(rh, mpz_limbs_read(R) + r.first, (mp_size_t)r.second);
    final_on_device(g, Qd, q.first, Rd, r.first, rh, w, pre, out, nullptr);
}

namespace {
void final_from_qr(const DeviceQR& qr, size_t w, const FinalConstants& pre, mpz_ptr out, DeviceInt* dout) {
    check_final_args(w, pre);
    if (qr.r_sign < 0) throw std::runtime_error("device final step: R must be positive");
    const size_t W = w / 64 + 2;
    Dev g(36 * W + 64);
    // Views of the top W limbs (Q and R carry exact lengths: top limbs nonzero).
    const size_t qf = qr.Q.n > W ? qr.Q.n - W : 0, rf = qr.R.n > W ? qr.R.n - W : 0;
    const Num Qd{qr.Q.d + qf, qr.Q.n - qf};
    const Num Rd{qr.R.d + rf, qr.R.n - rf};
    // The top limbs of R that the host steps of the reciprocal read.
    const size_t k = std::min(Rd.n, kHostLimbs + kGuard + 1);
    const std::vector<uint64_t> top = g.fetch(Rd, Rd.n - k, k);
    if (top.back() == 0) throw std::runtime_error("device final step: R has a zero top limb");
    mpz_t rh;
    mpz_roinit_n(rh, reinterpret_cast<const mp_limb_t*>(top.data()), (mp_size_t)k);
    final_on_device(g, Qd, qf, Rd, rf, rh, w, pre, out, dout);
}
}  // namespace

void pi_scaled_newton_device(mpz_t out, const DeviceQR& qr, size_t w, const FinalConstants& pre) {
    final_from_qr(qr, w, pre, out, nullptr);
}

void pi_scaled_newton_device(DeviceInt& out, const DeviceQR& qr, size_t w, const FinalConstants& pre) {
    final_from_qr(qr, w, pre, nullptr, &out);
}

}  // namespace pi
