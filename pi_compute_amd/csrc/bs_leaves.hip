// bs_leaves.hip - the bottom of the Chudnovsky binary-splitting tree on the
// device (gfx950): one workgroup per subtree of 64 terms, from the leaf terms
// up to the subtree's P, Q, R (the host's compute_small,
// src/engine/binary_splitting.cpp:137-160 in the reference).
//
// Leaf k: P = -(6k-5)(2k-1)(6k-1), Q = k^3 * 10939058860032000,
// R = P * (13591409 + 545140134 k); leaf 0: P = Q = 1, R = 13591409; terms
// past the end of the range are the identity P = Q = 1, R = 0.  Merge:
// P = Pl Pr, Q = Ql Qr, R = Qr Rl + Pl Rr.  Magnitudes in LDS slots of fixed
// capacity (3 limbs at a leaf, 2c + 1 after each merge) with the limbs in use
// tracked, so products cost what the operands hold; signs apart: R has
// the sign of t = Qr Rl, and when t and u = Pl Rr differ in sign |t| >= |u|
// (the series alternates with terms shrinking ~1.5e14-fold); a borrow out of
// t - u would contradict that and is reported.  Each level: every column of
// every product of the level summed by one thread (three-word accumulator),
// one thread per product carrying its columns, one thread per merge adding.
#include "ntt_device.h"
#include "bigint_add.h"

namespace pintt {
namespace {

constexpr int kT = kBsLeafThreads;
constexpr int kS = kBsLeafTerms;
constexpr int kL = kBsLeafLevels;
__host__ __device__ constexpr int cap_of(int level) { return (4 << level) - 1; }  // 3, 7, 15, ...
constexpr int kNodeWords = 3 * 4 * kS;          // >= (kS >> L) * 3 * cap_of(L) for every L
constexpr int kCols = 4 * 4 * kS;               // >= (kS / 2 >> L) * 4 * (2 cap_of(L) - 1)
constexpr int kProdWords = 4 * 4 * kS;          // >= (kS / 2 >> L) * 4 * 2 cap_of(L)
static_assert(cap_of(kL) == kBsLeafLimbs, "output capacity");

__device__ __forceinline__ void mul64(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
    lo = a * b;
    hi = __umul64hi(a, b);
}

__global__ void __launch_bounds__(kT) k_bs_leaves(uint64_t a, uint64_t b, uint64_t* __restrict__ out,
                                                  uint32_t* __restrict__ lengths, int32_t* __restrict__ signs,
                                                  uint32_t* __restrict__ error) {
    __shared__ uint64_t nodes[2][kNodeWords];
    __shared__ int sgn[2][2][kS];             // [buffer][P or R][node]
    __shared__ uint16_t len[2][3][kS];        // [buffer][P, Q or R][node]: limbs in use
    __shared__ uint16_t plen[4 * kS / 2];     // limbs of each product of a level
    __shared__ uint64_t C0[kCols], C1[kCols], C2[kCols];
    __shared__ uint64_t prod[kProdWords];
    const int tid = threadIdx.x;
    const uint64_t k0 = a + (uint64_t)blockIdx.x * kS;
    // Leaves.
    if (tid < kS) {
        const uint64_t k = k0 + tid;
        uint64_t* P = &nodes[0][tid * 3 * cap_of(0)];
        uint64_t* Q = P + cap_of(0);
        uint64_t* R = Q + cap_of(0);
        for (int i = 0; i < 3 * cap_of(0); ++i) P[i] = 0;
        int sp = 1, sr = 1;
        if (k >= b) {
            P[0] = 1;
            Q[0] = 1;
        } else if (k == 0) {
            P[0] = 1;
            Q[0] = 1;
            R[0] = 13591409;
        } else {
            // (6k-5)(2k-1) < 2^50 for k < 2^24; times (6k-1) < 2^27.
            const uint64_t f = (6 * k - 5) * (2 * k - 1);
            mul64(f, 6 * k - 1, P[0], P[1]);
            // k^3 < 2^72: k^2 (< 2^48) * k, then * 640320^3 / 24 (< 2^54).
            uint64_t q0, q1, x0, x1;
            mul64(k * k, k, q0, q1);                 // k^3 = q1 2^64 + q0
            mul64(q0, 10939058860032000ull, x0, x1);
            Q[0] = x0;
            uint64_t y0, y1;
            mul64(q1, 10939058860032000ull, y0, y1);
            Q[1] = x1 + y0;
            Q[2] = y1 + (Q[1] < x1 ? 1u : 0u);
            // R = |P| * (13591409 + 545140134 k), the multiplier < 2^53.
            const uint64_t m = 545140134ull * k + 13591409ull;
            uint64_t r0, r1, s0, s1;
            mul64(P[0], m, r0, r1);
            mul64(P[1], m, s0, s1);
            R[0] = r0;
            R[1] = r1 + s0;
            R[2] = s1 + (R[1] < r1 ? 1u : 0u);
            sp = -1;
            sr = -1;
        }
        sgn[0][0][tid] = sp;
        sgn[0][1][tid] = sr;
        for (int v = 0; v < 3; ++v) {
            const uint64_t* x = P + v * cap_of(0);
            int n = cap_of(0);
            while (n > 1 && x[n - 1] == 0) --n;
            len[0][v][tid] = (uint16_t)n;
        }
    }
    __syncthreads();
    int cur = 0;
    for (int L = 0; L < kL; ++L) {
        const int c = cap_of(L), C = 2 * c + 1;
        const int M = (kS / 2) >> L;  // merges
        const int ncol = 2 * c - 1;   // columns of one product
        const uint64_t* nd = nodes[cur];
        // Operands of product p of merge m: 0 Pl Pr, 1 Ql Qr, 2 Qr Rl, 3 Pl Rr
        // (value v of node: 0 P, 1 Q, 2 R).
        auto which = [](int p, int side, int& node, int& v) {
            const int nl[4] = {0, 0, 1, 0}, vl[4] = {0, 1, 1, 0};  // side 0: node offset, value
            const int nr[4] = {1, 1, 0, 1}, vr[4] = {0, 1, 2, 2};  // side 1
            node = side ? nr[p] : nl[p];
            v = side ? vr[p] : vl[p];
        };
        for (int idx = tid; idx < M * 4 * ncol; idx += kT) {
            const int m = idx / (4 * ncol), rem = idx % (4 * ncol), p = rem / ncol, j = rem % ncol;
            int nx_, vx, ny_, vy;
            which(p, 0, nx_, vx);
            which(p, 1, ny_, vy);
            const uint64_t* x = nd + (2 * m + nx_) * 3 * c + vx * c;
            const uint64_t* y = nd + (2 * m + ny_) * 3 * c + vy * c;
            const int lx = len[cur][vx][2 * m + nx_], ly = len[cur][vy][2 * m + ny_];
            uint64_t s0 = 0, s1 = 0, s2 = 0;
            const int lo = j >= ly ? j - ly + 1 : 0, hi = j < lx ? j : lx - 1;
            for (int i = lo; i <= hi; ++i) {
                uint64_t pl, ph;
                mul64(x[i], y[j - i], pl, ph);
                s0 += pl;
                const uint64_t c0 = s0 < pl ? 1u : 0u;
                const uint64_t t = s1 + ph;
                const uint64_t c1 = t < s1 ? 1u : 0u;
                s1 = t + c0;
                s2 += c1 + (s1 < c0 ? 1u : 0u);
            }
            C0[idx] = s0;
            C1[idx] = s1;
            C2[idx] = s2;
        }
        __syncthreads();
        for (int q = tid; q < M * 4; q += kT) {  // product q: lx + ly <= 2c limbs
            const int m = q / 4, p = q % 4;
            int nx_, vx, ny_, vy;
            which(p, 0, nx_, vx);
            which(p, 1, ny_, vy);
            const int n = len[cur][vx][2 * m + nx_] + len[cur][vy][2 * m + ny_] - 1;  // columns in use
            uint64_t c0 = 0, c1 = 0;
            uint64_t* o = prod + q * 2 * c;
            for (int j = 0; j < n; ++j) {
                const int idx = q * ncol + j;
                const uint64_t t0 = C0[idx] + c0;
                const uint64_t k0c = t0 < c0 ? 1u : 0u;
                const uint64_t t1 = C1[idx] + c1;
                const uint64_t k1 = t1 < c1 ? 1u : 0u;
                const uint64_t u1 = t1 + k0c;
                const uint64_t k2 = u1 < k0c ? 1u : 0u;
                o[j] = t0;
                c0 = u1;
                c1 = C2[idx] + k1 + k2;
            }
            o[n] = c0;  // the product is below 2^(64 (lx + ly))
            plen[q] = (uint16_t)(n + 1);
        }
        __syncthreads();
        uint64_t* nx = nodes[cur ^ 1];
        for (int m = tid; m < M; m += kT) {
            const uint64_t* pp = prod + (m * 4 + 0) * 2 * c;
            const uint64_t* qq = prod + (m * 4 + 1) * 2 * c;
            const uint64_t* t = prod + (m * 4 + 2) * 2 * c;
            const uint64_t* u = prod + (m * 4 + 3) * 2 * c;
            const int lp = plen[m * 4 + 0], lq = plen[m * 4 + 1], lt = plen[m * 4 + 2], lu = plen[m * 4 + 3];
            uint64_t* P = nx + m * 3 * C;
            uint64_t* Q = P + C;
            uint64_t* R = Q + C;
            for (int i = 0; i < lp; ++i) P[i] = pp[i];
            for (int i = 0; i < lq; ++i) Q[i] = qq[i];
            const int st = sgn[cur][1][2 * m];                            // sign of t = Qr Rl
            const int su = sgn[cur][0][2 * m] * sgn[cur][1][2 * m + 1];  // sign of u = Pl Rr
            const int lr = lt > lu ? lt : lu;
            int nr;
            if (st == su) {
                uint64_t carry = 0;
                for (int i = 0; i < lr; ++i) {
                    const uint64_t a = i < lt ? t[i] : 0, b = i < lu ? u[i] : 0;
                    const uint64_t s = a + b;
                    const uint64_t r = s + carry;
                    carry = (s < a || r < s) ? 1u : 0u;
                    R[i] = r;
                }
                R[lr] = carry;
                nr = lr + 1;
            } else {
                uint64_t borrow = 0;
                for (int i = 0; i < lr; ++i) {
                    const uint64_t a = i < lt ? t[i] : 0, b = i < lu ? u[i] : 0;
                    const uint64_t d = a - b;
                    const uint64_t r = d - borrow;
                    borrow = (a < b || d < borrow) ? 1u : 0u;
                    R[i] = r;
                }
                nr = lr;
                if (borrow) atomicOr(error, 1u);  // |t| < |u|: the sign rule does not hold
            }
            int np = lp, nq = lq;
            while (np > 1 && P[np - 1] == 0) --np;
            while (nq > 1 && Q[nq - 1] == 0) --nq;
            while (nr > 1 && R[nr - 1] == 0) --nr;
            len[cur ^ 1][0][m] = (uint16_t)np;
            len[cur ^ 1][1][m] = (uint16_t)nq;
            len[cur ^ 1][2][m] = (uint16_t)nr;
            sgn[cur ^ 1][0][m] = sgn[cur][0][2 * m] * sgn[cur][0][2 * m + 1];
            sgn[cur ^ 1][1][m] = st;
        }
        __syncthreads();
        cur ^= 1;
    }
    // The subtree's P, Q, R: magnitudes, exact lengths and signs.
    const int cf = cap_of(kL);
    const uint64_t* nd = nodes[cur];
    uint64_t* o = out + (size_t)blockIdx.x * 3 * cf;
    for (int i = tid; i < 3 * cf; i += kT) o[i] = nd[i];
    if (tid < 3) lengths[blockIdx.x * 3 + tid] = len[cur][tid][0];
    if (tid == 0) {
        signs[blockIdx.x * 2 + 0] = sgn[cur][0][0];
        signs[blockIdx.x * 2 + 1] = sgn[cur][1][0];
    }
}

}  // namespace

hipError_t launch_bs_leaves(uint64_t a, uint64_t b, uint32_t count, uint64_t* out, uint32_t* lengths, int32_t* signs,
                            uint32_t* error, hipStream_t s) {
    hipLaunchKernelGGL(k_bs_leaves, dim3(count), dim3(kT), 0, s, a, b, out, lengths, signs, error);
    return hipGetLastError();
}

}  // namespace pintt
