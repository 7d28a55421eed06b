/*
 * ntt_oracle.c - CPU restatement of the three-prime NTT product.
 *
 * TEST INFRASTRUCTURE ONLY (see ntt_oracle.h).  Written for clarity, not
 * speed: 64-bit '%' for every modular product, one radix-2 layer at a time.
 *
 * Reference anchors:
 *   pipeline order ............ src/gpu/int_ntt.cu:301-380 (IntNttEngine::multiply)
 *   modular product ........... src/gpu/int_ntt.cu:47-51 (dev_mod_mul: (a*b)%p)
 *   modular power ............. src/gpu/int_ntt.cu:32-41 (host_mod_pow)
 *   butterfly layers .......... src/gpu/int_ntt.cu:57-73,192-220 (radix-2 layers)
 *   pointwise product ......... src/gpu/int_ntt.cu:75-80
 *   CRT ....................... src/gpu/int_ntt.cu:226-245 (Garner, 3 primes)
 *   carry ..................... src/gpu/int_ntt.cu:352-377
 *   unpack/pack ............... src/arithmetic/int_ntt_multiplier.cpp:61-113
 *                                (the restatement uses 32-bit digits: base 2^32)
 */
#include "ntt_oracle.h"

#include <stdlib.h>
#include <string.h>

/* Prime index i: set i / 3, prime i % 3.  Set 0 (wide): transforms up to 2^25;
 * set 1 (narrow, every p < 2^29, so 8p < 2^32): transforms up to 2^22. */
static const uint32_t P[6] = {2013265921u, 1811939329u, 2113929217u, 469762049u, 415236097u, 377487361u};
static const uint32_t G[6] = {31u, 13u, 5u, 3u, 5u, 7u};
static const int TWO_ADIC[6] = {27, 26, 25, 26, 22, 23};

int orc_set_for(int log2n) { return log2n <= ORC_NARROW_MAX_LOG2N ? 1 : 0; }

uint32_t orc_prime(int i) { return P[i]; }
uint32_t orc_generator(int i) { return G[i]; }
int orc_two_adicity(int i) { return TWO_ADIC[i]; }

uint32_t orc_mulmod(uint32_t a, uint32_t b, uint32_t p) {
    return (uint32_t)(((uint64_t)a * b) % p);
}

uint32_t orc_powmod(uint32_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1 % p, x = a % p;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

uint32_t orc_invmod(uint32_t a, uint32_t p) { return orc_powmod(a, (uint64_t)p - 2, p); }

uint32_t orc_root(int prime, int log2n) {
    uint32_t p = P[prime];
    return orc_powmod(G[prime], ((uint64_t)p - 1) >> log2n, p);
}

static uint32_t bitrev(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

void orc_zeta_table(int prime, int log2n, uint32_t* z, uint32_t* zinv) {
    uint32_t p = P[prime];
    z[0] = 0;
    if (zinv) zinv[0] = 0;
    for (int s = 0; s < log2n; ++s) {
        uint32_t w = orc_root(prime, s + 1); /* primitive 2^(s+1)-th root */
        for (uint32_t b = 0; b < (1u << s); ++b) {
            uint32_t v = orc_powmod(w, bitrev(b, s), p);
            z[(1u << s) + b] = v;
            if (zinv) zinv[(1u << s) + b] = orc_invmod(v, p);
        }
    }
}

void orc_dft_textbook(const uint32_t* x, uint32_t* X, int log2n, int prime) {
    uint32_t p = P[prime];
    size_t n = (size_t)1 << log2n;
    uint32_t w = orc_root(prime, log2n);
    for (size_t k = 0; k < n; ++k) {
        uint32_t wk = orc_powmod(w, k, p), acc = 0, wjk = 1;
        for (size_t j = 0; j < n; ++j) {
            acc = (uint32_t)((acc + (uint64_t)x[j] * wjk) % p);
            wjk = orc_mulmod(wjk, wk, p);
        }
        X[k] = acc;
    }
}

/* Layer s splits every block of size N/2^s (block b holds x mod X^m - c_b with
 * m = N/2^(s+1)) into its two halves with the block's twiddle zeta:
 *   (lo, hi) -> (lo + zeta*hi, lo - zeta*hi). */
void orc_ntt_ct_layers(uint32_t* x, int log2n, int prime, int s_lo, int s_hi) {
    uint32_t p = P[prime];
    size_t n = (size_t)1 << log2n;
    uint32_t* z = (uint32_t*)malloc(n * sizeof(uint32_t));
    orc_zeta_table(prime, log2n, z, NULL);
    for (int s = s_lo; s < s_hi; ++s) {
        size_t half = n >> (s + 1);
        for (size_t b = 0; b < ((size_t)1 << s); ++b) {
            uint32_t zeta = z[((size_t)1 << s) + b];
            uint32_t* blk = x + b * 2 * half;
            for (size_t j = 0; j < half; ++j) {
                uint32_t lo = blk[j];
                uint32_t t = orc_mulmod(blk[j + half], zeta, p);
                blk[j] = (uint32_t)(((uint64_t)lo + t) % p);
                blk[j + half] = (uint32_t)(((uint64_t)lo + p - t) % p);
            }
        }
    }
    free(z);
}

/* Inverse of one layer up to a factor 2: (u, v) -> (u + v, (u - v) / zeta). */
void orc_ntt_gs_layers(uint32_t* x, int log2n, int prime, int s_lo, int s_hi) {
    uint32_t p = P[prime];
    size_t n = (size_t)1 << log2n;
    uint32_t* z = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint32_t* zi = (uint32_t*)malloc(n * sizeof(uint32_t));
    orc_zeta_table(prime, log2n, z, zi);
    for (int s = s_hi - 1; s >= s_lo; --s) {
        size_t half = n >> (s + 1);
        for (size_t b = 0; b < ((size_t)1 << s); ++b) {
            uint32_t zinv = zi[((size_t)1 << s) + b];
            uint32_t* blk = x + b * 2 * half;
            for (size_t j = 0; j < half; ++j) {
                uint32_t u = blk[j], v = blk[j + half];
                blk[j] = (uint32_t)(((uint64_t)u + v) % p);
                blk[j + half] = orc_mulmod((uint32_t)(((uint64_t)u + p - v) % p), zinv, p);
            }
        }
    }
    free(z);
    free(zi);
}

/* Garner: x = y0 + p0*y1 + p0*p1*y2 with y_i in [0, p_i). */
void orc_crt3_set(int set, const uint32_t r[3], uint32_t out[3]) {
    uint32_t p0 = P[3 * set], p1 = P[3 * set + 1], p2 = P[3 * set + 2];
    uint32_t y0 = r[0] % p0;
    uint32_t inv_p0_m1 = orc_invmod(p0 % p1, p1);
    uint32_t inv_p0_m2 = orc_invmod(p0 % p2, p2);
    uint32_t inv_p1_m2 = orc_invmod(p1 % p2, p2);
    uint32_t y1 = orc_mulmod((uint32_t)(((uint64_t)r[1] + p1 - y0 % p1) % p1), inv_p0_m1, p1);
    uint32_t t = orc_mulmod((uint32_t)(((uint64_t)r[2] + p2 - y0 % p2) % p2), inv_p0_m2, p2);
    uint32_t y2 = orc_mulmod((uint32_t)(((uint64_t)t + p2 - y1 % p2) % p2), inv_p1_m2, p2);
    unsigned __int128 x = (unsigned __int128)y2 * p1 + y1;
    x = x * p0 + y0;
    out[0] = (uint32_t)x;
    out[1] = (uint32_t)(x >> 32);
    out[2] = (uint32_t)(x >> 64);
}

void orc_crt3(const uint32_t r[3], uint32_t out[3]) { orc_crt3_set(0, r, out); }

/* ---- Lazy (Harvey) butterflies of the narrow set ------------------------- */

uint32_t orc_shoup_quotient(uint32_t w, uint32_t p) { return (uint32_t)(((uint64_t)w << 32) / p); }

static uint32_t min_u32(uint32_t a, uint32_t b) { return a < b ? a : b; }

uint32_t orc_narrow_load(uint32_t c, uint32_t p) { return min_u32(c, c - 8u * p); }

/* x plus and minus Shoup's w b (-(w b - q p) mod 2^32 kept negated). */
static void narrow_ct_core(uint32_t x, uint32_t* a, uint32_t* b, uint32_t w, uint32_t p) {
    const uint32_t wq = orc_shoup_quotient(w, p);
    const uint32_t q = (uint32_t)(((uint64_t)*b * wq) >> 32);
    const uint32_t tn = q * p - *b * w;
    *a = x - tn;
    *b = x + tn + 2u * p;
}

void orc_narrow_ct(uint32_t* a, uint32_t* b, uint32_t w, uint32_t p) {
    narrow_ct_core(min_u32(*a, *a - 4u * p), a, b, w, p);
}

void orc_narrow_ct_lazy(uint32_t* a, uint32_t* b, uint32_t w, uint32_t p) { narrow_ct_core(*a, a, b, w, p); }

/* a, b in [0, 4p): the sum reduced from [0, 8p) to [0, 4p); a - b + 4p in
 * (0, 8p) by Shoup to [0, 2p). */
void orc_narrow_gs(uint32_t* a, uint32_t* b, uint32_t winv, uint32_t p) {
    const uint32_t p4 = 4u * p, wq = orc_shoup_quotient(winv, p);
    const uint32_t u = *a, v = *b;
    const uint32_t s = u + v;
    const uint32_t d = u - v + p4;
    const uint32_t q = (uint32_t)(((uint64_t)d * wq) >> 32);
    *a = min_u32(s, s - p4);
    *b = d * winv - q * p;
}

/* a, b in [0, 2p): the sum kept in [0, 4p); the difference as above. */
void orc_narrow_gs_lazy(uint32_t* a, uint32_t* b, uint32_t winv, uint32_t p) {
    const uint32_t wq = orc_shoup_quotient(winv, p);
    const uint32_t u = *a, v = *b;
    const uint32_t d = u - v + 4u * p;
    const uint32_t q = (uint32_t)(((uint64_t)d * wq) >> 32);
    *a = u + v;
    *b = d * winv - q * p;
}

/* The shared operand of a row's products (b of a0 b and a1 b): y in [0, 8p)
 * reduced once to [0, 2p); each x is then reduced only to [0, 4p). */
uint32_t orc_narrow_shared_operand(uint32_t y, uint32_t p) {
    y = min_u32(y, y - 4u * p);
    return min_u32(y, y - 2u * p);
}

uint32_t orc_narrow_pointwise_shared(uint32_t x, uint32_t ys, uint32_t p) {
    uint32_t pinv = p; /* -p^-1 mod 2^32 by Newton's iteration */
    for (int i = 0; i < 5; ++i) pinv *= 2u - p * pinv;
    pinv = 0u - pinv;
    x = min_u32(x, x - 4u * p);
    const uint64_t t = (uint64_t)x * ys;
    const uint32_t m = (uint32_t)t * pinv;
    return (uint32_t)((t + (uint64_t)m * p) >> 32);
}

uint32_t orc_narrow_pointwise(uint32_t x, uint32_t y, uint32_t p) {
    const uint32_t p2 = 2u * p, p4 = 4u * p;
    uint32_t pinv = p; /* -p^-1 mod 2^32 by Newton's iteration */
    for (int i = 0; i < 5; ++i) pinv *= 2u - p * pinv;
    pinv = 0u - pinv;
    x = min_u32(x, x - p4);
    x = min_u32(x, x - p2);
    y = min_u32(y, y - p4);
    y = min_u32(y, y - p2);
    const uint64_t t = (uint64_t)x * y;
    const uint32_t m = (uint32_t)t * pinv;
    return (uint32_t)((t + (uint64_t)m * p) >> 32);
}

void orc_narrow_ct_layers(uint32_t* x, int log2n, int prime, int s_lo, int s_hi) {
    uint32_t p = P[prime];
    size_t n = (size_t)1 << log2n;
    uint32_t* z = (uint32_t*)malloc(n * sizeof(uint32_t));
    orc_zeta_table(prime, log2n, z, NULL);
    int bound = 8; /* inputs below 8p; a layer skips the reduction while every value is below 6p */
    for (int s = s_lo; s < s_hi; ++s) {
        size_t half = n >> (s + 1);
        const int lazy = bound <= 6;
        for (size_t b = 0; b < ((size_t)1 << s); ++b)
            for (size_t j = 0; j < half; ++j) {
                uint32_t* u = &x[b * 2 * half + j];
                uint32_t* v = &x[b * 2 * half + j + half];
                if (lazy) orc_narrow_ct_lazy(u, v, z[((size_t)1 << s) + b], p);
                else orc_narrow_ct(u, v, z[((size_t)1 << s) + b], p);
            }
        bound = lazy ? bound + 2 : 6;
    }
    free(z);
}

void orc_narrow_gs_layers(uint32_t* x, int log2n, int prime, int s_lo, int s_hi) {
    uint32_t p = P[prime];
    size_t n = (size_t)1 << log2n;
    uint32_t* z = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint32_t* zi = (uint32_t*)malloc(n * sizeof(uint32_t));
    orc_zeta_table(prime, log2n, z, zi);
    /* Bound of every value in units of p: 2 while it is a difference (or an
     * input below 2p), 4 otherwise; a butterfly of two values below 2p keeps
     * its sum unreduced. */
    unsigned char* bnd = (unsigned char*)malloc(n);
    for (size_t i = 0; i < n; ++i) bnd[i] = x[i] < 2u * p ? 2 : 4;
    for (int s = s_hi - 1; s >= s_lo; --s) {
        size_t half = n >> (s + 1);
        for (size_t b = 0; b < ((size_t)1 << s); ++b)
            for (size_t j = 0; j < half; ++j) {
                const size_t iu = b * 2 * half + j, iv = iu + half;
                if (bnd[iu] == 2 && bnd[iv] == 2) {
                    orc_narrow_gs_lazy(&x[iu], &x[iv], zi[((size_t)1 << s) + b], p);
                } else {
                    orc_narrow_gs(&x[iu], &x[iv], zi[((size_t)1 << s) + b], p);
                }
                bnd[iu] = 4;
                bnd[iv] = 2;
            }
    }
    free(bnd);
    free(z);
    free(zi);
}

size_t orc_coeff_count(const uint64_t* a, size_t n) {
    if (n == 0) return 0;
    return 2 * n - ((a[n - 1] >> 32) == 0 ? 1 : 0);
}

int orc_transform_log2(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, int min_log2) {
    size_t need = orc_coeff_count(a, an) + orc_coeff_count(b, bn) - 1;
    int k = min_log2;
    while (((size_t)1 << k) < need) ++k;
    return k;
}

static uint32_t coeff(const uint64_t* a, size_t n, size_t k) {
    size_t limb = k >> 1;
    if (limb >= n) return 0;
    return (uint32_t)(a[limb] >> (32 * (k & 1)));
}

void orc_conv_residues(const uint64_t* a, size_t an, const uint64_t* b, size_t bn,
                       int log2n, uint32_t* res) {
    size_t n = (size_t)1 << log2n;
    uint32_t* fb = (uint32_t*)malloc(n * sizeof(uint32_t));
    const int base = 3 * orc_set_for(log2n);
    for (int i = 0; i < 3; ++i) {
        uint32_t p = P[base + i];
        uint32_t* fa = res + (size_t)i * n;
        for (size_t k = 0; k < n; ++k) {
            fa[k] = coeff(a, an, k) % p;
            fb[k] = coeff(b, bn, k) % p;
        }
        orc_ntt_ct_layers(fa, log2n, base + i, 0, log2n);
        orc_ntt_ct_layers(fb, log2n, base + i, 0, log2n);
        for (size_t k = 0; k < n; ++k) fa[k] = orc_mulmod(fa[k], fb[k], p);
        orc_ntt_gs_layers(fa, log2n, base + i, 0, log2n);
        uint32_t ninv = orc_invmod((uint32_t)(n % p), p);
        for (size_t k = 0; k < n; ++k) fa[k] = orc_mulmod(fa[k], ninv, p);
    }
    free(fb);
}

int orc_ntt_mul(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, uint64_t* out) {
    memset(out, 0, (an + bn) * sizeof(uint64_t));
    if (an == 0 || bn == 0) return 0;
    int log2n = orc_transform_log2(a, an, b, bn, 1);
    if (log2n > ORC_MAX_LOG2N) return -1;
    size_t n = (size_t)1 << log2n;
    uint32_t* res = (uint32_t*)malloc(3 * n * sizeof(uint32_t));
    orc_conv_residues(a, an, b, bn, log2n, res);
    /* Sum the 96-bit convolution values at 32-bit offsets with carry. */
    size_t nwords = 2 * (an + bn);
    uint32_t* w = (uint32_t*)out; /* little-endian host: limb i = words 2i, 2i+1 */
    unsigned __int128 acc = 0;
    for (size_t j = 0; j < nwords; ++j) {
        /* word j collects x_j, x_{j-1} >> 32, x_{j-2} >> 64; a running 128-bit
         * accumulator with the carry shifted down each step does exactly that. */
        if (j < n) {
            uint32_t r[3] = {res[j], res[n + j], res[2 * n + j]}, x[3];
            orc_crt3_set(orc_set_for(log2n), r, x);
            acc += ((unsigned __int128)x[2] << 64) | ((unsigned __int128)x[1] << 32) | x[0];
        }
        w[j] = (uint32_t)acc;
        acc >>= 32;
    }
    free(res);
    return acc == 0 ? 0 : -2;
}
