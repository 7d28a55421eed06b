/*
 * ntt_oracle.h - CPU restatement of the exact big-integer product path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py use it, and only as the checker.
 *
 * What it restates (reference pipeline order, src/gpu/int_ntt.cu:301-380):
 *   limbs -> coefficients -> residues mod three NTT primes -> forward NTT ->
 *   pointwise product -> inverse NTT -> CRT -> carry -> limbs.
 * The reference uses base-2^15 digits, 64-bit '%' arithmetic, primes of 2-adic
 * order 23/23/24 and a host CRT (src/gpu/int_ntt.cu:226-245,352-377).  The
 * restatement follows the product the engine computes instead: 32-bit
 * coefficients; two prime sets, the wide set 15*2^27+1, 27*2^26+1, 63*2^25+1
 * for transforms of 2^23 .. 2^25 points and the narrow set 7*2^26+1,
 * 99*2^22+1, 45*2^23+1 (all below 2^29) up to 2^22 points; a Cooley-Tukey
 * forward transform (natural order in, bit-reversed out) and a Gentleman-Sande
 * inverse, Garner CRT to an exact 96-bit value and a sequential carry.  The
 * narrow set's lazy butterflies (values in [0, 8p) forward, [0, 2p) inverse)
 * are restated separately, for range and congruence checks.  The whole product is pinned against GMP mpz_mul
 * (GmpMultiplier, src/arithmetic/gmp_multiplier.cpp:10-19).
 */
#ifndef PINTT_ORACLE_H
#define PINTT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NPRIMES 3
#define ORC_MAX_LOG2N 25
#define ORC_NARROW_MAX_LOG2N 22

/* Prime index i in 0..5: set i / 3 (0 wide, 1 narrow), prime i % 3.
 * p_i, generator g_i of Z_p^*, and the 2-adic order of p_i - 1. */
uint32_t orc_prime(int i);
uint32_t orc_generator(int i);
int orc_two_adicity(int i);

uint32_t orc_mulmod(uint32_t a, uint32_t b, uint32_t p);
uint32_t orc_powmod(uint32_t a, uint64_t e, uint32_t p);
uint32_t orc_invmod(uint32_t a, uint32_t p);
/* Primitive 2^log2n-th root of unity modulo prime i: g^((p-1)/2^log2n). */
uint32_t orc_root(int prime, int log2n);

/* Twiddle table in Cooley-Tukey order: z[2^s + b] = w_{2^(s+1)}^{bitrev_s(b)}
 * for 0 <= s < log2n, 0 <= b < 2^s; zinv holds the inverses.  Entry 0 unused.
 * Arrays hold 2^log2n entries. */
void orc_zeta_table(int prime, int log2n, uint32_t* z, uint32_t* zinv);

/* Textbook O(N^2) DFT over Z_p: X[k] = sum_j x[j] w^(jk), natural order. */
void orc_dft_textbook(const uint32_t* x, uint32_t* X, int log2n, int prime);

/* Forward Cooley-Tukey layers s in [s_lo, s_hi) on x (length 2^log2n, values < p).
 * All layers [0, log2n): x[b] becomes X[bitrev(b)]. */
void orc_ntt_ct_layers(uint32_t* x, int log2n, int prime, int s_lo, int s_hi);
/* Inverse Gentleman-Sande layers s = s_hi-1 down to s_lo (unscaled: the full
 * inverse returns N * x). */
void orc_ntt_gs_layers(uint32_t* x, int log2n, int prime, int s_lo, int s_hi);

/* Prime set serving a transform of 2^log2n points (1: narrow, 0: wide). */
int orc_set_for(int log2n);

/* Garner CRT of residues (r0 mod p0, r1 mod p1, r2 mod p2) of a set to the
 * unique x < p0*p1*p2, written as three little-endian 32-bit words.
 * orc_crt3: the wide set. */
void orc_crt3_set(int set, const uint32_t r[3], uint32_t out[3]);
void orc_crt3(const uint32_t r[3], uint32_t out[3]);

/* Lazy butterflies of a narrow prime p < 2^29 with Shoup's quotient
 * floor(w 2^32 / p).  ct: a in [0, 8p), any b -> (a + w b, a - w b) mod p,
 * both in [0, 6p) (a reduced to [0, 4p) first).  ct_lazy: a in [0, 6p), any b
 * -> the same congruences, both in [0, 8p) (no reduction).  gs: a, b in
 * [0, 4p) -> (a + b, (a - b) winv) mod p in [0, 4p) and [0, 2p).  gs_lazy:
 * a, b in [0, 2p) -> the same, the sum unreduced in [0, 4p).  pointwise: x, y in
 * [0, 8p) -> x y 2^-32 mod p in [0, 2p).  shared_operand: y in [0, 8p) ->
 * y mod p in [0, 2p); pointwise_shared: x in [0, 8p), ys in [0, 2p) -> x ys
 * 2^-32 mod p in [0, 2p) (x reduced only to [0, 4p)).  load: any 32-bit c -> c mod p in
 * [0, 8p).  The layer functions run whole layers as orc_ntt_ct_layers /
 * orc_ntt_gs_layers do, with unreduced values; the forward ones start below
 * 8p and skip the reduction while every value is below 6p (every other layer),
 * as the device's register stages do; the inverse ones take inputs below 4p
 * and keep the sum of two values below 2p unreduced. */
uint32_t orc_shoup_quotient(uint32_t w, uint32_t p);
void orc_narrow_ct(uint32_t* a, uint32_t* b, uint32_t w, uint32_t p);
void orc_narrow_ct_lazy(uint32_t* a, uint32_t* b, uint32_t w, uint32_t p);
void orc_narrow_gs(uint32_t* a, uint32_t* b, uint32_t winv, uint32_t p);
void orc_narrow_gs_lazy(uint32_t* a, uint32_t* b, uint32_t winv, uint32_t p);
uint32_t orc_narrow_pointwise(uint32_t x, uint32_t y, uint32_t p);
uint32_t orc_narrow_shared_operand(uint32_t y, uint32_t p);
uint32_t orc_narrow_pointwise_shared(uint32_t x, uint32_t ys, uint32_t p);
uint32_t orc_narrow_load(uint32_t c, uint32_t p);
void orc_narrow_ct_layers(uint32_t* x, int log2n, int prime, int s_lo, int s_hi);
void orc_narrow_gs_layers(uint32_t* x, int log2n, int prime, int s_lo, int s_hi);

/* Number of 32-bit coefficients of an n-limb magnitude whose top limb is
 * non-zero (2n, or 2n-1 if the top limb's high half is zero). */
size_t orc_coeff_count(const uint64_t* a, size_t n);
/* log2 of the transform length used for a (an limbs) times b (bn limbs):
 * smallest N = 2^k >= ca + cb - 1, at least 2^min_log2. */
int orc_transform_log2(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, int min_log2);

/* Full product through the restated NTT path.  out receives an + bn limbs.
 * Returns 0, or -1 when the transform would exceed 2^25. */
int orc_ntt_mul(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, uint64_t* out);

/* Convolution residues, for intermediate-buffer checks: res[i*N + k] =
 * (sum_j a_j b_{k-j}) mod p_i with 32-bit coefficients, N = 2^log2n, over the
 * primes of the set that serves N. */
void orc_conv_residues(const uint64_t* a, size_t an, const uint64_t* b, size_t bn,
                       int log2n, uint32_t* res);

/* GMP oracle (GmpMultiplier::multiply == mpz_mul, src/arithmetic/gmp_multiplier.cpp:10-12)
 * on raw magnitudes.  out receives an + bn limbs.  Defined in gmp_oracle.c. */
void orc_gmp_mul(const uint64_t* a, size_t an, const uint64_t* b, size_t bn, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif
