// ntt_device.h - device building blocks shared by the NTT passes (gfx950).
//
// A workgroup owns a tile of SEQ sequences of length L in LDS; every thread
// keeps 16 elements in registers and runs up to four butterfly layers on them
// between LDS exchanges (radix-16 steps).  Forward layers are Cooley-Tukey
// (natural order in, bit-reversed out), inverse layers Gentleman-Sande; the
// twiddle of layer s, block b is Z[2^s + b] (see ntt_kernels.h).
#pragma once

#include "ntt_kernels.h"

namespace pintt {
template <int SET>
__device__ __forceinline__ uint32_t prime_of(int i) {
    return i == 0 ? kPrimeSet[SET][0] : (i == 1 ? kPrimeSet[SET][1] : kPrimeSet[SET][2]);
}

// Pointers read from job descriptors are generic to the compiler, which then
// emits flat loads (counted against both the vector-memory and the LDS wait
// counters, so every LDS wait also waits for them).  Device buffers here are
// always global memory: say so.
template <class T>
using global_ptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ global_ptr<T> as_global(T* p) {
    return (global_ptr<T>)p;
}
// Four words as one 128-bit access (a clang vector type: it loads and stores
// through address-space-qualified pointers, unlike uint4).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Twiddle table of one prime, picked with scalar selects: indexing the
// kernel-argument arrays with a run-time prime would copy them to scratch.
// A table entry (w, w') is read as one 64-bit word (w in the low half).
using TwTable = global_ptr<const unsigned long long>;
__device__ __forceinline__ TwTable pick(const uint2* t0, const uint2* t1, const uint2* t2, int prime) {
    return (TwTable)(prime == 0 ? t0 : (prime == 1 ? t1 : t2));
}
__device__ __forceinline__ uint2 tw_at(TwTable z, uint32_t i) {
    const unsigned long long v = z[i];
    return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
}
__device__ __forceinline__ uint2 tw_at(const uint2* z, uint32_t i) { return z[i]; }
// The same table read through the constant address space: the tables never
// change while a kernel runs, so a wave-uniform index becomes a scalar load
// (no vector registers held for twiddles shared by the whole wave).
using ConstTable = __attribute__((address_space(4))) const unsigned long long*;
__device__ __forceinline__ ConstTable as_const(TwTable z) { return (ConstTable)(const unsigned long long*)z; }
__device__ __forceinline__ uint2 tw_at(ConstTable z, uint32_t i) {
    const unsigned long long v = z[i];
    return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
}
#define PINTT_FWD_TABLE(tw, prime) pick((tw).fwd[0], (tw).fwd[1], (tw).fwd[2], (prime))
#define PINTT_INV_TABLE(tw, prime) pick((tw).inv[0], (tw).inv[1], (tw).inv[2], (prime))

// Row factors.  Of a transform of 2^n points as 2^Lr rows of 2^Lc elements,
// row r holds at global layer Lr + m the blocks r 2^m + j (j < 2^m), whose
// twiddles factor as Z[2^(Lr+m) + r 2^m + j] = Z[2^m + j] * a_r^(2^(Lc-1-m)),
// a_r = w_(2^n)^bitrev_Lr(r).  Layer m pairs elements 2^(Lc-1-m) apart, so
// scaling element c of the row by a_r^c first makes the row's layers those of
// block 0 (twiddles shared by every row: sb = row0 = 0 below), and the inverse
// layers of block 0 followed by a scaling by a_r^-c undo them.  a_r^c = wK^e,
// wK the set's 2^K-th root, e = bitrev_Lr(r) c 2^(K-n) < 2^K, comes from two
// tables: lo[i] = wK^i 2^32 mod p (i < 2^12, Montgomery form) and, after it,
// hi[i] = (wK^(i 2^12), its Shoup quotient); one Shoup product gives wK^e in
// Montgomery form, ready for mul_mont.
using RowTable = global_ptr<const uint32_t>;
__device__ __forceinline__ RowTable pick_row(const uint32_t* t0, const uint32_t* t1, const uint32_t* t2, int prime) {
    return (RowTable)(prime == 0 ? t0 : (prime == 1 ? t1 : t2));
}
__device__ __forceinline__ uint32_t row_factor(RowTable t, uint32_t e, uint32_t p) {
    const uint32_t lo = t[e & (kPowLo - 1)];
    const unsigned long long h = ((global_ptr<const unsigned long long>)(t + kPowLo))[e >> kPowLoBits];
    return mul_shoup(lo, (uint32_t)h, (uint32_t)(h >> 32), p);
}
#define PINTT_ROW_TABLE(tw, prime) pick_row((tw).row[0], (tw).row[1], (tw).row[2], (prime))

// Last index i < n with key(i) <= v, for keys sorted ascending and key(0) <= v,
// called by every lane of a wave: each round every lane reads one sampled key
// and a ballot keeps the last sample at or below v, so 4,096 entries take two
// dependent loads instead of a twelve-step binary search.
template <class Key>
__device__ __forceinline__ uint32_t wave_search(uint32_t n, uint32_t v, Key key) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t lo = 0, span = n;  // the answer lies in [lo, lo + span)
    while (span > 1) {
        const uint32_t step = (span + 63u) >> 6;
        const bool ok = lane * step < span && key(lo + lane * step) <= v;
        const uint64_t m = __ballot(ok);  // lane 0 is always set
        const uint32_t nlo = lo + (63u - (uint32_t)__builtin_clzll(m)) * step;
        span = min(step, lo + span - nlo);
        lo = nlo;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// Workgroup order of a forward column grid (column tiles, prime, job).  The
// dispatcher deals workgroups to the 8 XCDs round-robin in linear order
// (x fastest), and every XCD has its own L2.  The three primes of one
// (tile, job) read the same limbs, so workgroup L runs the (tile, prime, job)
// below: the pairs u = (tile, job), u = tile + nx * job, go in groups of 8,
// pair 8 q + x on XCD x, its primes 0, 1, 2 that XCD's workgroups 3 q .. 3 q
// + 2; the limbs come from HBM once and from that L2 twice.  The last
// (nx nz mod 8) pairs, fewer than 24 workgroups, go in linear order.  A
// bijection of [0, 3 nx nz) for every nx, nz (tests/test_col_blocks.py).
struct ColBlock {
    uint32_t tile, prime, job;
};
__host__ __device__ constexpr ColBlock prime_grouped_block(uint32_t L, uint32_t nx, uint32_t nz) {
    const uint32_t full = (nx * nz) >> 3;  // whole groups of 8 pairs (24 workgroups)
    uint32_t u = 0, prime = 0;
    if (L < 24 * full) {
        const uint32_t s = L >> 3;  // L's place on its XCD
        prime = s % 3;
        u = (s / 3) * 8 + (L & 7u);
    } else {
        const uint32_t r = L - 24 * full;
        prime = r % 3;
        u = 8 * full + r / 3;
    }
    return ColBlock{u % nx, prime, u / nx};
}

// Bank layout of a tile in LDS.  Linear position x of a tile sits at word
// x ^ T(x), T(x) = ((x >> s1) & m1) << t1  ^  ((x >> s2) & m2) << t2, which
// permutes the words of each line of 32 (T changes bits 0..4 only and is
// linear over GF(2)).  LDS serves a wave in two halves of 32 lanes over 32
// banks (bank = word mod 32); a radix-16 exchange has the lanes touch words a
// power-of-two stride apart, which the identity layout puts on few banks.  The
// terms below leave no bank conflict in any load, store or stage exchange of
// the kernels, for every compiled shape (counted lane by lane on a model of
// those accesses; a search over one- and two-term maps):
//   rows: (x >> 4) & 31 into bits 0..4, and a second term by log2(L) mod 4;
//   columns: the line number (x >> 5) into bits 0..4 as listed.
struct LdsMap {
    uint32_t s1, m1, t1, s2, m2, t2;
};
__host__ __device__ constexpr LdsMap lds_map(int logl, int logseq, bool cols) {
    if (!cols) {
        switch (logl & 3) {
            case 1: return {4, 31, 0, 4, 3, 3};
            case 2: return {4, 31, 0, 4, 7, 2};
            case 3: return {4, 31, 0, 4, 15, 1};
            default: return {4, 31, 0, 9, 1, 4};
        }
    }
    if ((logl == 7 && logseq <= 4) || (logl == 11 && logseq == 3)) return {5, 7, 2, 0, 0, 0};
    if (logl == 8 && logseq <= 4) return {5, 15, 1, 0, 0, 0};
    if (logl == 9 && logseq == 3) return {5, 15, 1, 5, 1, 3};
    if (logl == 9 && logseq == 4) return {5, 1, 4, 0, 0, 0};
    if ((logl == 6 && logseq == 4) || (logl == 10 && logseq <= 4)) return {5, 3, 3, 0, 0, 0};
    return {5, 1, 0, 0, 0, 0};
}

// A tile of SEQ sequences of length L in LDS.  COLS: the column pass, where
// element (seq w, index r) has linear position r * SEQ + w so that
// neighbouring threads touch neighbouring words; otherwise sequences are
// stored one after another.
template <int LOGL_, int LOGSEQ_, bool COLS_>
struct Tile {
    static constexpr int LOGL = LOGL_;
    static constexpr int LOGSEQ = LOGSEQ_;
    static constexpr bool COLS = COLS_;
    static constexpr int L = 1 << LOGL;
    static constexpr int SEQ = 1 << LOGSEQ;
    static constexpr int NT = L * SEQ / kElemsPerThread;
    static constexpr int WORDS = L * SEQ;
    static constexpr uint32_t STRIDE = COLS ? SEQ : 1;  // linear distance of neighbouring indices
    static constexpr LdsMap M = lds_map(LOGL, LOGSEQ, COLS);
    static_assert(LOGL + LOGSEQ >= 5 && (M.m1 << M.t1) < 32 && (M.m2 << M.t2) < 32,
                  "the map permutes whole lines of 32 words");
    __host__ __device__ static constexpr uint32_t word(uint32_t x) {
        return x ^ (((x >> M.s1) & M.m1) << M.t1) ^ (((x >> M.s2) & M.m2) << M.t2);
    }
    __device__ static __forceinline__ uint32_t addr(uint32_t seq, uint32_t idx) {
        return word(COLS ? (idx << LOGSEQ) + seq : (seq << LOGL) + idx);
    }
    // Word of position x0 + c, from w0 = word(x0), for a c with no set bit in
    // common with x0: then x0 + c = x0 ^ c and word(x0 + c) = w0 ^ word(c).
    // Above bit 4 that is an addition, so a constant c costs at most one XOR
    // of the low bits; the rest goes into the LDS instruction's offset.
    __device__ static __forceinline__ uint32_t word_at(uint32_t w0, uint32_t c) {
        const uint32_t wc = word(c);
        return (w0 ^ (wc & 31u)) + (wc & ~31u);
    }
};

// Group g of a stage covering local layers [S0, S0 + K): the 2^K elements
// base + t * DD (t < 2^K) of one sequence, DD = L / 2^(S0+K).  `bblk` is the
// block index of the group at local layer S0.
template <class TL, int S0, int K>
struct Stage {
    static constexpr int G = 1 << K;
    static constexpr int NG = kElemsPerThread / G;
    static constexpr int DD = TL::L >> (S0 + K);
    static constexpr int GPS = TL::L / G;  // groups per sequence
    __device__ static __forceinline__ void locate(uint32_t gi, uint32_t& seq, uint32_t& base,
                                                  uint32_t& bblk) {
        uint32_t grp;
        if constexpr (TL::COLS) {
            seq = gi & (TL::SEQ - 1);
            grp = gi >> TL::LOGSEQ;
        } else {
            grp = gi & (GPS - 1);
            seq = gi / GPS;
        }
        uint32_t j = grp % DD;
        bblk = grp / DD;
        base = bblk * (DD * G) + j;
    }
};

// Forward (Cooley-Tukey) layers [S0, S0+K) on NB tiles (1..3) that share
// their twiddles, with the butterflies of prime set SET.  Twiddle of local
// layer S0+l, sub-block `sub` of the group: global layer sb + S0 + l, global
// block ((row0 + seq) << (S0 + l)) + (bblk << l) + sub, where row0 is the
// block index of sequence 0 at layer sb (0 for columns, the row for rows).
// The run's values enter below B0 p (narrow set): layer S0 + l skips the
// reduction of its first input when lazy_ok(B0, S0 + l) (the bound of a layer
// is the same for every element).
template <int SET, class TL, int S0, int K, int NB, int B0, class ZP>
__device__ __forceinline__ void ct_stage(uint32_t* const (&t)[NB], ZP Z, int sb, uint32_t row0, uint32_t p) {
    using ST = Stage<TL, S0, K>;
#pragma unroll
    for (int q = 0; q < ST::NG; ++q) {
        uint32_t seq, base, bblk;
        ST::locate(q * TL::NT + threadIdx.x, seq, base, bblk);
        uint32_t blk = TL::COLS ? row0 : row0 + seq;
        const uint32_t w0 = TL::addr(seq, base);  // the group's elements: base + i DD, bits apart from base's
        uint32_t v[NB][ST::G];
#pragma unroll
        for (int i = 0; i < ST::G; ++i)
#pragma unroll
            for (int u = 0; u < NB; ++u) v[u][i] = t[u][TL::word_at(w0, i * ST::DD * TL::STRIDE)];
#pragma unroll
        for (int l = 0; l < K; ++l) {
            const int h = ST::G >> (l + 1);
#pragma unroll
            for (int sub = 0; sub < (1 << l); ++sub) {
                uint2 z = tw_at(Z, (1u << (sb + S0 + l)) + (blk << (S0 + l)) + (bblk << l) + sub);
#pragma unroll
                for (int i = 0; i < h; ++i)
#pragma unroll
                    for (int u = 0; u < NB; ++u)
                        Arith<SET>::ct_sel(v[u][sub * 2 * h + i], v[u][sub * 2 * h + i + h], z, p, lazy_ok(B0, S0 + l));
            }
        }
#pragma unroll
        for (int i = 0; i < ST::G; ++i)
#pragma unroll
            for (int u = 0; u < NB; ++u) t[u][TL::word_at(w0, i * ST::DD * TL::STRIDE)] = v[u][i];
    }
}

// Inverse (Gentleman-Sande) layers S0+K-1 down to S0 on NB tiles (1..2).
// Narrow set: values below 4p; a butterfly keeps its sum unreduced when both
// inputs are below 2p: in the stage's first layer when FRESH (every value of
// the tile below 2p), in a later one where the layer before (register bit
// log2(h) - 1) left two differences, i & (h >> 1).
template <int SET, class TL, int S0, int K, int NB, bool FRESH, class ZP>
__device__ __forceinline__ void gs_stage(uint32_t* const (&t)[NB], ZP Zi, int sb, uint32_t row0, uint32_t p) {
    using ST = Stage<TL, S0, K>;
#pragma unroll
    for (int q = 0; q < ST::NG; ++q) {
        uint32_t seq, base, bblk;
        ST::locate(q * TL::NT + threadIdx.x, seq, base, bblk);
        uint32_t blk = TL::COLS ? row0 : row0 + seq;
        const uint32_t w0 = TL::addr(seq, base);  // the group's elements: base + i DD, bits apart from base's
        uint32_t v[NB][ST::G];
#pragma unroll
        for (int i = 0; i < ST::G; ++i)
#pragma unroll
            for (int u = 0; u < NB; ++u) v[u][i] = t[u][TL::word_at(w0, i * ST::DD * TL::STRIDE)];
#pragma unroll
        for (int l = K - 1; l >= 0; --l) {
            const int h = ST::G >> (l + 1);
#pragma unroll
            for (int sub = 0; sub < (1 << l); ++sub) {
                uint2 z = tw_at(Zi, (1u << (sb + S0 + l)) + (blk << (S0 + l)) + (bblk << l) + sub);
#pragma unroll
                for (int i = 0; i < h; ++i)
#pragma unroll
                    for (int u = 0; u < NB; ++u)
                        Arith<SET>::gs_sel(v[u][sub * 2 * h + i], v[u][sub * 2 * h + i + h], z, p,
                                           l == K - 1 ? FRESH : (i & (h >> 1)) != 0);
            }
        }
#pragma unroll
        for (int i = 0; i < ST::G; ++i)
#pragma unroll
            for (int u = 0; u < NB; ++u) t[u][TL::word_at(w0, i * ST::DD * TL::STRIDE)] = v[u][i];
    }
}

// Forward layers S0.. of a tile whose values enter the run below B0 p.
template <int SET, class TL, int S0, int NB, int B0 = 8, class ZP>
__device__ __forceinline__ void ct_run(uint32_t* const (&t)[NB], ZP Z, int sb, uint32_t row0, uint32_t p) {
    if constexpr (S0 < TL::LOGL) {
        constexpr int K = (TL::LOGL - S0) < 4 ? (TL::LOGL - S0) : 4;
        ct_stage<SET, TL, S0, K, NB, B0>(t, Z, sb, row0, p);
        __syncthreads();
        ct_run<SET, TL, S0 + K, NB, B0>(t, Z, sb, row0, p);
    }
}

// Inverse layers down to S0; FRESH: the run's values enter below 2p (the
// first stage it runs, the deepest, may start lazy; after an LDS exchange the
// sums and differences of a register are mixed, so later stages reduce first).
template <int SET, class TL, int S0, int NB, bool FRESH = false, class ZP>
__device__ __forceinline__ void gs_run(uint32_t* const (&t)[NB], ZP Zi, int sb, uint32_t row0, uint32_t p) {
    if constexpr (S0 < TL::LOGL) {
        constexpr int K = (TL::LOGL - S0) < 4 ? (TL::LOGL - S0) : 4;
        gs_run<SET, TL, S0 + K, NB, FRESH>(t, Zi, sb, row0, p);
        gs_stage<SET, TL, S0, K, NB, FRESH && S0 + K == TL::LOGL>(t, Zi, sb, row0, p);
        __syncthreads();
    }
}

}  // namespace pintt
