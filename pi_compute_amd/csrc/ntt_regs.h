// ntt_regs.h - register-resident NTT stages with cross-lane exchanges (gfx950).
//
// A lane holds 16 residues v[0..15].  A 6-layer stage works on groups of 64
// elements spread over 4 lanes (lanes l, l+16, l+32, l+48 of a wave): the first
// four layers pair registers (i3, i2, i1, i0 in turn), then two gfx950 lane
// swaps exchange register bits with lane bits and the last two layers pair
// registers again.  No LDS traffic and no barrier inside a stage.
//
// Element index t (6 bits) of a group, by register bit i_b and lane bit L_b
// (lane = threadIdx.x & 63):
//   entry layout ("phase 1"):  t = (i3 i2 i1 i0 L5 L4)
//   after the swaps ("post"):  SB = 1: t = (L5 i2 L4 i0 i3 i1)
//                              SB = 2: t = (L5 L4 i1 i0 i3 i2)
// v_permlane32_swap (vdst = v[i], vsrc = v[i | 8], i3 = 0) exchanges lanes
// 32..63 of vdst with lanes 0..31 of vsrc: register bit 3 and lane bit 5 trade
// places.  v_permlane16_swap (vdst = v[i], vsrc = v[i | 2^SB], bit SB of i
// clear) exchanges the odd 16-lane rows of vdst with the even rows of vsrc:
// register bit SB and lane bit 4 trade places.  SB is the caller's choice of
// which element bit lands on lane bit 4, which decides the LDS banks of the
// post layout.  Both swaps are involutions and act on different bits, so
// applying them again restores the entry layout.
//
// A five-layer stage (groups of 32 elements on 2 lanes, l and l+32) swaps
// only register bit 3 with lane bit 5: entry t = (i3 i2 i1 i0 L5), post
// t = (L5 i2 i1 i0 i3).
//
// Forward layers are Cooley-Tukey (layer l of the stage pairs t bit 5 - l),
// inverse layers Gentleman-Sande in reverse order; the twiddle of stage layer
// l and block `blk` (the top l bits of t) comes from a functor tw(l, blk).
#pragma once

#include "ntt_device.h"

namespace pintt {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

template <int NB>
__device__ __forceinline__ void swap32(uint32_t (&v)[NB][16]) {
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const auto r = __builtin_amdgcn_permlane32_swap(v[u][i], v[u][i | 8], false, false);
            v[u][i] = r[0];
            v[u][i | 8] = r[1];
        }
}

template <int NB, int SB = 1>
__device__ __forceinline__ void lane_swaps(uint32_t (&v)[NB][16]) {
    static_assert(SB == 1 || SB == 2, "v_permlane16_swap on register bit 1 or 2");
    swap32<NB>(v);
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i & (1 << SB)) continue;
            const auto r = __builtin_amdgcn_permlane16_swap(v[u][i], v[u][i | (1 << SB)], false, false);
            v[u][i] = r[0];
            v[u][i | (1 << SB)] = r[1];
        }
}

// Register r holds, in the post layout, element post_t<SB>(r, lane) of its group.
template <int SB = 1>
__device__ __forceinline__ uint32_t post_t(int r, uint32_t lane) {
    const uint32_t i3 = (r >> 3) & 1, i2 = (r >> 2) & 1, i1 = (r >> 1) & 1, i0 = r & 1;
    const uint32_t l5 = (lane >> 5) & 1u, l4 = (lane >> 4) & 1u;
    if constexpr (SB == 1) return (l5 << 5) | (i2 << 4) | (l4 << 3) | (i0 << 2) | (i3 << 1) | i1;
    else return (l5 << 5) | (l4 << 4) | (i1 << 3) | (i0 << 2) | (i3 << 1) | i2;
}
// ... and of the entry layout.
__device__ __forceinline__ uint32_t entry_t(int r, uint32_t lane) { return ((uint32_t)r << 2) | (lane >> 4); }
// Block index (top l bits of t) of post-layout register r at stage layers 4
// and 5, for SB.
template <int SB>
__device__ __forceinline__ uint32_t post_block(int l, int r, uint32_t lane) {
    return post_t<SB>(r, lane) >> (6 - l);
}

// Six forward layers on NB tiles sharing twiddles; entry layout in, post
// layout out.  ZT: the group's top half (t >= 32, registers 8..15) is zero,
// so layer 0 only copies the bottom half up.
// Forward values enter below B0 p (narrow set): layer l of the stage skips the
// reduction of its first input when lazy_ok says that input is below 6p.  A
// stage exits below ct6_exit(B0, ZT) p.
__host__ __device__ constexpr int ct6_exit(int B0, bool ZT) { return lazy_bound(B0, ZT ? 5 : 6); }

template <int SET, int NB, bool ZT = false, int SB = 1, int B0 = 8, class TW>
__device__ __forceinline__ void ct6(uint32_t (&v)[NB][16], const TW& tw, uint32_t p) {
    // Bound before stage layer l: the zero-top copy of layer 0 changes none.
    auto lazy = [](int l) { return ZT ? lazy_ok(B0, l - 1) : lazy_ok(B0, l); };
    if constexpr (ZT) {
#pragma unroll
        for (int u = 0; u < NB; ++u)
#pragma unroll
            for (int i = 0; i < 8; ++i) v[u][i + 8] = v[u][i];
    }
#pragma unroll
    for (int l = ZT ? 1 : 0; l < 4; ++l) {
        const int h = 8 >> l;
#pragma unroll
        for (int sub = 0; sub < (1 << l); ++sub) {
            const uint2 z = tw(l, sub);
#pragma unroll
            for (int x = 0; x < h; ++x)
#pragma unroll
                for (int u = 0; u < NB; ++u)
                    Arith<SET>::ct_sel(v[u][sub * 2 * h + x], v[u][sub * 2 * h + x + h], z, p, lazy(l));
        }
    }
    lane_swaps<NB, SB>(v);
    const uint32_t lane = lane_id();
    // Layer 4 pairs t1 = register bit 3; layer 5 pairs t0 = register bit SB.
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & 8) continue;
        if (r & (1 << SB)) continue;  // bit SB (t0) is not in layer 4's block: both pairs share its twiddle
        const uint2 z = tw(4, post_block<SB>(4, r, lane));
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int rr = r | (q << SB);
#pragma unroll
            for (int u = 0; u < NB; ++u) Arith<SET>::ct_sel(v[u][rr], v[u][rr | 8], z, p, lazy(4));
        }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & (1 << SB)) continue;
        const uint2 z = tw(5, post_block<SB>(5, r, lane));
#pragma unroll
        for (int u = 0; u < NB; ++u) Arith<SET>::ct_sel(v[u][r], v[u][r | (1 << SB)], z, p, lazy(5));
    }
}

// Inverse values (narrow set) stay below 4p; a butterfly skips the reduction
// of its sum when both inputs are below 2p.  Inputs of the stage are below 2p
// when FRESH (pointwise products, scaled rows), else below 4p.  After a layer
// the differences (the register of the pair with the layer's bit set) are
// below 2p and the sums below 4p, so the next layer of the same registers
// pairs two differences exactly where the previous layer's bit is set in the
// register index; a lane swap mixes sums and differences of one register
// across lanes, so the layer after it reduces everywhere.
//
// The six layers undone: post layout in, entry layout out.
template <int SET, int NB, int SB = 1, bool FRESH = false, class TW>
__device__ __forceinline__ void gs6(uint32_t (&v)[NB][16], const TW& tw, uint32_t p) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & (1 << SB)) continue;
        const uint2 z = tw(5, post_block<SB>(5, r, lane));
#pragma unroll
        for (int u = 0; u < NB; ++u) Arith<SET>::gs_sel(v[u][r], v[u][r | (1 << SB)], z, p, FRESH);
    }
    // Layer 4 pairs registers with the same bit SB (layer 5's): differences
    // of layer 5 where it is set.
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & 8) continue;
        if (r & (1 << SB)) continue;
        const uint2 z = tw(4, post_block<SB>(4, r, lane));
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int rr = r | (q << SB);
#pragma unroll
            for (int u = 0; u < NB; ++u) Arith<SET>::gs_sel(v[u][rr], v[u][rr | 8], z, p, q == 1);
        }
    }
    lane_swaps<NB, SB>(v);
    // Layer l pairs register bit log2(h); the layer before paired bit
    // log2(h) - 1, set in x where x & (h >> 1).
#pragma unroll
    for (int l = 3; l >= 0; --l) {
        const int h = 8 >> l;
#pragma unroll
        for (int sub = 0; sub < (1 << l); ++sub) {
            const uint2 z = tw(l, sub);
#pragma unroll
            for (int x = 0; x < h; ++x)
#pragma unroll
                for (int u = 0; u < NB; ++u)
                    Arith<SET>::gs_sel(v[u][sub * 2 * h + x], v[u][sub * 2 * h + x + h], z, p,
                                       l < 3 && (x & (h >> 1)) != 0);
        }
    }
}

// Five forward layers on groups of 32 elements held by lanes l and l + 32:
// entry t = (i3 i2 i1 i0 L5) in, post t = (L5 i2 i1 i0 i3) out.
template <int SET, int NB, int B0 = 8, class TW>
__device__ __forceinline__ void ct5(uint32_t (&v)[NB][16], const TW& tw, uint32_t p) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const int h = 8 >> l;
#pragma unroll
        for (int sub = 0; sub < (1 << l); ++sub) {
            const uint2 z = tw(l, sub);
#pragma unroll
            for (int x = 0; x < h; ++x)
#pragma unroll
                for (int u = 0; u < NB; ++u)
                    Arith<SET>::ct_sel(v[u][sub * 2 * h + x], v[u][sub * 2 * h + x + h], z, p, lazy_ok(B0, l));
        }
    }
    swap32<NB>(v);
    const uint32_t l5 = (lane_id() >> 5) & 1u;
    // Layer 4 pairs t0 = register bit 3; block = (L5 i2 i1 i0).
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const uint2 z = tw(4, (l5 << 3) | (uint32_t)r);
#pragma unroll
        for (int u = 0; u < NB; ++u) Arith<SET>::ct_sel(v[u][r], v[u][r | 8], z, p, lazy_ok(B0, 4));
    }
}

// Bounds as in gs6.
template <int SET, int NB, bool FRESH = false, class TW>
__device__ __forceinline__ void gs5(uint32_t (&v)[NB][16], const TW& tw, uint32_t p) {
    const uint32_t l5 = (lane_id() >> 5) & 1u;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const uint2 z = tw(4, (l5 << 3) | (uint32_t)r);
#pragma unroll
        for (int u = 0; u < NB; ++u) Arith<SET>::gs_sel(v[u][r], v[u][r | 8], z, p, FRESH);
    }
    swap32<NB>(v);
#pragma unroll
    for (int l = 3; l >= 0; --l) {
        const int h = 8 >> l;
#pragma unroll
        for (int sub = 0; sub < (1 << l); ++sub) {
            const uint2 z = tw(l, sub);
#pragma unroll
            for (int x = 0; x < h; ++x)
#pragma unroll
                for (int u = 0; u < NB; ++u)
                    Arith<SET>::gs_sel(v[u][sub * 2 * h + x], v[u][sub * 2 * h + x + h], z, p,
                                       l < 3 && (x & (h >> 1)) != 0);
        }
    }
}

// K <= 4 layers inside each lane: register i = (g << K) | t holds element t of
// the lane's group g (16 >> K groups).  tw(g, l, blk).
template <int SET, int K, int B0 = 8, class TW>
__device__ __forceinline__ void ct_lane(uint32_t (&v)[16], const TW& tw, uint32_t p) {
    static_assert(K >= 1 && K <= 4, "in-lane stage of 1..4 layers");
#pragma unroll
    for (int g = 0; g < (16 >> K); ++g)
#pragma unroll
        for (int l = 0; l < K; ++l) {
            const int h = (1 << K) >> (l + 1);
#pragma unroll
            for (int sub = 0; sub < (1 << l); ++sub) {
                const uint2 z = tw(g, l, sub);
#pragma unroll
                for (int x = 0; x < h; ++x)
                    Arith<SET>::ct_sel(v[(g << K) + sub * 2 * h + x], v[(g << K) + sub * 2 * h + x + h], z, p,
                                       lazy_ok(B0, l));
            }
        }
}

// Bounds as in gs6: the first layer is lazy when FRESH, every later one where
// the previous layer's bit is set.
template <int SET, int K, bool FRESH = false, class TW>
__device__ __forceinline__ void gs_lane(uint32_t (&v)[16], const TW& tw, uint32_t p) {
    static_assert(K >= 1 && K <= 4, "in-lane stage of 1..4 layers");
#pragma unroll
    for (int g = 0; g < (16 >> K); ++g)
#pragma unroll
        for (int l = K - 1; l >= 0; --l) {
            const int h = (1 << K) >> (l + 1);
#pragma unroll
            for (int sub = 0; sub < (1 << l); ++sub) {
                const uint2 z = tw(g, l, sub);
#pragma unroll
                for (int x = 0; x < h; ++x)
                    Arith<SET>::gs_sel(v[(g << K) + sub * 2 * h + x], v[(g << K) + sub * 2 * h + x + h], z, p,
                                       l == K - 1 ? FRESH : (x & (h >> 1)) != 0);
            }
        }
}


}  // namespace pintt
