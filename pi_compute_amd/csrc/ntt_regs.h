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
//   after the swaps ("post"):  t = (L5 i2 L4 i0 i3 i1)
// v_permlane32_swap (vdst = v[i], vsrc = v[i | 8], i3 = 0) exchanges lanes
// 32..63 of vdst with lanes 0..31 of vsrc: register bit 3 and lane bit 5 trade
// places.  v_permlane16_swap (vdst = v[i], vsrc = v[i | 2], i1 = 0) exchanges
// the odd 16-lane rows of vdst with the even rows of vsrc: register bit 1 and
// lane bit 4 trade places.  Both are involutions and act on different bits,
// so applying them again restores the entry layout.
//
// Forward layers are Cooley-Tukey (layer l of the stage pairs t bit 5 - l),
// inverse layers Gentleman-Sande in reverse order; the twiddle of stage layer
// l and block `blk` (the top l bits of t) comes from a functor tw(l, blk).
#pragma once

#include "ntt_device.h"

namespace pintt {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

template <int NB>
__device__ __forceinline__ void lane_swaps(uint32_t (&v)[NB][16]) {
#pragma unroll
    for (int u = 0; u < NB; ++u) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const auto r = __builtin_amdgcn_permlane32_swap(v[u][i], v[u][i | 8], false, false);
            v[u][i] = r[0];
            v[u][i | 8] = r[1];
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i & 2) continue;
            const auto r = __builtin_amdgcn_permlane16_swap(v[u][i], v[u][i | 2], false, false);
            v[u][i] = r[0];
            v[u][i | 2] = r[1];
        }
    }
}

// Six forward layers on NB tiles sharing twiddles; entry layout in, post
// layout out.  ZT: the group's top half (t >= 32, registers 8..15) is zero,
// so layer 0 only copies the bottom half up.
template <int SET, int NB, bool ZT = false, class TW>
__device__ __forceinline__ void ct6(uint32_t (&v)[NB][16], const TW& tw, uint32_t p) {
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
                for (int u = 0; u < NB; ++u) Arith<SET>::ct(v[u][sub * 2 * h + x], v[u][sub * 2 * h + x + h], z, p);
        }
    }
    lane_swaps<NB>(v);
    const uint32_t l5 = (lane_id() >> 5) & 1u, l4 = (lane_id() >> 4) & 1u;
    // Layer 4 pairs t1 = i3; block = (L5 i2 L4 i0).
#pragma unroll
    for (int i2 = 0; i2 < 2; ++i2)
#pragma unroll
        for (int i0 = 0; i0 < 2; ++i0) {
            const uint2 z = tw(4, (l5 << 3) | (i2 << 2) | (l4 << 1) | i0);
#pragma unroll
            for (int i1 = 0; i1 < 2; ++i1) {
                const int r = (i2 << 2) | (i1 << 1) | i0;
#pragma unroll
                for (int u = 0; u < NB; ++u) Arith<SET>::ct(v[u][r], v[u][r | 8], z, p);
            }
        }
    // Layer 5 pairs t0 = i1; block = (L5 i2 L4 i0 i3).
#pragma unroll
    for (int i3 = 0; i3 < 2; ++i3)
#pragma unroll
        for (int i2 = 0; i2 < 2; ++i2)
#pragma unroll
            for (int i0 = 0; i0 < 2; ++i0) {
                const uint2 z = tw(5, (l5 << 4) | (i2 << 3) | (l4 << 2) | (i0 << 1) | i3);
                const int r = (i3 << 3) | (i2 << 2) | i0;
#pragma unroll
                for (int u = 0; u < NB; ++u) Arith<SET>::ct(v[u][r], v[u][r | 2], z, p);
            }
}

// The six layers undone: post layout in, entry layout out.
template <int SET, int NB, class TW>
__device__ __forceinline__ void gs6(uint32_t (&v)[NB][16], const TW& tw, uint32_t p) {
    const uint32_t l5 = (lane_id() >> 5) & 1u, l4 = (lane_id() >> 4) & 1u;
#pragma unroll
    for (int i3 = 0; i3 < 2; ++i3)
#pragma unroll
        for (int i2 = 0; i2 < 2; ++i2)
#pragma unroll
            for (int i0 = 0; i0 < 2; ++i0) {
                const uint2 z = tw(5, (l5 << 4) | (i2 << 3) | (l4 << 2) | (i0 << 1) | i3);
                const int r = (i3 << 3) | (i2 << 2) | i0;
#pragma unroll
                for (int u = 0; u < NB; ++u) Arith<SET>::gs(v[u][r], v[u][r | 2], z, p);
            }
#pragma unroll
    for (int i2 = 0; i2 < 2; ++i2)
#pragma unroll
        for (int i0 = 0; i0 < 2; ++i0) {
            const uint2 z = tw(4, (l5 << 3) | (i2 << 2) | (l4 << 1) | i0);
#pragma unroll
            for (int i1 = 0; i1 < 2; ++i1) {
                const int r = (i2 << 2) | (i1 << 1) | i0;
#pragma unroll
                for (int u = 0; u < NB; ++u) Arith<SET>::gs(v[u][r], v[u][r | 8], z, p);
            }
        }
    lane_swaps<NB>(v);
#pragma unroll
    for (int l = 3; l >= 0; --l) {
        const int h = 8 >> l;
#pragma unroll
        for (int sub = 0; sub < (1 << l); ++sub) {
            const uint2 z = tw(l, sub);
#pragma unroll
            for (int x = 0; x < h; ++x)
#pragma unroll
                for (int u = 0; u < NB; ++u) Arith<SET>::gs(v[u][sub * 2 * h + x], v[u][sub * 2 * h + x + h], z, p);
        }
    }
}

// K <= 4 layers inside each lane: register i = (g << K) | t holds element t of
// the lane's group g (16 >> K groups).  tw(g, l, blk).
template <int SET, int K, class TW>
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
                    Arith<SET>::ct(v[(g << K) + sub * 2 * h + x], v[(g << K) + sub * 2 * h + x + h], z, p);
            }
        }
}

template <int SET, int K, class TW>
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
                    Arith<SET>::gs(v[(g << K) + sub * 2 * h + x], v[(g << K) + sub * 2 * h + x + h], z, p);
            }
        }
}

// Register r of the post layout holds element t = post_t(r, lane) of its group.
__device__ __forceinline__ uint32_t post_t(int r, uint32_t lane) {
    const uint32_t i3 = (r >> 3) & 1, i2 = (r >> 2) & 1, i1 = (r >> 1) & 1, i0 = r & 1;
    return (((lane >> 5) & 1u) << 5) | (i2 << 4) | (((lane >> 4) & 1u) << 3) | (i0 << 2) | (i3 << 1) | i1;
}
// ... and of the entry layout.
__device__ __forceinline__ uint32_t entry_t(int r, uint32_t lane) { return ((uint32_t)r << 2) | (lane >> 4); }

}  // namespace pintt
