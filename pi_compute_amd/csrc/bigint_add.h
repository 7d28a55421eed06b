// bigint_add.h - descriptors and launcher of the batched device add/subtract.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pintt {

constexpr int kAddThreads = 256;
constexpr int kAddLimbsPerThread = 16;
constexpr int kAddLimbsPerBlock = kAddThreads * kAddLimbsPerThread;

struct AddJob {
    const uint64_t* x;
    const uint64_t* y;
    uint64_t* out;
    uint64_t xn, yn;  // limbs (>= 1)
    uint32_t sub;     // 1: out = x - y over max(xn, yn) limbs (x >= y); 0: out = x + y over max(xn, yn) + 1 limbs
    uint32_t blk0;    // first block of the job in the launch
    uint32_t nblk;    // blocks of the job: ceil(max(xn, yn) / kAddLimbsPerBlock)
    uint32_t pad;
};

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s);

// jobs: device array sorted by blk0; maps, carry_in: nblocks bytes each;
// error: set to nonzero when a subtraction's x < y; status[i] (njobs words,
// all written): 1 when job i is such a subtraction, else 0.
hipError_t launch_add_batch(const AddJob* jobs, uint32_t njobs, uint32_t nblocks, uint8_t* maps, uint8_t* carry_in,
                            uint32_t* error, uint32_t* status, hipStream_t s);

// Decimal leaves (decimal_leaves.hip): job i writes exactly width ASCII
// digits of x (xn limbs, x < 10^width, width <= kLeafMaxDigits) at
// out + offset, left-padded with '0'.
constexpr int kLeafThreads = 64;
constexpr int kLeafMaxDigits = 4096;
constexpr int kLeafMaxLimbs = 214;  // ceil(4096 log2(10) / 64) + 1

struct LeafJob {
    const uint64_t* x;
    uint64_t xn, width, offset;
};

hipError_t launch_decimal_leaves(const LeafJob* jobs, uint32_t njobs, uint8_t* out, hipStream_t s);

// Binary-splitting subtrees (bs_leaves.hip): subtree i covers terms
// [a + 64 i, a + 64 (i + 1)) with terms at or past b as the identity; out
// holds its P, Q, R magnitudes (kBsLeafLimbs limbs each, only the first
// lengths[] of them meaningful), lengths
// their exact limb counts, signs those of P and R.  b < 2^24.
constexpr int kBsLeafThreads = 256;
constexpr int kBsLeafTerms = 64;
constexpr int kBsLeafLevels = 6;
constexpr int kBsLeafLimbs = 255;  // (4 << 6) - 1
hipError_t launch_bs_leaves(uint64_t a, uint64_t b, uint32_t count, uint64_t* out, uint32_t* lengths, int32_t* signs,
                            uint32_t* error, hipStream_t s);

}  // namespace pintt
