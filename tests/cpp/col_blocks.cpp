// Prints, for each "nx nz" pair given on the command line, the (tile, prime,
// job) that every linear workgroup L of a forward column grid (nx, 3, nz) runs
// under prime_grouped_block: one line per L, "tile prime job", and an empty
// line after each grid.  Built host-only from ntt_device.h by
// tests/test_col_blocks.py.
#include <cstdio>
#include <cstdlib>

#include "ntt_device.h"

int main(int argc, char** argv) {
    for (int i = 1; i + 1 < argc; i += 2) {
        const uint32_t nx = (uint32_t)std::atoi(argv[i]), nz = (uint32_t)std::atoi(argv[i + 1]);
        for (uint32_t L = 0; L < 3 * nx * nz; ++L) {
            const pintt::ColBlock b = pintt::prime_grouped_block(L, nx, nz);
            std::printf("%u %u %u\n", b.tile, b.prime, b.job);
        }
        std::printf("\n");
    }
    return 0;
}
