// Prints the LDS bank map of every tile shape given on the command line as
// "logl logseq cols" triples: one line "s1 m1 t1 s2 m2 t2" each.  Built
// host-only from ntt_device.h by tests/test_lds_maps.py.
#include <cstdio>
#include <cstdlib>

#include "ntt_device.h"

int main(int argc, char** argv) {
    for (int i = 1; i + 2 < argc; i += 3) {
        const pintt::LdsMap m = pintt::lds_map(std::atoi(argv[i]), std::atoi(argv[i + 1]), std::atoi(argv[i + 2]) != 0);
        std::printf("%u %u %u %u %u %u\n", m.s1, m.m1, m.t1, m.s2, m.m2, m.t2);
    }
    return 0;
}
