// Host check of the engine's bit-sliced stem scan (csrc/cd_stem.h): reads
// sequences of nucleotide codes from stdin (one per line, digits 0-3), runs the
// lane routine for every lane of a workgroup in turn with a shared threshold,
// and prints "l p q" of the best stem per sequence.
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "cd_stem.h"

int main(int argc, char **argv) {
    const int NT = argc > 1 ? std::stoi(argv[1]) : 256;
    const int seed_t = argc > 2 ? std::stoi(argv[2]) : 0;
    std::string line;
    while (std::getline(std::cin, line)) {
        const int n = (int)line.size();
        std::vector<uint8_t> s(n);
        for (int i = 0; i < n; ++i) s[i] = (uint8_t)(line[i] - '0');
        const int nw = cd::stem_words(n);
        std::vector<uint32_t> sl(nw, 0), sh(nw, 0), rl(nw, 0), rh(nw, 0);
        for (int i = 0; i < n; ++i) {
            sl[i >> 5] |= (uint32_t)(s[i] & 1) << (i & 31);
            sh[i >> 5] |= (uint32_t)(s[i] >> 1) << (i & 31);
            const int rc = s[n - 1 - i] ^ 3;
            rl[i >> 5] |= (uint32_t)(rc & 1) << (i & 31);
            rh[i >> 5] |= (uint32_t)(rc >> 1) << (i & 31);
        }
        const int nblk = cd::stem_blocks(n);
        std::vector<int> gpre(2 * nblk + 1, 0);
        for (int b = 0; b < 2 * nblk; ++b) gpre[b + 1] = gpre[b] + cd::stem_groups_of(n, 32 * (b % nblk));
        cd::StemView V{s.data(), sl.data(), sh.data(), rl.data(), rh.data(), n, nblk};
        int shared_t = seed_t;
        uint64_t best = 0;
        const int G = gpre[2 * nblk];
        for (int t = 0; t < NT; ++t) {
            const int g0 = (int)((long)t * G / NT), g1 = (int)((long)(t + 1) * G / NT);
            uint64_t k = cd::stem_lane(
                V, gpre.data(), g0, g1, [&]() { return shared_t; },
                [&](int l) {
                    if (l > shared_t) shared_t = l;
                });
            if (k > best) best = k;
        }
        int l = (int)(best >> 40), p = 0, q = 0;
        if (l > 0) {
            p = 0xFFFFF - (int)((best >> 20) & 0xFFFFF);
            q = 0xFFFFF - (int)(best & 0xFFFFF);
        }
        std::printf("%d %d %d\n", l, p, q);
    }
    return 0;
}
