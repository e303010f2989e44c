// Host check of the engine's stem scan (csrc/cd_scan.h): reads sequences of
// nucleotide codes from stdin (one per line, digits 0-3, optionally followed
// by a space and a seed length) and runs the lane loop of the stem kernel
// (stem_cds in cd_kernels.hip) for PARTS workgroups of NT lanes: lane t of
// part p takes anti-diagonals S = 5 + p*NT + t + j*PARTS*NT for the trip
// count of its wave's first lane, and after every trip the 64 lanes of a
// wave share the largest K, as the kernel's shuffle does.  Prints "l p q" of
// the best stem per sequence.  Planes are allocated at exactly their
// documented size so that the address sanitizer catches any read outside
// them.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cd_scan.h"

int main(int argc, char **argv) {
    const int NT = argc > 1 ? std::stoi(argv[1]) : 256;  // a multiple of 64, or fewer than 64
    const int PARTS = argc > 2 ? std::stoi(argv[2]) : 1;
    const int MIN_HAIRPIN = 4;
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string seq;
        int seed = 0;
        in >> seq >> seed;
        const int n = (int)seq.size();
        std::vector<uint8_t> s(n);
        for (int i = 0; i < n; ++i) s[i] = (uint8_t)(seq[i] - '0');
        const int nw = cd::scan_words(n);
        std::vector<uint2> sp(nw), rp(nw);
        for (auto &w : sp) w = make_uint2(0, 0);
        for (auto &w : rp) w = make_uint2(0, 0);
        for (int i = 0; i < n; ++i) {
            sp[i >> 5].x |= (uint32_t)(s[i] & 1) << (i & 31);
            sp[i >> 5].y |= (uint32_t)(s[i] >> 1) << (i & 31);
            const int rc = s[n - 1 - i] ^ 3;
            rp[i >> 5].x |= (uint32_t)(rc & 1) << (i & 31);
            rp[i >> 5].y |= (uint32_t)(rc >> 1) << (i & 31);
        }
        uint64_t best = 0;
        const int stride = PARTS * NT, last = 2 * n - 3;
        for (int part = 0; part < PARTS; ++part)
            for (int w0 = 0; w0 < NT; w0 += 64) {
                const int lanes = NT - w0 < 64 ? NT - w0 : 64;
                std::vector<int> K(lanes, seed > 1 ? seed : 1);
                std::vector<uint64_t> b(lanes, 0);
                const int Sw = MIN_HAIRPIN + 1 + part * NT + w0;
                const int trips = Sw <= last ? (last - Sw) / stride + 1 : 0;
                for (int j = 0; j < trips; ++j) {
                    int kmax = 0;
                    for (int l = 0; l < lanes; ++l) {
                        const int S = MIN_HAIRPIN + 1 + part * NT + w0 + l + j * stride;
                        if (S <= last) cd::scan_diagonal(sp.data(), rp.data(), n, S, MIN_HAIRPIN, K[l], b[l]);
                        kmax = K[l] > kmax ? K[l] : kmax;
                    }
                    for (int l = 0; l < lanes; ++l) K[l] = kmax;
                }
                for (int l = 0; l < lanes; ++l) best = b[l] > best ? b[l] : best;
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
