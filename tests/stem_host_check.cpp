// Host check of the engine's stem scan (csrc/cd_scan.h): reads sequences of
// nucleotide codes from stdin (one per line, digits 0-3, optionally followed
// by a space and a seed length), runs the scan for every lane of a workgroup
// of NT lanes in turn (lane t takes anti-diagonals S = 5 + t, 5 + t + NT, ...,
// as the kernel does) and prints "l p q" of the best stem per sequence.
// Planes are allocated at exactly their documented size so that the address
// sanitizer catches any read outside them.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cd_scan.h"

int main(int argc, char **argv) {
    const int NT = argc > 1 ? std::stoi(argv[1]) : 256;
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
        for (int t = 0; t < NT; ++t) {
            int K = seed > 1 ? seed : 1;
            uint64_t b = 0;
            for (int S = MIN_HAIRPIN + 1 + t; S <= 2 * n - 3; S += NT)
                cd::scan_diagonal(sp.data(), rp.data(), n, S, MIN_HAIRPIN, K, b);
            if (b > best) best = b;
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
