// Host check of the engine's stem scans (csrc/cd_scan.h, csrc/cd_tile.h):
// reads sequences of nucleotide codes from stdin (one per line, digits 0-3,
// optionally followed by a space and a seed length) and runs the lane loops
// of the stem kernel (stem_cds in cd_kernels.hip) for PARTS workgroups of NT
// lanes.  Mode "diag": lane t of part p takes anti-diagonals
// S = 5 + p*NT + t + j*PARTS*NT.  Mode "tile": lane L = p*NT + t takes the
// flat tiles [L*TT/T, (L+1)*TT/T), and a part with no seed of 4 and no run of
// 4 in its tiles falls back to the "diag" lanes of that part with K = 1.  In
// both, after every trip the 64 lanes of a wave share the largest K, as the
// kernel's shuffle does.  Prints "l p q" of the best stem per sequence.  Planes are allocated at exactly their
// documented size so that the address sanitizer catches any read outside
// them.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cd_scan.h"
#include "cd_tile.h"

int main(int argc, char **argv) {
    const int NT = argc > 1 ? std::stoi(argv[1]) : 256;  // a multiple of 64, or fewer than 64
    const int PARTS = argc > 2 ? std::stoi(argv[2]) : 1;
    const bool TILED = argc > 3 && std::string(argv[3]) == "tile";
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
        if (TILED) {
            // the tiled lane loop of the stem kernel: lane L of PARTS*NT takes flat
            // tiles [L*TT/T, (L+1)*TT/T), trips = ceil(TT/T) for every lane, and the
            // 64 lanes of a wave share the largest K after each trip
            const int nblk = cd::tile_blocks(n);
            int pre = 0;  // the closed-form prefix must equal the explicit sum
            for (int b = 0; b <= nblk; ++b) {
                if (cd::tile_prefix(n, b) != pre) {
                    std::printf("prefix mismatch n=%d b=%d\n", n, b);
                    return 1;
                }
                if (b < nblk) pre += cd::tile_groups(n, 32 * b);
            }
            const int TT = 2 * cd::tile_prefix(n, nblk), T = PARTS * NT;
            const int trips = TT > 0 ? (TT + T - 1) / T : 0;
            const int stride = PARTS * NT, last = 2 * n - 3;
            for (int part = 0; part < PARTS; ++part) {
                uint64_t bp = 0;  // the part's workgroup maximum
                for (int w0 = 0; w0 < NT; w0 += 64) {
                    const int lanes = NT - w0 < 64 ? NT - w0 : 64;
                    std::vector<int> K(lanes, seed > cd::TILE_MIN_RUN ? seed : cd::TILE_MIN_RUN);
                    std::vector<uint64_t> b(lanes, 0);
                    std::vector<cd::TileCursor> cur(lanes);
                    for (int j = 0; j < trips; ++j) {
                        int kmax = 0;
                        for (int l = 0; l < lanes; ++l) {
                            const int L = part * NT + w0 + l;
                            const int t0 = (int)((int64_t)L * TT / T), t1 = (int)((int64_t)(L + 1) * TT / T);
                            if (t0 + j < t1) {
                                if (j == 0) cd::tile_seek(sp.data(), rp.data(), n, t0, cur[l]);
                                cd::tile_step(sp.data(), rp.data(), n, cur[l], K[l], b[l]);
                            }
                            kmax = K[l] > kmax ? K[l] : kmax;
                        }
                        for (int l = 0; l < lanes; ++l) K[l] = kmax;
                    }
                    for (int l = 0; l < lanes; ++l) bp = b[l] > bp ? b[l] : bp;
                }
                if (seed < cd::TILE_MIN_RUN && (int)(bp >> 40) < cd::TILE_MIN_RUN) {
                    // no seed of 4 and no run of 4 in the share: the part's exact scan
                    bp = 0;
                    for (int w0 = 0; w0 < NT; w0 += 64) {
                        const int lanes = NT - w0 < 64 ? NT - w0 : 64;
                        std::vector<int> K(lanes, 1);
                        std::vector<uint64_t> b(lanes, 0);
                        const int Sw = MIN_HAIRPIN + 1 + part * NT + w0;
                        const int ftrips = Sw <= last ? (last - Sw) / stride + 1 : 0;
                        for (int j = 0; j < ftrips; ++j) {
                            int kmax = 0;
                            for (int l = 0; l < lanes; ++l) {
                                const int S = Sw + l + j * stride;
                                if (S <= last) cd::scan_diagonal(sp.data(), rp.data(), n, S, MIN_HAIRPIN, K[l], b[l]);
                                kmax = K[l] > kmax ? K[l] : kmax;
                            }
                            for (int l = 0; l < lanes; ++l) K[l] = kmax;
                        }
                        for (int l = 0; l < lanes; ++l) bp = b[l] > bp ? b[l] : bp;
                    }
                }
                best = bp > best ? bp : best;
            }
        } else {
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
