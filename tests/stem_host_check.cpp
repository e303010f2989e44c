// Host check of the engine's stem scans (csrc/cd_scan.h, csrc/cd_tile.h,
// csrc/cd_qgram.h): reads sequences of nucleotide codes from stdin (one per
// line, digits 0-3, optionally followed by a seed length and the parent's
// stem length) and runs the lane loops of the stem kernel (stem_cds in
// cd_kernels.hip) for PARTS workgroups of NT lanes.  Mode "diag": lane t of
// part p takes anti-diagonals S = 5 + p*NT + t + j*PARTS*NT.  Mode "tile":
// lane L = p*NT + t takes the flat tiles [L*TT/T, (L+1)*TT/T), and a part
// with no seed of 4 and no run of 4 in its tiles falls back to the "diag"
// lanes of that part with K = 1.  In both, after every trip the 64 lanes of a
// wave share the largest K, as the kernel's shuffle does.  Mode "qgram": a
// changed CDS as the offspring stem kernel takes it (stem_cds on one
// workgroup): q-gram searches with K from the seed, or guessed from the
// parent's length and then 7, and the tiles after a fruitless guess of 7 or
// a search out of budget; an optional fifth argument replaces every lane's
// hit budget, and stderr counts the searches by outcome.  Prints "l p q" of the best stem per
// sequence.  Planes and index arrays are allocated at exactly their
// documented size so that the address sanitizer catches any access outside
// them.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cd_qgram.h"
#include "cd_scan.h"
#include "cd_tile.h"

int main(int argc, char **argv) {
    const int NT = argc > 1 ? std::stoi(argv[1]) : 256;  // a multiple of 64, or fewer than 64
    const int PARTS = argc > 2 ? std::stoi(argv[2]) : 1;
    const std::string MODE = argc > 3 ? argv[3] : "diag";
    const bool TILED = MODE == "tile" || MODE == "qgram", QGRAM = MODE == "qgram";
    const int BUDGET = argc > 4 ? std::stoi(argv[4]) : -1;
    const int MIN_HAIRPIN = 4;
    long n_qgram = 0, n_miss = 0, n_overflow = 0;  // q-gram searches: exact, guess not found, out of budget
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string seq;
        int seed = 0, parent_l = 0;
        in >> seq >> seed >> parent_l;
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
            // the tile share of part `part` of `parts` (its workgroup maximum)
            auto tile_part = [&](int part, int parts) -> uint64_t {
                const int TT = 2 * cd::tile_prefix(n, nblk), T = parts * NT;
                const int trips = TT > 0 ? (TT + T - 1) / T : 0;
                const int stride = parts * NT, last = 2 * n - 3;
                uint64_t bp = 0;
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
                return bp;
            };
            // the q-gram search at K by NT lanes (chunks of 8 rc positions, every
            // NT-th to a lane); returns whether every lookup finished, the
            // maximum key in *qbest
            auto qgram_search = [&](int K, uint64_t *qbest) -> bool {
                const int st = cd::qg_stride(K), NI = cd::qg_index_count(n, st);
                if (NI > cd::qg_index_max(n)) {
                    std::printf("index larger than its LDS share n=%d K=%d\n", n, K);
                    std::exit(1);
                }
                std::vector<uint16_t> head(cd::QG_CODES, (uint16_t)cd::QG_NIL);
                std::vector<uint16_t> next(NI);
                for (int k = 0; k < NI; ++k) cd::qg_insert(sp.data(), st, k, head.data(), next.data());
                const int xe = n - cd::QG_Q + 1, nch = (xe + 7) >> 3;
                bool ok = true;
                uint64_t bq = 0;
                for (int t = 0; t < NT; ++t) {  // lane t's walk; its hits measured as they come
                    int budget = BUDGET >= 0 ? BUDGET : cd::qg_budget(8 * ((nch + NT - 1) / NT));
                    cd::QgLane ln;
                    cd::qg_lane_start(ln, rp.data(), head.data(), t, nch, xe);
                    while (!cd::qg_lane_done(ln, nch) && budget > 0) {
                        int a = 0, x = 0;
                        if (cd::qg_lane_step(ln, rp.data(), head.data(), next.data(), n, st, nch, NT, budget, a, x)) {
                            const uint32_t v = cd::qg_pack(a, x);
                            cd::qg_measure(sp.data(), rp.data(), n, (int)(v >> 16), (int)(v & 0xFFFFu), K, bq);
                        }
                    }
                    ok = ok && cd::qg_lane_done(ln, nch);
                }
                *qbest = bq;
                return ok;
            };
            if (QGRAM) {
                // a changed CDS as the offspring stem kernel takes it: one workgroup,
                // q-gram searches from the seed or the guesses, else its tiles
                bool done = false;
                for (int K = cd::qg_first_k(n, seed, parent_l); K && !done;) {
                    uint64_t qb = 0;
                    if (!qgram_search(K, &qb)) {
                        ++n_overflow;
                        break;
                    }
                    if (seed >= cd::QG_KMIN || (int)(qb >> 40) >= K) {
                        best = qb;
                        done = true;
                        ++n_qgram;
                    } else {
                        ++n_miss;
                        K = cd::qg_next_k(K);
                    }
                }
                if (!done) best = tile_part(0, 1);
            } else {
                for (int part = 0; part < PARTS; ++part) {
                    const uint64_t bp = tile_part(part, PARTS);
                    best = bp > best ? bp : best;
                }
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
    std::fprintf(stderr, "qgram %ld miss %ld overflow %ld\n", n_qgram, n_miss, n_overflow);
    return 0;
}
