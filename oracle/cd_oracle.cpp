/*
 * cd_oracle.cpp -- CPU oracle (test infrastructure) for the NSGA-III
 * codon-design generation loop.
 *
 * TEST INFRASTRUCTURE ONLY: this file restates the reference's algorithm on
 * the CPU so that tests can check the HIP engine against it.  The engine never
 * links, loads or calls it.  Citations are file:line in the reference
 * repository (CAU-HPCL/Protein_encoding_v2_NSGA-III).
 *
 * Deterministic replacements for the reference's nondeterminism (the engine
 * reproduces each of them bit for bit):
 *   R1  cuRAND XORWOW streams -> Philox4x32-10 (Salmon et al., SC'11) keyed by
 *       the job seed, counter = (generation, design/child/ref, slot, tag).
 *   R2  uniform draws: integers U{0..k-1} = (x*k)>>32; Bernoulli(p) is
 *       x < floor(p*2^32), where x is one 32-bit Philox output word.
 *   R3  CAI and CPB are accumulated exactly in 2^-40 fixed point (per-codon
 *       log weights, pair scores), then CAI = exp(mean log w) with a
 *       Taylor-series exp built from IEEE double operations only.
 *   R4  stem (SL) tie rule: max length, then the last CDS, then the smallest
 *       p, then the smallest q (the reference races a spin lock,
 *       utils.cuh:865-884).
 *   R5  fronts are listed in ascending pool index; ASF and association ties
 *       go to the first candidate / lowest reference point.
 *   R6  niching sweeps reference points in ascending index per level and
 *       truncates the final level there (the reference appends with
 *       atomics, sorting.cuh:1185-1279); random picks follow a per-design
 *       Philox priority; the survivor list is in ascending pool index.
 *   R7  defined behaviour where the reference reads out of bounds or hangs:
 *       no left neighbour at codon 0 (mutation.cuh:167,183,312); mutationSL
 *       restores the codon it changed (mutation.cuh:953); a zero
 *       normalisation range maps to 0; a peel pass that ranks nobody ranks
 *       all remaining designs.
 */
#include "cd_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int NOBJ = 6;            // common.cuh:23
constexpr int NCODON = 64;         // common.cuh:17
constexpr int NSENSE = 61;         // TOTAL_CODON_NUM - STOP_CODON_NUM
constexpr int STOP_AA = 20;        // 'Z', info.cuh:8
constexpr int MIN_HAIRPIN = 4;     // common.cuh:34
constexpr float F_PRECISION = 0.000001f;  // common.cuh:53
constexpr int GC_UP = 1, GC_DOWN = -1;    // common.cuh:42-43

// info.cuh:9-10 -- synonym block start and size per amino acid (A..Y, Z=stop)
const int8_t SYN_START[21] = {0, 4, 6, 8, 10, 12, 16, 18, 21, 23, 29, 30, 32, 36, 38, 44, 50, 54, 58, 59, 61};
const int8_t SYN_NUM[21] = {4, 2, 2, 2, 2, 4, 2, 3, 2, 6, 1, 2, 4, 2, 6, 6, 4, 4, 1, 2, 3};

// Philox stream tags (counter word 3)
constexpr uint32_t TAG_INIT = 0x494E4954u;
constexpr uint32_t TAG_TYPE = 0x54595045u;
constexpr uint32_t TAG_MUT = 0x4D555441u;
constexpr uint32_t TAG_SL = 0x534C5452u;
constexpr uint32_t TAG_NICHE = 0x4E494348u;

inline uint32_t uniform_int(uint32_t x, uint32_t k) { return (uint32_t)(((uint64_t)x * k) >> 32); }
inline uint64_t bern_threshold(float p) {
    double t = (double)p * 4294967296.0;
    if (!(t > 0.0)) return 0;
    if (t >= 18446744073709551615.0) return UINT64_MAX;
    return (uint64_t)t;
}
inline bool bern(uint32_t x, uint64_t thr) { return (uint64_t)x < thr; }

void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed, uint32_t out[4]) {
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    or_philox4x32_10(ctr, key, out);
}

inline int comp_codes(int a, int b) { return (a ^ b) == 3; }  // A=0,C=1,G=2,U=3: A-U, C-G

int nt_code(char c) {
    switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'U': return 3;
    default: return -1;
    }
}

// ------------------------------------------------------------------------
// Per-job tables derived from the codon table (info.cuh:11-54)
// ------------------------------------------------------------------------
struct Tables {
    uint8_t nt[NCODON][3];
    float w[NCODON];
    float cps[NSENSE * NSENSE];
    int64_t lw[NCODON];              // R3: llrint(log(w) * 2^40)
    int64_t cpfx[NSENSE * NSENSE];   // R3: llrint(cps * 2^40)
    uint8_t hs[NCODON][NCODON];      // hidden stops across the junction a|b
    uint8_t nd[NCODON][NCODON];      // nucleotide mismatches between codons
    uint8_t gc3[NCODON];
    uint8_t gcnt[NCODON];
    float gc3_ref;

    bool is_stop(int a, int b, int c) const {  // utils.cuh:101-112: the table's last three codons
        for (int i = 1; i <= 3; ++i) {
            const uint8_t *s = nt[NCODON - i];
            if (s[0] == a && s[1] == b && s[2] == c) return true;
        }
        return false;
    }

    bool build(const char *codon_nt, const float *wt, const float *cp, float gref) {
        for (int g = 0; g < NCODON; ++g)
            for (int k = 0; k < 3; ++k) {
                int c = nt_code(codon_nt[3 * g + k]);
                if (c < 0) return false;
                nt[g][k] = (uint8_t)c;
            }
        for (int g = 0; g < NCODON; ++g) {
            w[g] = wt[g];
            if (!(w[g] > 0.0f) || w[g] > 1.0f) return false;
            lw[g] = llrint(std::log((double)w[g]) * 0x1p40);
        }
        for (int i = 0; i < NSENSE * NSENSE; ++i) {
            cps[i] = cp[i];
            if (!(std::fabs(cps[i]) <= 1000.0f)) return false;
            cpfx[i] = llrint((double)cps[i] * 0x1p40);
        }
        for (int a = 0; a < NCODON; ++a) {
            for (int b = 0; b < NCODON; ++b) {
                // countRightSideHSC, utils.cuh:138-161
                int h = 0;
                if (is_stop(nt[a][1], nt[a][2], nt[b][0])) ++h;
                if (is_stop(nt[a][2], nt[b][0], nt[b][1])) ++h;
                hs[a][b] = (uint8_t)h;
                int d = 0;
                for (int k = 0; k < 3; ++k) d += nt[a][k] != nt[b][k];
                nd[a][b] = (uint8_t)d;
            }
            gc3[a] = (nt[a][2] == 1 || nt[a][2] == 2);  // utils.cuh:562 third base G or C
            gcnt[a] = (uint8_t)((nt[a][0] == 1 || nt[a][0] == 2) + (nt[a][1] == 1 || nt[a][1] == 2) +
                                (nt[a][2] == 1 || nt[a][2] == 2));
        }
        gc3_ref = gref;
        return true;
    }
};

// R3: exp(x) for double x, from range reduction x = k ln2 + r and the Taylor
// series of exp(r) to degree 14 in Horner form with fused multiply-adds.
struct ExpCoef {
    double c[15];
    ExpCoef() {
        double f = 1.0;
        for (int j = 0; j < 15; ++j) {
            if (j > 0) f *= (double)j;
            c[j] = 1.0 / f;
        }
    }
};
const ExpCoef EXPC;

double det_exp(double x) {
    const double INV_LN2 = 0x1.71547652b82fep0;
    const double LN2_HI = 0x1.62e42fefa39efp-1;
    const double LN2_LO = 0x1.abc9e3b39803fp-56;
    double k = std::nearbyint(x * INV_LN2);
    double r = std::fma(-k, LN2_HI, x);
    r = std::fma(-k, LN2_LO, r);
    double p = EXPC.c[14];
    for (int j = 13; j >= 0; --j) p = std::fma(p, r, EXPC.c[j]);
    return std::ldexp(p, (int)k);
}

// ------------------------------------------------------------------------
// Hairpin stem, calMaximumSL (utils.cuh:712-888)
// ------------------------------------------------------------------------
struct Stem {
    int32_t l, p, q;
};

// Closed form: l = max over innermost pairs (i, j), j - i > MIN_HAIRPIN, of
// the number of consecutive Watson-Crick pairs (i-t, j+t).  On the
// anti-diagonal i + j = S walked from the outermost pair inwards this is the
// longest run of complementary pairs with j - i >= 5.  Tie rule R4 (smallest
// p = first base of the left arm, then smallest q = innermost base of the
// right arm).
Stem sl_closed(const uint8_t *s, int n) {
    Stem best = {0, 0, 0};
    for (int S = MIN_HAIRPIN + 1; S <= 2 * n - 3; ++S) {
        int i0 = std::max(0, S - (n - 1));
        int i1 = (S - (MIN_HAIRPIN + 1));
        if (i1 < 0) continue;
        i1 >>= 1;
        int run = 0;
        for (int i = i0; i <= i1; ++i) {
            int j = S - i;
            if (comp_codes(s[i], s[j])) {
                ++run;
                if (run >= best.l) {
                    int p = i - run + 1, q = j;
                    if (run > best.l || p < best.p || (p == best.p && q < best.q)) best = {run, p, q};
                }
            } else {
                run = 0;
            }
        }
    }
    return best;
}

// Transliteration of the reference's per-thread anti-diagonal walk
// (utils.cuh:724-836), run sequentially over every t_idx; returns l only
// (the reference resolves (p, q, cds) ties through a race).
int sl_walk(const uint8_t *s, int n) {
    int l = 0;
    auto compat = [&](int a, int b) -> bool {
        if (a < 0 || b < 0 || a >= n || b >= n) return false;
        return comp_codes(s[a], s[b]);
    };
    for (int t = 0; t < 2 * n + 1; ++t) {
        int diag_len, st_r, st_c;
        if (t < n + 1) {
            diag_len = t + 1;
            st_r = n - diag_len;
            st_c = n;
        } else {
            diag_len = 2 * n + 1 - t;
            st_r = -1;
            st_c = diag_len - 1;
        }
        int tmp_l = 0;
        for (int j = 0; j < diag_len; ++j) {
            if (j == 0) {
                tmp_l = 0;
            } else if (compat(st_r + j, st_c - j)) {
                ++tmp_l;
                int r = st_r + j, c = st_c - j;
                if (c > r) {
                    if (c - r > MIN_HAIRPIN && tmp_l >= l) l = tmp_l;
                } else if ((r - tmp_l + 1) > (c + tmp_l - 1)) {
                    if (((r - tmp_l + 1) - (c + tmp_l - 1)) > MIN_HAIRPIN && tmp_l >= l) l = tmp_l;
                } else {
                    tmp_l = 0;
                }
            } else {
                tmp_l = 0;
            }
        }
    }
    return l;
}

int sl_brute(const uint8_t *s, int n) {
    int best = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + MIN_HAIRPIN + 1; j < n; ++j) {
            int k = 0;
            while (i - k >= 0 && j + k < n && comp_codes(s[i - k], s[j + k])) ++k;
            best = std::max(best, k);
        }
    return best;
}

// ------------------------------------------------------------------------
// Canonical evaluation of one design (utils.cuh:271-888 semantics)
// ------------------------------------------------------------------------
struct EvalOut {
    float obj[NOBJ];
    int8_t idx[2 * NOBJ];
    int32_t pql[3];
};

void decode_cds(const Tables &T, const int8_t *aa, int L, const int8_t *syn_cds, int *g_out, uint8_t *nt_out) {
    for (int k = 0; k < L; ++k) {
        int g = SYN_START[(int)aa[k]] + syn_cds[k];
        g_out[k] = g;
        if (nt_out) {
            nt_out[3 * k] = T.nt[g][0];
            nt_out[3 * k + 1] = T.nt[g][1];
            nt_out[3 * k + 2] = T.nt[g][2];
        }
    }
}

void eval_design(const Tables &T, const int8_t *aa, int L, int m, const int8_t *syn, EvalOut &o) {
    const int n = 3 * L;
    std::vector<int> g((size_t)m * L);
    std::vector<uint8_t> nt((size_t)n);
    std::memset(&o, 0, sizeof(o));
    float cai_v = 0, cpb_v = 0, hsc_v = 0, gc_v = 0;
    Stem sl_best = {0, 0, 0};
    int sl_cds = m - 1;
    for (int c = 0; c < m; ++c) {
        decode_cds(T, aa, L, syn + (size_t)c * L, g.data() + (size_t)c * L, nullptr);
    }
    for (int c = 0; c < m; ++c) {
        const int *gc = g.data() + (size_t)c * L;
        // calMinimumCAI, utils.cuh:271-337 (R3 exact accumulation)
        int64_t sfx = 0;
        for (int k = 0; k < L; ++k) sfx += T.lw[gc[k]];
        float cai = (float)det_exp(((double)sfx * 0x1p-40) / (double)L);
        if (c == 0 || -cai >= cai_v) {  // ties -> last CDS (utils.cuh:327)
            cai_v = -cai;
            o.idx[0] = (int8_t)c;
        }
        // calMinimumCBP, utils.cuh:339-409: pairs k, k+1 for k < L-2
        int64_t pfx = 0;
        for (int k = 0; k < L - 2; ++k) pfx += T.cpfx[gc[k] * NSENSE + gc[k + 1]];
        float cpb = (float)(((double)pfx * 0x1p-40) / (double)(L - 2));
        if (c == 0 || -cpb >= cpb_v) {
            cpb_v = -cpb;
            o.idx[2] = (int8_t)c;
        }
        // calMinimumHSC, utils.cuh:411-473: junctions k|k+1 for k < L-1
        int h = 0;
        for (int k = 0; k < L - 1; ++k) h += T.hs[gc[k]][gc[k + 1]];
        float hq = (float)h / (float)(L - 1);
        if (c == 0 || -hq >= hsc_v) {
            hsc_v = -hq;
            o.idx[4] = (int8_t)c;
        }
        // calMaximumGC3, utils.cuh:545-626
        int gcn = 0;
        for (int k = 0; k < L; ++k) gcn += T.gc3[gc[k]];
        float frac = (float)gcn / (float)L;
        float dev = std::fabs(frac - T.gc3_ref);
        if (c == 0 || dev >= gc_v) {
            gc_v = dev;
            o.idx[8] = (int8_t)c;
            o.idx[9] = (int8_t)((frac - T.gc3_ref) < 0 ? GC_UP : GC_DOWN);
        }
        // calMaximumSL, utils.cuh:712-888 (R4)
        for (int k = 0; k < L; ++k) {
            nt[3 * k] = T.nt[gc[k]][0];
            nt[3 * k + 1] = T.nt[gc[k]][1];
            nt[3 * k + 2] = T.nt[gc[k]][2];
        }
        Stem st = sl_closed(nt.data(), n);
        if (st.l >= sl_best.l) {
            sl_best = st;
            sl_cds = c;
        }
    }
    // calMinimumHD, utils.cuh:475-543: pairs i<j in order, ties -> last pair
    float hd_v = 0;
    for (int i = 0; i < m - 1; ++i)
        for (int j = i + 1; j < m; ++j) {
            const int *gi = g.data() + (size_t)i * L;
            const int *gj = g.data() + (size_t)j * L;
            int d = 0;
            for (int k = 0; k < L; ++k) d += T.nd[gi[k]][gj[k]];
            float dq = (float)d / (float)n;
            if ((i == 0 && j == 1) || -dq >= hd_v) {
                hd_v = -dq;
                o.idx[6] = (int8_t)i;
                o.idx[7] = (int8_t)j;
            }
        }
    o.obj[0] = cai_v;
    o.obj[1] = cpb_v;
    o.obj[2] = hsc_v;
    o.obj[3] = hd_v;
    o.obj[4] = gc_v;
    o.obj[5] = (float)sl_best.l / (float)((n - 4) / 2);
    o.idx[10] = (int8_t)sl_cds;
    o.pql[0] = sl_best.p;
    o.pql[1] = sl_best.q;
    o.pql[2] = sl_best.l;
}

// ------------------------------------------------------------------------
// Mutation operators (mutation.cuh) on int8 synonym indices.  `child` is the
// child slot; gen the generation being produced; s the child genome [m][L].
// ------------------------------------------------------------------------
struct MutCtx {
    const Tables *T;
    const int8_t *aa;
    int L, m;
    uint64_t seed;
    uint32_t gen, child;
    uint64_t thr_pm, thr_2pm;
};

inline int syn_of(const MutCtx &c, int k) { return SYN_NUM[(int)c.aa[k]]; }
inline int gidx(const MutCtx &c, const int8_t *s, int cds, int k) {
    return SYN_START[(int)c.aa[k]] + s[(size_t)cds * c.L + k];
}

// Fisher-Yates shuffle of 0..n-1 (indexArrayShuffling, utils.cuh:203-218):
// for i = n-1 down to 1, swap entries i and j_i, with the j_i the
// mixed-radix digits (radix i+1) of uniform_int(x, n!) -- one draw gives a
// uniformly random order of the n <= 6 synonyms.
void shuffle(uint32_t x, int n, int8_t *arr) {
    static const uint32_t FACT[7] = {1, 1, 2, 6, 24, 120, 720};
    for (int i = 0; i < 6; ++i) arr[i] = (int8_t)(i < n ? i : -1);
    uint32_t r = uniform_int(x, FACT[n]);
    for (int i = n - 1; i > 0; --i) {
        const uint32_t j = r % (uint32_t)(i + 1);
        r /= (uint32_t)(i + 1);
        std::swap(arr[i], arr[j]);
    }
}

// R1 draws of the operators below: the i-th slot an operator visits in CDS c
// (every codon, or every other codon for the parity operators) takes words
// 2(i&1) -- the mutation test -- and 2(i&1)+1 -- the choice or the shuffle --
// of Philox(gen, child, c*L + (i >> 1), TAG_MUT).
void slot_draws(const MutCtx &c, int cds, int i, uint32_t &test, uint32_t &pick) {
    uint32_t r[4];
    philox(c.gen, c.child, (uint32_t)(cds * c.L + (i >> 1)), TAG_MUT, c.seed, r);
    test = r[2 * (i & 1)];
    pick = r[2 * (i & 1) + 1];
}

// mutationRandom, mutation.cuh:16-63: every codon of every CDS
void mut_random(const MutCtx &c, int8_t *s) {
    for (int cds = 0; cds < c.m; ++cds)
        for (int k = 0; k < c.L; ++k) {
            uint32_t test, pick;
            slot_draws(c, cds, k, test, pick);
            int ns = syn_of(c, k);
            if (!bern(test, c.thr_pm) || ns == 1) continue;
            int8_t &cur = s[(size_t)cds * c.L + k];
            int nw = (int)uniform_int(pick, (uint32_t)(ns - 1));
            if (nw >= cur) ++nw;
            cur = (int8_t)nw;
        }
}

// mutationCAI, SELECT_UPPER_RANDOM, mutation.cuh:65-142 (main.cu:187)
void mut_cai(const MutCtx &c, int8_t *s, int cds) {
    for (int k = 0; k < c.L; ++k) {
        uint32_t test, pick;
        slot_draws(c, cds, k, test, pick);
        int ns = syn_of(c, k);
        if (!bern(test, c.thr_pm) || ns == 1) continue;
        int8_t &cur = s[(size_t)cds * c.L + k];
        if (cur != ns - 1) cur = (int8_t)(cur + 1 + (int)uniform_int(pick, (uint32_t)(ns - 1 - cur)));
    }
}

// mutationCBP, SELECT_UPPER_RANDOM, mutation.cuh:144-262 (main.cu:190)
void mut_cbp(const MutCtx &c, int8_t *s, int cds) {
    const Tables &T = *c.T;
    const int parity = (int)(c.gen & 1u);
    for (int k = 0; k < c.L - 1; ++k) {
        if ((k % 2) != parity) continue;
        uint32_t test, pick;
        slot_draws(c, cds, k >> 1, test, pick);
        int ns = syn_of(c, k);
        if (!bern(test, c.thr_2pm) || ns == 1) continue;
        const int base = SYN_START[(int)c.aa[k]];
        const int cur = s[(size_t)cds * c.L + k];
        const int left = (k > 0) ? gidx(c, s, cds, k - 1) : -1;  // R7: no left neighbour at k=0
        const int right = gidx(c, s, cds, k + 1);
        auto score = [&](int g) -> float {
            float v = 0.f;
            if (left >= 0) v += T.cps[left * NSENSE + g];
            if (right < NSENSE) v += T.cps[g * NSENSE + right];  // mutation.cuh:187
            return v;
        };
        float curv = score(base + cur);
        int cand[6], nc = 0;
        for (int j = 0; j < ns; ++j)
            if (j != cur && score(base + j) > curv) cand[nc++] = j;
        if (nc > 0) s[(size_t)cds * c.L + k] = (int8_t)cand[uniform_int(pick, (uint32_t)nc)];
    }
}

// mutationHSC, SELECT_UPPER_RANDOM, mutation.cuh:264-388 (main.cu:193)
void mut_hsc(const MutCtx &c, int8_t *s, int cds) {
    const Tables &T = *c.T;
    const int parity = (int)(c.gen & 1u);
    for (int k = 0; k < c.L; ++k) {
        if ((k % 2) != parity) continue;
        uint32_t test, pick;
        slot_draws(c, cds, k >> 1, test, pick);
        int ns = syn_of(c, k);
        if (!bern(test, c.thr_2pm) || ns == 1) continue;
        const int base = SYN_START[(int)c.aa[k]];
        const int cur = s[(size_t)cds * c.L + k];
        const int left = (k > 0) ? gidx(c, s, cds, k - 1) : -1;  // R7
        const int right = (k < c.L - 1) ? gidx(c, s, cds, k + 1) : -1;  // mutation.cuh:300-313
        auto count = [&](int g) -> int {
            int v = 0;
            if (left >= 0) v += T.hs[left][g];
            if (right >= 0) v += T.hs[g][right];
            return v;
        };
        int curv = count(base + cur);
        int cand[6], nc = 0;
        for (int j = 0; j < ns; ++j)
            if (j != cur && count(base + j) > curv) cand[nc++] = j;
        if (nc > 0) s[(size_t)cds * c.L + k] = (int8_t)cand[uniform_int(pick, (uint32_t)nc)];
    }
}

// mutationHD, mutation.cuh:390-534: codons of cds1, scored against all CDS
void mut_hd(const MutCtx &c, int8_t *s, int cds1, int cds2) {
    const Tables &T = *c.T;
    for (int k = 0; k < c.L; ++k) {
        uint32_t test, pick;
        slot_draws(c, cds1, k, test, pick);
        int ns = syn_of(c, k);
        if (!bern(test, c.thr_pm) || ns == 1) continue;
        int8_t order[6];
        shuffle(pick, ns, order);
        const int base = SYN_START[(int)c.aa[k]];
        const int cur = s[(size_t)cds1 * c.L + k];
        const int g2 = gidx(c, s, cds2, k);
        auto min_hd = [&](int g) -> int {
            int mn = 127;
            for (int o = 0; o < c.m; ++o) {
                if (o == cds1) continue;
                int t = T.nd[g][gidx(c, s, o, k)];
                if (t < mn) mn = t;
            }
            return mn;
        };
        int cur_pair = T.nd[base + cur][g2];
        int cur_mhd = min_hd(base + cur);
        int best_pair = -1, best_mhd = -1, best = -1;
        for (int j = 0; j < ns; ++j) {
            int sh = order[j];
            if (sh == cur) continue;
            int np = T.nd[base + sh][g2];
            int nm = min_hd(base + sh);
            if (nm > cur_mhd && nm > best_mhd) {
                best_mhd = nm;
                best = sh;
            } else if (best_mhd == -1 && nm == cur_mhd && np > cur_pair && np > best_pair) {
                best_pair = np;
                best = sh;
            }
        }
        if (best_mhd != -1 || best_pair != -1) s[(size_t)cds1 * c.L + k] = (int8_t)best;
    }
}

// mutationGC3, mutation.cuh:536-653: probability = current GC3 deviation
void mut_gc3(const MutCtx &c, int8_t *s, int cds, int dir, float dev) {
    const Tables &T = *c.T;
    const uint64_t thr = bern_threshold(dev);
    for (int k = 0; k < c.L; ++k) {
        uint32_t test, pick;
        slot_draws(c, cds, k, test, pick);
        int ns = syn_of(c, k);
        if (!bern(test, thr) || ns == 1) continue;
        int8_t order[6];
        shuffle(pick, ns, order);
        const int base = SYN_START[(int)c.aa[k]];
        const int cur = s[(size_t)cds * c.L + k];
        const int curg = T.gc3[base + cur];
        for (int j = 0; j < ns; ++j) {
            int sh = order[j];
            if (sh == cur) continue;
            int ng = T.gc3[base + sh];
            if ((dir == GC_UP && ng > curg) || (dir == GC_DOWN && ng < curg)) {
                s[(size_t)cds * c.L + k] = (int8_t)sh;
                break;
            }
        }
    }
}

// Full-hairpin test: does CDS `cds` hold a stem of length (n-4)/2?  A stem of
// that length spans the whole CDS (outer pair at distance >= n-2).
bool sl_is_full(const uint8_t *nt, int n) {
    const int den = (n - 4) / 2;
    auto run = [&](int a, int b) {  // pairs (a+t, b-t), t < den
        for (int t = 0; t < den; ++t)
            if (!comp_codes(nt[a + t], nt[b - t])) return false;
        return true;
    };
    if (den <= 0) return false;
    if (run(0, n - 1) && (n - 1 - 2 * (den - 1)) > MIN_HAIRPIN) return true;
    if (n % 2 == 1) {
        if (run(0, n - 2) && (n - 2 - 2 * (den - 1)) > MIN_HAIRPIN) return true;
        if (run(1, n - 1) && (n - 2 - 2 * (den - 1)) > MIN_HAIRPIN) return true;
    }
    return false;
}

bool sl_any_stem(const uint8_t *nt, int n) {
    // exists i < j, j - i > MIN_HAIRPIN with complementary bases
    int first[4] = {-1, -1, -1, -1}, last[4] = {-1, -1, -1, -1};
    for (int i = 0; i < n; ++i) {
        if (first[nt[i]] < 0) first[nt[i]] = i;
        last[nt[i]] = i;
    }
    for (int a = 0; a < 4; ++a) {
        int b = a ^ 3;
        if (first[a] >= 0 && last[b] >= 0 && last[b] - first[a] > MIN_HAIRPIN) return true;
    }
    return false;
}

// mutationSL, mutation.cuh:775-969, with the calOneCDS_SL test of
// mutation.cuh:843,896,947.  That test compares the parent's SL objective with
// an integer quotient l_new / ((n-4)/2), which is 0 unless the CDS is one full
// hairpin; it therefore ends the loop after the first trial whenever the
// parent has a stem.  Only the outcome of the test is needed, so the trial is
// followed by the two O(n) checks above instead of a full rescan.
void mut_sl(const MutCtx &c, int8_t *s, const EvalOut &parent) {
    const Tables &T = *c.T;
    const int cds = parent.idx[10];
    const int n = 3 * c.L;
    const int P = parent.pql[0], Lst = parent.pql[2];
    const int l_idx = P / 3;
    const int r_idx = (P + Lst - 1) / 3;  // C truncation toward zero
    const int mid = (r_idx + l_idx) / 2;
    const int stem_len = r_idx - l_idx + 1;
    const int half = stem_len / 2;
    const bool odd = (stem_len % 2) == 1;
    int l_st = odd ? mid - 1 : mid;
    int r_st = mid + 1;
    int32_t cur_L = Lst;
    const float parent_sl = parent.obj[5];
    const int den = (n - 4) / 2;
    std::vector<uint8_t> nt((size_t)n);
    uint32_t trial = 0;

    auto do_trial = [&](int pos) -> bool {  // returns true when the loop terminates
        uint32_t r[4];
        philox(c.gen, c.child, trial++, TAG_SL, c.seed, r);
        int8_t &cur = s[(size_t)cds * c.L + pos];
        const int ns = syn_of(c, pos);
        int8_t saved = cur;
        bool changed = false;
        if (bern(r[0], c.thr_pm) && ns > 1) {
            int nw = (int)uniform_int(r[1], (uint32_t)(ns - 1));
            if (nw == cur) ++nw;  // mutation.cuh:827 (== not >=)
            cur = (int8_t)nw;
            changed = true;
        }
        // calOneCDS_SL: only whether l_new is 0 (no write) or (n-4)/2 matters
        for (int k = 0; k < c.L; ++k) {
            int g = SYN_START[(int)c.aa[k]] + s[(size_t)cds * c.L + k];
            nt[3 * k] = T.nt[g][0];
            nt[3 * k + 1] = T.nt[g][1];
            nt[3 * k + 2] = T.nt[g][2];
        }
        if (sl_any_stem(nt.data(), n)) cur_L = sl_is_full(nt.data(), n) ? den : 0;
        float rhs = (float)(den > 0 ? cur_L / den : 0);
        if (parent_sl > rhs) return true;
        if (changed) cur = saved;  // R7: restore the codon that was changed
        return false;
    };
    if (odd) {
        if (do_trial(mid)) return;
    }
    for (int i = 1; i <= half; ++i) {
        bool t1 = do_trial(l_st);
        l_st -= 1;
        if (t1) return;
        bool t2 = do_trial(r_st);
        r_st += 1;
        if (t2) return;
    }
}

// ------------------------------------------------------------------------
// Selection (sorting.cuh)
// ------------------------------------------------------------------------
bool dominates(const float *a, const float *b) {  // sorting.cuh:94-130
    bool any = false;
    for (int k = 0; k < NOBJ; ++k) {
        if (std::fabs(a[k] - b[k]) > F_PRECISION) {
            any = true;
            if (a[k] > b[k]) return false;
        }
    }
    return any;
}

// nonDominatedSorting, sorting.cuh:974-1089.  Returns fronts (ascending index
// inside each front); stops as soon as >= N designs are ranked.
std::vector<std::vector<int>> nds(const float *obj, int P, int N) {
    std::vector<int> np(P, 0);
    std::vector<std::vector<int>> sp(P);
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < P; ++j) {
            if (i == j) continue;
            if (dominates(&obj[i * NOBJ], &obj[j * NOBJ])) sp[i].push_back(j);
            if (dominates(&obj[j * NOBJ], &obj[i * NOBJ])) np[i] += 1;
        }
    std::vector<std::vector<int>> fronts;
    std::vector<char> ranked(P, 0);
    std::vector<int> f0;
    for (int i = 0; i < P; ++i)
        if (np[i] == 0) f0.push_back(i);
    int count = 0;
    if (f0.empty()) {  // R7 (cannot happen for finite data; kept for symmetry)
        for (int i = 0; i < P; ++i) f0.push_back(i);
    }
    for (int i : f0) ranked[i] = 1;
    count += (int)f0.size();
    fronts.push_back(f0);
    while (count < N) {
        std::vector<int> nxt;
        for (int k : fronts.back())
            for (int j : sp[k]) {
                if (ranked[j]) continue;
                if (--np[j] == 0) nxt.push_back(j);
            }
        if (nxt.empty()) {  // R7: a dominance cycle; rank everything left
            for (int i = 0; i < P; ++i)
                if (!ranked[i]) nxt.push_back(i);
        }
        std::sort(nxt.begin(), nxt.end());
        for (int i : nxt) ranked[i] = 1;
        count += (int)nxt.size();
        fronts.push_back(nxt);
    }
    return fronts;
}

float asf(const float *x, const float *ideal, const float *wv) {  // sorting.cuh:428-449
    float mx = -FLT_MAX;
    for (int k = 0; k < NOBJ; ++k) {
        float t = x[k] - ideal[k];
        if ((double)t < 0.001) t = 0.f;
        float v = t / wv[k];
        if (v > mx) mx = v;
    }
    return mx;
}

// GaussianElimination, sorting.cuh:451-599: solves E x = 1 for the 6x6 matrix
// of (untranslated) extreme points; x returned in sol[6].
void gauss(const float E[NOBJ][NOBJ], float sol[NOBJ]) {
    const int R = NOBJ, C = NOBJ + 1;
    float AB[NOBJ][NOBJ + 1];
    for (int r = 0; r < R; ++r) {
        for (int c = 0; c < NOBJ; ++c) AB[r][c] = E[r][c];
        AB[r][NOBJ] = 1.0f;
    }
    for (int col = 0; col < C - 1; ++col) {
        // replace_zero_gpu, sorting.cuh:451-471
        if (std::fabs(AB[col][col]) <= 1e-6f) {
            int row = col + 1;
            for (; row < R; ++row)
                if (std::fabs(AB[row][col]) > 1e-6f) break;
            if (row < R)
                for (int c = col; c < C; ++c) AB[col][c] += AB[row][c];
        }
        // column_elimination_gpu, sorting.cuh:473-490
        for (int row = col + 1; row < R; ++row) {
            float f = AB[row][col] / AB[col][col];
            for (int c = col; c < C; ++c) AB[row][c] -= f * AB[col][c];
        }
    }
    // reverse_row_elimination + multiple_column, sorting.cuh:492-553
    for (int row = R - 1; row >= 0; --row) {
        float s = 0.f;
        for (int c = row + 1; c < NOBJ; ++c) s += AB[row][c];
        float x = AB[row][NOBJ];
        if (row != R - 1) x -= s;
        x /= AB[row][row];
        AB[row][NOBJ] = x;
        AB[row][row] = 1.0f;
        for (int r = 0; r < row; ++r) AB[r][row] *= x;
    }
    for (int r = 0; r < R; ++r) sol[r] = AB[r][NOBJ];
}

float perp_dist(const float *f, const float *w) {  // utils.cuh:1149-1171
    float num = 0.f, den = 0.f;
    for (int i = 0; i < NOBJ; ++i) {
        num += w[i] * f[i];
        den += w[i] * w[i];
    }
    float k = num / den;
    float d = 0.f;
    for (int i = 0; i < NOBJ; ++i) {
        float t = k * w[i] - f[i];
        d += t * t;
    }
    return std::sqrt(d);
}

struct SelState {
    float ext[NOBJ][NOBJ];
    float ideal[NOBJ], worst[NOBJ], nadir[NOBJ];
    int n_cut, hyp_exc;
};

// sortingKernel, main.cu:409-454
void select_pool(const float *obj, int N, const std::vector<float> &refpts, int R, uint32_t gen,
                 uint64_t seed, SelState &st, std::vector<int> &sorted, std::vector<int> &front_of) {
    const int P = 2 * N;
    auto fronts = nds(obj, P, N);
    front_of.assign(P, -1);
    int ranked = 0;
    for (size_t f = 0; f < fronts.size(); ++f) {
        for (int i : fronts[f]) front_of[i] = (int)f;
        ranked += (int)fronts[f].size();
    }
    st.n_cut = (ranked == N);
    // updateIdealValue / updateWorstValue, sorting.cuh:132-426
    for (int k = 0; k < NOBJ; ++k) {
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (int i = 0; i < P; ++i) {
            mn = std::min(mn, obj[i * NOBJ + k]);
            mx = std::max(mx, obj[i * NOBJ + k]);
        }
        st.ideal[k] = mn;
        st.worst[k] = mx;
    }
    // findExtremePoints, sorting.cuh:601-683
    const std::vector<int> &F0 = fronts[0];
    for (int i = 0; i < NOBJ; ++i) {
        float wv[NOBJ];
        for (int j = 0; j < NOBJ; ++j) wv[j] = (i == j) ? 1.f : F_PRECISION;
        float best = FLT_MAX;
        int bi = -1;
        for (size_t t = 0; t < F0.size(); ++t) {
            float v = asf(&obj[F0[t] * NOBJ], st.ideal, wv);
            if (v < best) {
                best = v;
                bi = (int)t;
            }
        }
        for (int e = 0; e < NOBJ; ++e) {
            float v = asf(st.ext[e], st.ideal, wv);
            if (v < best) {
                best = v;
                bi = (int)F0.size() + e;
            }
        }
        float tmp[NOBJ];
        if (bi < 0) continue;
        if (bi >= (int)F0.size())
            std::memcpy(tmp, st.ext[bi - F0.size()], sizeof(tmp));
        else
            std::memcpy(tmp, &obj[F0[bi] * NOBJ], sizeof(tmp));
        std::memcpy(st.ext[i], tmp, sizeof(tmp));
    }
    // updateNadirValue_HYP, sorting.cuh:911-972
    float x[NOBJ];
    gauss(st.ext, x);
    bool exc = false;
    float intercept[NOBJ];
    for (int t = 0; t < NOBJ; ++t) {
        float res = 0.f;
        for (int i = 0; i < NOBJ; ++i) res += x[i] * (st.ext[t][i] - st.ideal[i]);
        if (std::fabs(res - 1.f) > F_PRECISION || std::isnan(res) || std::isinf(res)) exc = true;
        intercept[t] = 1.f / x[t];
        if (intercept[t] <= F_PRECISION || std::isnan(intercept[t]) || std::isinf(intercept[t])) exc = true;
    }
    st.hyp_exc = exc;
    if (exc) {
        // updateNadirValue_MNDF, sorting.cuh:685-884
        for (int k = 0; k < NOBJ; ++k) {
            float mx = -FLT_MAX;
            for (int i : F0) mx = std::max(mx, obj[i * NOBJ + k]);
            st.nadir[k] = ((mx - st.ideal[k]) > F_PRECISION) ? mx : st.worst[k];
        }
    } else {
        for (int k = 0; k < NOBJ; ++k) st.nadir[k] = st.ideal[k] + intercept[k];
    }

    sorted.clear();
    const int last = (int)fronts.size() - 1;
    for (int f = 0; f < last; ++f)
        for (int i : fronts[f]) sorted.push_back(i);
    if (st.n_cut) {
        for (int i : fronts[last]) sorted.push_back(i);
        std::sort(sorted.begin(), sorted.end());
        return;
    }
    // referenceBasedSorting, sorting.cuh:1091-1282
    float den[NOBJ];
    for (int k = 0; k < NOBJ; ++k) den[k] = st.nadir[k] - st.ideal[k];
    std::vector<int> count(R, 0);
    std::vector<std::vector<int>> bucket(R);
    std::vector<float> dist(P, 0.f);
    for (int f = 0; f <= last; ++f)
        for (int i : fronts[f]) {
            float nf[NOBJ];
            for (int k = 0; k < NOBJ; ++k)
                nf[k] = (den[k] == 0.f) ? 0.f : (obj[i * NOBJ + k] - st.ideal[k]) / den[k];
            float best = FLT_MAX;
            int br = 0;
            for (int r = 0; r < R; ++r) {
                float d = perp_dist(nf, &refpts[(size_t)r * NOBJ]);
                if (d < best) {
                    best = d;
                    br = r;
                }
            }
            dist[i] = best;
            if (f < last)
                count[br] += 1;
            else
                bucket[br].push_back(i);  // ascending pool index
        }
    // Level sweeps, sorting.cuh:1182-1279.  Level 0 takes the candidate
    // nearest to its reference line; later levels take candidates in
    // ascending order of a per-design Philox priority (a uniformly random
    // order, replacing the reference's cuRAND index into its candidate list).
    auto prio = [&](int i) {
        uint32_t rr[4];
        philox(gen, (uint32_t)i, 0u, TAG_NICHE, seed, rr);
        return ((uint64_t)rr[0] << 32) | (uint32_t)i;
    };
    int have = (int)sorted.size();
    for (int level = 0; have < N; ++level) {
        for (int r = 0; r < R && have < N; ++r) {
            if (count[r] != level || bucket[r].empty()) continue;
            std::vector<int> &b = bucket[r];
            int pos = 0;
            for (int t = 1; t < (int)b.size(); ++t) {
                bool better;
                if (level == 0)
                    better = dist[b[t]] < dist[b[pos]] || (dist[b[t]] == dist[b[pos]] && b[t] < b[pos]);
                else
                    better = prio(b[t]) < prio(b[pos]);
                if (better) pos = t;
            }
            sorted.push_back(b[pos]);
            b.erase(b.begin() + pos);
            count[r] += 1;
            ++have;
        }
    }
    std::sort(sorted.begin(), sorted.end());
}

// combinationRepetition / NUnrank / bincoef, sorting.cuh:1285-1354
uint64_t bincoef(int n, int r) {
    if (r < 0 || n < r) return 0;
    if (2 * r > n) r = n - r;
    uint64_t res = 1;
    for (int i = 0; i <= r - 1; ++i) res = (res * (uint64_t)(n - i)) / (uint64_t)(i + 1);
    return res;
}

void nunrank(uint64_t r, float *ref, int m, int p) {
    int x = 1;
    for (int i = 1; i <= p; ++i) {
        uint64_t y = bincoef(m - x - i + 1, p - i);
        while (y <= r) {
            r -= y;
            ++x;
            if (x > m) return;
            y = bincoef(m - x - i + 1, p - i);
        }
        ref[x - 1] += (1.f / (float)p);
    }
}

std::vector<float> dasdennis(int M, int H) {
    uint64_t nc = bincoef(M + H - 1, H);
    std::vector<float> out((size_t)nc * M, 0.f);  // main.cu:665 zeroed buffer
    for (uint64_t i = 0; i < nc; ++i) nunrank(i, &out[(size_t)i * M], M + H - 1, H);
    return out;
}

// ------------------------------------------------------------------------
// Whole job (main.cu:467-787)
// ------------------------------------------------------------------------
struct Job {
    int L, m, N, H, R;
    float Pm;
    uint64_t seed;
    Tables T;
    std::vector<int8_t> aa;
    std::vector<float> refpts;
    std::vector<int8_t> genome[2];
    std::vector<EvalOut> rec[2];
    int cur = 0;
    std::vector<int> sorted, front_of;
    SelState st;
    uint32_t gen = 0;
};

void eval_into(Job &J, int buf, int slot) {
    eval_design(J.T, J.aa.data(), J.L, J.m, &J.genome[buf][(size_t)slot * J.m * J.L], J.rec[buf][slot]);
}

// The operator of child `child` in generation `gen` (main.cu:172-211) applied
// to s, a copy of its parent's genome, given the parent's record.
void mutate_child(const Tables &T, const int8_t *aa, int L, int m, uint64_t seed, float Pm, uint32_t gen,
                  uint32_t child, const EvalOut &pr, int8_t *s) {
    MutCtx mc{&T, aa, L, m, seed, gen, child, bern_threshold(Pm), bern_threshold(Pm * 2.0f)};
    uint32_t r[4];
    philox(gen, child, 0u, TAG_TYPE, seed, r);
    const int type = (int)uniform_int(r[0], (uint32_t)(NOBJ + 1));  // main.cu:172-178
    switch (type) {  // main.cu:181-211
    case 0: mut_random(mc, s); break;
    case 1: mut_cai(mc, s, pr.idx[0]); break;
    case 2: mut_cbp(mc, s, pr.idx[2]); break;
    case 3: mut_hsc(mc, s, pr.idx[4]); break;
    case 4: mut_hd(mc, s, pr.idx[6], pr.idx[7]); break;
    case 5:
        if (pr.idx[9] == GC_UP || pr.idx[9] == GC_DOWN) mut_gc3(mc, s, pr.idx[8], pr.idx[9], pr.obj[4]);
        break;
    case 6: mut_sl(mc, s, pr); break;
    }
}

void offspring(Job &J, int src, int dst, int c) {
    const size_t G = (size_t)J.m * J.L;
    const int parent = J.sorted[c];
    // copySolution x2, main.cu:169-170
    std::memcpy(&J.genome[dst][(size_t)c * G], &J.genome[src][(size_t)parent * G], G);
    J.rec[dst][c] = J.rec[src][parent];
    int8_t *s = &J.genome[dst][(size_t)(J.N + c) * G];
    std::memcpy(s, &J.genome[src][(size_t)parent * G], G);
    mutate_child(J.T, J.aa.data(), J.L, J.m, J.seed, J.Pm, J.gen, (uint32_t)c, J.rec[src][parent], s);
    eval_into(J, dst, J.N + c);
}

int g_threads = 0;

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================
extern "C" {

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

double or_det_exp(double x) { return det_exp(x); }

int or_dasdennis(int M, int H, float *out) {
    auto v = dasdennis(M, H);
    std::memcpy(out, v.data(), v.size() * sizeof(float));
    return (int)(v.size() / M);
}

int or_sl_cds(const uint8_t *codes, int n, int mode, int32_t *l, int32_t *p, int32_t *q) {
    for (int i = 0; i < n; ++i)
        if (codes[i] > 3) return -1;
    if (mode == 0) {
        Stem s = sl_closed(codes, n);
        *l = s.l;
        if (p) *p = s.p;
        if (q) *q = s.q;
    } else if (mode == 1) {
        *l = sl_walk(codes, n);
    } else {
        *l = sl_brute(codes, n);
    }
    return 0;
}

int or_eval(int L, int m, const int8_t *aa, const char *codon_nt, const float *w, const float *cps,
            float gc3_ref, const int8_t *syn, int count, float *obj, int8_t *idx, int32_t *pql) {
    Tables T;
    if (!T.build(codon_nt, w, cps, gc3_ref)) return -1;
    if (L < 3 || m < 2) return -1;
#pragma omp parallel for schedule(dynamic) if (count > 1)
    for (int d = 0; d < count; ++d) {
        EvalOut o;
        eval_design(T, aa, L, m, syn + (size_t)d * m * L, o);
        std::memcpy(obj + (size_t)d * NOBJ, o.obj, sizeof(o.obj));
        std::memcpy(idx + (size_t)d * 2 * NOBJ, o.idx, sizeof(o.idx));
        std::memcpy(pql + (size_t)d * 3, o.pql, sizeof(o.pql));
    }
    return 0;
}

int or_cai_cpb_ref(int L, int m, const int8_t *aa, const char *codon_nt, const float *w, const float *cps,
                   const int8_t *syn, int count, int threads, float *out32, double *out64) {
    Tables T;
    if (!T.build(codon_nt, w, cps, 0.381f)) return -1;
    std::vector<int> g((size_t)L);
    std::vector<float> buf((size_t)threads);
    for (int d = 0; d < count; ++d)
        for (int c = 0; c < m; ++c) {
            decode_cds(T, aa, L, syn + ((size_t)d * m + c) * L, g.data(), nullptr);
            auto tree = [&](bool mul) {  // utils.cuh:297-318 halving tree with odd fold
                int j = threads / 2;
                while (true) {
                    for (int t = 0; t < j; ++t) buf[t] = mul ? buf[t] * buf[t + j] : buf[t] + buf[t + j];
                    if (j == 1) break;
                    if (j % 2 == 1) buf[0] = mul ? buf[0] * buf[j - 1] : buf[0] + buf[j - 1];
                    j /= 2;
                }
                return buf[0];
            };
            for (int t = 0; t < threads; ++t) buf[t] = 1.f;
            for (int k = 0; k < L; ++k)
                buf[k % threads] *= (float)std::pow((double)T.w[g[k]], 1.0 / L);
            float cai32 = tree(true);
            for (int t = 0; t < threads; ++t) buf[t] = 0.f;
            for (int k = 0; k < L - 2; ++k) buf[k % threads] += T.cps[g[k] * NSENSE + g[k + 1]];
            float cpb32 = tree(false) / (float)(L - 2);
            double lsum = 0.0, psum = 0.0;
            for (int k = 0; k < L; ++k) lsum += std::log((double)T.w[g[k]]);
            for (int k = 0; k < L - 2; ++k) psum += (double)T.cps[g[k] * NSENSE + g[k + 1]];
            size_t o = ((size_t)d * m + c) * 2;
            out32[o] = cai32;
            out32[o + 1] = cpb32;
            out64[o] = std::exp(lsum / L);
            out64[o + 1] = psum / (double)(L - 2);
        }
    return 0;
}

int or_mutate(int L, int m, const int8_t *aa, const char *codon_nt, const float *w, const float *cps,
              float gc3_ref, float Pm, uint64_t seed, uint32_t gen, uint32_t child, const int8_t *parent,
              const float *pobj, const int8_t *pidx, const int32_t *ppql, int8_t *out) {
    Tables T;
    if (!T.build(codon_nt, w, cps, gc3_ref)) return -1;
    if (L < 3 || m < 2) return -1;
    EvalOut pr;
    std::memcpy(pr.obj, pobj, sizeof(pr.obj));
    std::memcpy(pr.idx, pidx, sizeof(pr.idx));
    std::memcpy(pr.pql, ppql, sizeof(pr.pql));
    std::memcpy(out, parent, (size_t)m * L);
    mutate_child(T, aa, L, m, seed, Pm, gen, child, pr, out);
    return 0;
}

int or_nds(const float *obj, int P, int N, int32_t *front_of, int32_t *n_fronts) {
    auto fr = nds(obj, P, N);
    for (int i = 0; i < P; ++i) front_of[i] = -1;
    for (size_t f = 0; f < fr.size(); ++f)
        for (int i : fr[f]) front_of[i] = (int32_t)f;
    *n_fronts = (int32_t)fr.size();
    return 0;
}

int or_select(const float *obj, int N, int H, uint32_t gen, uint64_t seed, float *extremes,
              int32_t *sorted_out, int32_t *front_of, float *norm, int32_t *flags) {
    auto ref = dasdennis(NOBJ, H);
    SelState st;
    std::memcpy(st.ext, extremes, sizeof(st.ext));
    std::vector<int> sorted, fo;
    select_pool(obj, N, ref, (int)(ref.size() / NOBJ), gen, seed, st, sorted, fo);
    std::memcpy(extremes, st.ext, sizeof(st.ext));
    for (int i = 0; i < N; ++i) sorted_out[i] = sorted[i];
    for (int i = 0; i < 2 * N; ++i) front_of[i] = fo[i];
    for (int k = 0; k < NOBJ; ++k) {
        norm[k] = st.ideal[k];
        norm[NOBJ + k] = st.worst[k];
        norm[2 * NOBJ + k] = st.nadir[k];
    }
    flags[0] = st.n_cut;
    flags[1] = st.hyp_exc;
    return 0;
}

void *or_job_create(int L, int m, int N, int H, float Pm, uint64_t seed, const int8_t *aa, const char *codon_nt,
                    const float *w, const float *cps, float gc3_ref) {
    // argv validation, main.cu:485-489
    if (N <= 0 || m <= 1 || m > 127 || !(Pm >= 0.f && Pm <= 1.f) || L < 3 || H < 1) return nullptr;
    for (int k = 0; k < L; ++k) {
        if (aa[k] < 0 || aa[k] > STOP_AA) return nullptr;
        if ((aa[k] == STOP_AA) != (k == L - 1)) return nullptr;
    }
    Job *J = new Job();
    if (!J->T.build(codon_nt, w, cps, gc3_ref)) {
        delete J;
        return nullptr;
    }
    J->L = L;
    J->m = m;
    J->N = N;
    J->H = H;
    J->Pm = Pm;
    J->seed = seed;
    J->aa.assign(aa, aa + L);
    J->refpts = dasdennis(NOBJ, H);
    J->R = (int)(J->refpts.size() / NOBJ);
    const size_t G = (size_t)m * L;
    for (int b = 0; b < 2; ++b) {
        J->genome[b].assign((size_t)2 * N * G, 0);
        J->rec[b].assign((size_t)2 * N, EvalOut{});
    }
    std::memset(J->st.ext, 0, sizeof(J->st.ext));
    return J;
}

// initializationKernel, main.cu:26-119
int or_job_init(void *h) {
    Job &J = *(Job *)h;
    const size_t G = (size_t)J.m * J.L;
    J.cur = 0;
    J.gen = 0;
#pragma omp parallel for schedule(dynamic)
    for (int d = 0; d < J.N; ++d) {
        int8_t *s = &J.genome[0][(size_t)d * G];
        for (int c = 0; c < J.m; ++c)
            for (int k = 0; k < J.L; ++k) {
                const int ns = SYN_NUM[(int)J.aa[k]];
                if (d == 0) {  // HIGHEST_CAI_GEN, utils.cuh:1082-1100
                    s[(size_t)c * J.L + k] = (int8_t)(ns - 1);
                } else {  // RANDOM_GEN, utils.cuh:1059-1080
                    uint32_t r[4];
                    philox(0u, (uint32_t)d, (uint32_t)(c * J.L + k), TAG_INIT, J.seed, r);
                    s[(size_t)c * J.L + k] = (int8_t)uniform_int(r[0], (uint32_t)ns);
                }
            }
        eval_into(J, 0, d);
    }
    J.sorted.resize(J.N);
    for (int i = 0; i < J.N; ++i) J.sorted[i] = i;
    J.front_of.assign((size_t)2 * J.N, -1);
    std::memset(J.st.ext, 0, sizeof(J.st.ext));
    for (int k = 0; k < NOBJ && k < J.N; ++k)  // main.cu:106-109
        std::memcpy(J.st.ext[k], J.rec[0][k].obj, sizeof(J.st.ext[k]));
    return 0;
}

// generation loop, main.cu:708-728
int or_job_step(void *h, int gens) {
    Job &J = *(Job *)h;
    for (int g = 0; g < gens; ++g) {
        const int src = J.cur, dst = 1 - J.cur;
#pragma omp parallel for schedule(dynamic)
        for (int c = 0; c < J.N; ++c) offspring(J, src, dst, c);
        std::vector<float> obj((size_t)2 * J.N * NOBJ);
        for (int i = 0; i < 2 * J.N; ++i) std::memcpy(&obj[(size_t)i * NOBJ], J.rec[dst][i].obj, sizeof(float) * NOBJ);
        select_pool(obj.data(), J.N, J.refpts, J.R, J.gen, J.seed, J.st, J.sorted, J.front_of);
        J.cur = dst;
        J.gen += 1;
    }
    return 0;
}

int or_job_read(void *h, int8_t *genome, float *obj, int8_t *idx, int32_t *pql, int32_t *sorted,
                int32_t *front_of, float *state) {
    Job &J = *(Job *)h;
    const size_t G = (size_t)J.m * J.L;
    const int P = 2 * J.N;
    if (genome) std::memcpy(genome, J.genome[J.cur].data(), (size_t)P * G);
    for (int i = 0; i < P; ++i) {
        const EvalOut &o = J.rec[J.cur][i];
        if (obj) std::memcpy(obj + (size_t)i * NOBJ, o.obj, sizeof(o.obj));
        if (idx) std::memcpy(idx + (size_t)i * 2 * NOBJ, o.idx, sizeof(o.idx));
        if (pql) std::memcpy(pql + (size_t)i * 3, o.pql, sizeof(o.pql));
    }
    if (sorted)
        for (int i = 0; i < J.N; ++i) sorted[i] = J.sorted[i];
    if (front_of)
        for (int i = 0; i < P; ++i) front_of[i] = J.front_of.empty() ? -1 : J.front_of[i];
    if (state) {
        std::memcpy(state, J.st.ext, sizeof(J.st.ext));
        std::memcpy(state + 36, J.st.ideal, sizeof(J.st.ideal));
        std::memcpy(state + 42, J.st.worst, sizeof(J.st.worst));
        std::memcpy(state + 48, J.st.nadir, sizeof(J.st.nadir));
    }
    return 0;
}

int or_job_generation(void *h) { return (int)((Job *)h)->gen; }

void or_job_destroy(void *h) { delete (Job *)h; }

int or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    g_threads = n;
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // extern "C"
