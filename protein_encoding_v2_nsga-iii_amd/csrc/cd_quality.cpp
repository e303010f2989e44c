// cd_quality.cpp -- quality indicators of a final pool, on the host.
//
// The reference prints the hypervolume of its normalised pool by running a
// prebuilt `hv` program (src/main.cu:839-858) and the minimum distance to the
// ideal point (MinEuclid, utils.cuh:42-58); its experiment script adds the
// maximum spread (auto_execution_experiments.py:119-134).  This file computes
// them itself: the hypervolume exactly, with the WFG algorithm (While,
// Bradstreet and Barone, IEEE TEVC 16(1), 2012), for minimisation against a
// reference point.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "codon_nsga3.h"

namespace {

using Pt = std::vector<double>;

// q weakly dominates p (minimisation)
bool weakly_dominates(const double *q, const double *p, int d) {
    for (int k = 0; k < d; ++k)
        if (q[k] > p[k]) return false;
    return true;
}

// Keep the points not weakly dominated by another (one copy of duplicates).
std::vector<Pt> nondominated(std::vector<Pt> s, int d) {
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    std::vector<Pt> out;
    for (size_t i = 0; i < s.size(); ++i) {
        bool dom = false;
        for (size_t j = 0; j < s.size() && !dom; ++j)
            if (i != j && weakly_dominates(s[j].data(), s[i].data(), d)) dom = true;
        if (!dom) out.push_back(s[i]);
    }
    return out;
}

double inclhv(const Pt &p, const double *ref, int d) {
    double v = 1.0;
    for (int k = 0; k < d; ++k) v *= (ref[k] - p[k]);
    return v;
}

double hv2(std::vector<Pt> s, const double *ref) {  // exact sweep in two dimensions
    std::sort(s.begin(), s.end(), [](const Pt &a, const Pt &b) { return a[0] < b[0]; });
    double v = 0.0, ylim = ref[1];
    for (const Pt &p : s) {
        if (p[1] < ylim) {
            v += (ref[0] - p[0]) * (ylim - p[1]);
            ylim = p[1];
        }
    }
    return v;
}

double wfg(std::vector<Pt> s, const double *ref, int d);

// Volume dominated by p and by none of `rest` (exclusive hypervolume).
double exclhv(const Pt &p, const std::vector<Pt> &rest, const double *ref, int d) {
    double v = inclhv(p, ref, d);
    if (rest.empty()) return v;
    std::vector<Pt> lim;
    lim.reserve(rest.size());
    for (const Pt &q : rest) {
        Pt m(d);
        for (int k = 0; k < d; ++k) m[k] = std::max(p[k], q[k]);
        lim.push_back(std::move(m));
    }
    lim = nondominated(std::move(lim), d);
    return v - wfg(std::move(lim), ref, d);
}

double wfg(std::vector<Pt> s, const double *ref, int d) {
    if (s.empty()) return 0.0;
    if (s.size() == 1) return inclhv(s[0], ref, d);
    if (d == 2) return hv2(std::move(s), ref);
    // decreasing order of the last objective keeps the limit sets small
    std::sort(s.begin(), s.end(), [d](const Pt &a, const Pt &b) { return a[d - 1] > b[d - 1]; });
    double v = 0.0;
    for (size_t i = 0; i < s.size(); ++i) {
        std::vector<Pt> rest(s.begin() + i + 1, s.end());
        v += exclhv(s[i], rest, ref, d);
    }
    return v;
}

}  // namespace

extern "C" {

/* Exact hypervolume (minimisation) of n points in d dimensions against ref;
 * points outside the reference box contribute nothing. */
int cd_hypervolume(const double *pts, int32_t n, int32_t d, const double *ref, double *out) {
    if (!pts || !ref || !out || n < 0 || d < 1 || d > 16) return CD_E_INVALID;
    std::vector<Pt> s;
    for (int i = 0; i < n; ++i) {
        Pt p(pts + (size_t)i * d, pts + (size_t)(i + 1) * d);
        bool inside = true;
        for (int k = 0; k < d; ++k)
            if (!(p[k] < ref[k])) inside = false;
        if (inside) s.push_back(std::move(p));
    }
    s = nondominated(std::move(s), d);
    *out = wfg(std::move(s), ref, d);
    return CD_OK;
}

/* Normalise raw objective records as src/main.cu:789-835 does before its
 * quality output: objectives 0-3 are stored negated and are flipped back,
 * every objective becomes (x - ideal) / (nadir - ideal) with the reference's
 * fixed ideal {1, 0.2, 0.4, 0.5, 0, 0} and nadir {0, -0.15, 0, 0, 0.6, 1},
 * then values >= 1 become 0.999999 and negative values 0.  out [n][6]. */
int cd_normalize(const float *obj, int32_t n, float *out) {
    if (!obj || !out || n < 0) return CD_E_INVALID;
    static const float ideal[6] = {1.f, 0.2f, 0.4f, 0.5f, 0.f, 0.f};
    static const float nadir[6] = {0.f, -0.15f, 0.f, 0.f, 0.6f, 1.f};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 6; ++k) {
            float v = obj[(size_t)i * 6 + k];
            if (k < 4) v = -v;
            v = (v - ideal[k]) / (nadir[k] - ideal[k]);
            if (v >= 1.f) v = 0.999999f;
            if (v < 0.f) v = 0.f;
            out[(size_t)i * 6 + k] = v;
        }
    return CD_OK;
}

/* Minimum Euclidean distance to the origin (MinEuclid, utils.cuh:42-58) and
 * maximum spread sqrt(sum_k (max_k - min_k)^2)
 * (auto_execution_experiments.py:119-134) of n normalised points. */
int cd_min_distance_spread(const float *pts, int32_t n, float *min_dist, float *spread) {
    if (!pts || n <= 0) return CD_E_INVALID;
    double best = 1e300, sp = 0.0;
    double lo[6], hi[6];
    for (int k = 0; k < 6; ++k) {
        lo[k] = 1e300;
        hi[k] = -1e300;
    }
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) {
            const double v = pts[(size_t)i * 6 + k];
            s += v * v;
            lo[k] = std::min(lo[k], v);
            hi[k] = std::max(hi[k], v);
        }
        best = std::min(best, std::sqrt(s));
    }
    for (int k = 0; k < 6; ++k) sp += (hi[k] - lo[k]) * (hi[k] - lo[k]);
    if (min_dist) *min_dist = (float)best;
    if (spread) *spread = (float)std::sqrt(sp);
    return CD_OK;
}

}  // extern "C"
