// cd_api.hip -- C ABI of the codon-design engine (include/codon_nsga3.h):
// job lifecycle, table preparation, work lists and the per-generation
// hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>
#include <mutex>
#include <string>
#include <vector>

#include "cd_host.h"

using cdhost::fail;
using cdhost::Launch;

namespace {

const int8_t SYN_START[21] = {0, 4, 6, 8, 10, 12, 16, 18, 21, 23, 29, 30, 32, 36, 38, 44, 50, 54, 58, 59, 61};
const int8_t SYN_NUM[21] = {4, 2, 2, 2, 2, 4, 2, 3, 2, 6, 1, 2, 4, 2, 6, 6, 4, 4, 1, 2, 3};
const char AMINO[] = "ACDEFGHIKLMNPQRSTVWYZ";
// Standard genetic code, codons enumerated U C A G per position ('Z' = stop).
const char CODE[] = "FFLLSSSSYYZZCCZWLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";

int nt_code(char c) {
    switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'U': return 3;
    default: return -1;
    }
}

char translate(const uint8_t *c3) {  // codes A0 C1 G2 U3 -> index in U C A G order
    static const int ucag[4] = {2, 1, 3, 0};
    return CODE[ucag[c3[0]] * 16 + ucag[c3[1]] * 4 + ucag[c3[2]]];
}

// Tables of a job indexed by codon code (first base << 4 | second << 2 |
// third, A=0 C=1 G=2 U=3), the form k_design works in.
struct HostTables {
    uint8_t code[64];        // code of each codon of the table
    int64_t lw6[64];         // 2^40-scaled log weights
    int64_t cpfx6[64 * 64];  // 2^40-scaled pair scores, 0 where either codon is a stop
    float cps6[64 * 64];     // pair scores, 0 where either codon is a stop
    uint64_t stop6;          // bit per code: the table's last three codons
};

int build_tables(const cd_tables &t, HostTables &h) {
    if (!t.codons || !t.weights || !t.cps) return fail(CD_E_INVALID, "null table pointer");
    uint8_t nt[192];
    for (int i = 0; i < 192; ++i) {
        const int c = nt_code(t.codons[i]);
        if (c < 0) return fail(CD_E_INVALID, "codon table letters must be A, C, G or U");
        nt[i] = (uint8_t)c;
    }
    for (int a = 0; a < 21; ++a)
        for (int g = SYN_START[a]; g < SYN_START[a] + SYN_NUM[a]; ++g)
            if (translate(&nt[3 * g]) != AMINO[a])
                return fail(CD_E_INVALID, "codon table block does not encode its amino acid");
    uint64_t seen = 0;
    for (int g = 0; g < 64; ++g) {
        h.code[g] = (uint8_t)(nt[3 * g] << 4 | nt[3 * g + 1] << 2 | nt[3 * g + 2]);
        seen |= 1ull << h.code[g];
    }
    if (seen != ~0ull) return fail(CD_E_INVALID, "codon table lists a codon twice");
    h.stop6 = 0;
    for (int g = 61; g < 64; ++g) h.stop6 |= 1ull << h.code[g];
    for (int g = 0; g < 64; ++g) {
        const float w = t.weights[g];
        if (!(w > 0.f) || w > 1.f) return fail(CD_E_INVALID, "CAI weights must lie in (0, 1]");
        h.lw6[h.code[g]] = llrint(std::log((double)w) * 0x1p40);
    }
    std::memset(h.cpfx6, 0, sizeof(h.cpfx6));
    std::memset(h.cps6, 0, sizeof(h.cps6));
    for (int a = 0; a < 61; ++a)
        for (int b = 0; b < 61; ++b) {
            const float v = t.cps[a * 61 + b];
            if (!(std::fabs(v) <= 1000.f)) return fail(CD_E_INVALID, "codon-pair scores must be finite and |v| <= 1000");
            h.cps6[h.code[a] * 64 + h.code[b]] = v;
            h.cpfx6[h.code[a] * 64 + h.code[b]] = llrint((double)v * 0x1p40);
        }
    return CD_OK;
}

// Das-Dennis simplex lattice in the order of combinationRepetition
// (sorting.cuh:1285-1354): multisets of H objectives in lexicographic order,
// coordinate = H-fold float sum of 1/H over the objective's multiplicity.
std::vector<float> das_dennis(int H) {
    std::vector<float> out;
    std::vector<int> x(H, 0);
    const float step = 1.f / (float)H;
    while (true) {
        float v[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < H; ++i) v[x[i]] += step;
        out.insert(out.end(), v, v + 6);
        int i = H - 1;
        while (i >= 0 && x[i] == 5) --i;
        if (i < 0) break;
        const int nv = x[i] + 1;
        for (int j = i; j < H; ++j) x[j] = nv;
    }
    return out;
}

int64_t n_ref(int H) {
    int64_t r = 1;
    for (int i = 1; i <= 5; ++i) r = r * (H + i) / i;
    return r;
}

// Bernoulli threshold: a 32-bit draw x succeeds when x < floor(p * 2^32).
uint64_t bern_threshold_host(float p) {
    double t = (double)p * 4294967296.0;
    if (!(t > 0.0)) return 0;
    if (t >= 18446744073709551615.0) return ~0ull;
    return (uint64_t)t;
}

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Sizes {
    size_t total = 0;
    size_t add(size_t b) {
        size_t o = total;
        total += a256(b);
        return o;
    }
};

}  // namespace

namespace {

// Lay out every device buffer of a job in one allocation.
struct Layout {
    size_t pos, code, lw6, cpfx6, cps6, ref, genome[2], obj[2], oidx[2], pql[2], cst[2], hdp[2], planes, stem_max, slist,
        sctl,
        front,
        np, dom,
        ref_of, dist, niche, st, acc, ranked, selm, prio, cutl, cutr, self, off, domi, asc;
    size_t total;
};

Layout layout(int L, int m, int N, int P, int R, int W, int n_dom, int n_asc, bool sel_only) {
    Sizes s;
    Layout o{};
    if (!sel_only) {
        o.pos = s.add(2 * (size_t)cd::row_stride(L));
        o.code = s.add(64);
        o.lw6 = s.add(64 * 8);
        o.cpfx6 = s.add(64 * 64 * 8);
        o.cps6 = s.add(64 * 64 * 4);
        for (int b = 0; b < 2; ++b) {
            o.genome[b] = s.add((size_t)P * m * cd::row_stride(L));
            o.oidx[b] = s.add((size_t)P * 12);
            o.pql[b] = s.add((size_t)P * 16);
            o.cst[b] = s.add((size_t)P * m * sizeof(cd::CdsStat));
            o.hdp[b] = s.add((size_t)P * (m * (m - 1) / 2) * 4);
        }
        o.off = s.add((size_t)N * sizeof(int2));
        o.planes = s.add((size_t)P * m * cd::cds_plane_bytes(L));
        o.stem_max = s.add((size_t)P * 4);
        o.slist = s.add(cd::stem_list_entries(N, m) * 4);
        o.sctl = s.add((size_t)cd::STEM_MAX_GROUPS * 4);
    }
    for (int b = 0; b < (sel_only ? 1 : 2); ++b) o.obj[b] = s.add((size_t)P * 6 * 4);
    if (sel_only) o.obj[1] = o.obj[0];
    o.ref = s.add((size_t)R * 6 * 4);
    o.front = s.add((size_t)P * 4);
    o.np = s.add((size_t)P * 4);
    o.dom = s.add((size_t)P * W * 8);
    o.ref_of = s.add((size_t)P * 4);
    o.dist = s.add((size_t)P * 4);
    o.niche = s.add((size_t)cd::NICHE_INTS * R * 4);
    o.st = s.add(sizeof(cd::SelState));
    o.acc = s.add(sizeof(cd::SelAcc));
    o.ranked = s.add((size_t)W * 8);
    o.selm = s.add((size_t)W * 8);
    o.prio = s.add((size_t)P * 4);
    o.cutl = s.add((size_t)P * 8);
    o.cutr = s.add((size_t)P * 4);
    o.self = s.add(sizeof(cd::DevJob));
    o.domi = s.add((size_t)n_dom * sizeof(int4));
    o.asc = s.add((size_t)n_asc * sizeof(int2));
    o.total = s.total;
    return o;
}

// Room for the work lists of a job of pool P in any shape (the smallest
// group gives the most items; the domination tiles are the same in every
// shape).
void work_counts(int P, int W, int *n_dom, int *n_asc) {
    (void)P;
    *n_dom = W * (W + 1) / 2;
    *n_asc = (P + cd::ASSOC_DESIGNS - 1) / cd::ASSOC_DESIGNS;
}

// Domination tiles and association groups of one job (job index j) in the
// given shapes.  The tiles are the 64 x 64 blocks (row block b, column block
// w) with b <= w; k_dominance takes dom_rows / 64 of them per workgroup.
void job_lists(int P, int W, const cd::Shapes &sh, std::vector<int4> &dom, std::vector<int2> &asc, int j = 0) {
    for (int b = 0; b < W; ++b)
        for (int w = b; w < W; ++w) dom.push_back(make_int4(j, 64 * b, w, 0));
    for (int t = 0; t < (P + sh.assoc_designs - 1) / sh.assoc_designs; ++t)
        asc.push_back(make_int2(j, t * sh.assoc_designs));
}

int validate(const cd_job_desc *d) {
    if (!d) return fail(CD_E_INVALID, "null descriptor");
    // main.cu:485-489
    if (d->N <= 0) return fail(CD_E_INVALID, "population size N must be positive");
    if (d->m < 2 || d->m > cd::MAX_M) return fail(CD_E_INVALID, "CDS count m must be in [2, 16]");
    if (!(d->Pm >= 0.f && d->Pm <= 1.f)) return fail(CD_E_INVALID, "mutation probability must be in [0, 1]");
    if (d->L < 3 || d->L > 100000) return fail(CD_E_INVALID, "L (codons incl. stop) must be in [3, 100000]");
    if (d->H < 1 || n_ref(d->H) > (1 << 20)) return fail(CD_E_INVALID, "H must be >= 1 with at most 2^20 reference points");
    if (2LL * d->N > 64LL * cd::MAX_W) return fail(CD_E_INVALID, "pool 2N exceeds 16384");
    if (!d->aa) return fail(CD_E_INVALID, "null amino-acid sequence");
    for (int k = 0; k < d->L; ++k) {  // findAminoIndex, utils.cuh:60-86
        if (d->aa[k] < 0 || d->aa[k] > 20) return fail(CD_E_INVALID, "amino-acid index out of range");
        if ((d->aa[k] == 20) != (k == d->L - 1)) return fail(CD_E_INVALID, "the stop (20) must be exactly the last codon");
    }
    if (3LL * d->L >= (1 << 20)) return fail(CD_E_INVALID, "sequence too long for stem positions");
    if (cd::design_lds_bytes(d->L, d->m) > cd::LDS_BYTES || cd::stem_list_lds(cd::stem_lds_bytes(d->L)) > cd::LDS_BYTES)
        return fail(CD_E_INVALID, "m*L too large for one workgroup's LDS (160 KiB)");
    return CD_OK;
}

// Pinned host staging for genome rows, which the device holds at a padded
// stride Ls: reads copy the padded block in one transfer and drop the pads on
// the host, writes add them on the host and copy in one transfer (a pitched
// copy to or from pageable memory runs row by row).  Grown on demand and kept
// for the life of the process; callers serialise on its lock.
struct Staging {
    std::mutex mu;
    uint8_t *p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n) {  // with mu held
        if (n <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const size_t c = std::max(n + n / 4, (size_t)1 << 20);
        hipError_t e = hipHostMalloc((void **)&p, c, hipHostMallocDefault);
        if (e == hipSuccess) cap = c;
        return e;
    }
};
Staging g_rows;

// Memory a job frees stays in the device's stream-ordered pool for the next
// job instead of going back to the system at each synchronisation (set once
// per device).
hipError_t keep_pool_memory() {
    static std::mutex mu;
    static bool done[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 64) return e;
    std::lock_guard<std::mutex> lk(mu);
    if (done[dev]) return hipSuccess;
    hipMemPool_t pool;
    e = hipDeviceGetDefaultMemPool(&pool, dev);
    uint64_t keep = ~0ull;
    if (e == hipSuccess) e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    if (e == hipSuccess) done[dev] = true;
    return e;
}

// Regions of a job's allocation to upload at creation; the rest of the
// allocation is zero (cleared on the device).
struct Upload {
    std::vector<char> bytes;
    std::vector<std::pair<size_t, std::pair<size_t, size_t>>> parts;  // (device offset, (host offset, size))
    template <class T>
    void add(size_t dev_off, const T *src, size_t count) {
        const size_t at = bytes.size(), n = count * sizeof(T);
        bytes.resize(at + n);
        std::memcpy(bytes.data() + at, src, n);
        parts.push_back({dev_off, {at, n}});
    }
};

int create_job(const cd_job_desc *d, bool sel_only, int P_override, cd_job **out) {
    cd_job *J = new cd_job();
    J->selection_only = sel_only;
    J->L = sel_only ? 0 : d->L;
    J->m = sel_only ? 0 : d->m;
    J->N = d->N;
    J->H = d->H;
    J->R = (int)n_ref(d->H);
    J->P = P_override > 0 ? P_override : 2 * d->N;
    J->W = (J->P + 63) / 64;
    work_counts(J->P, J->W, &J->n_dom, &J->n_asc);
    Layout o = layout(J->L, J->m, J->N, J->P, J->R, J->W, J->n_dom, J->n_asc, sel_only);
    HostTables t;
    if (!sel_only) {
        const int rc = build_tables(d->tables, t);
        if (rc != CD_OK) {
            delete J;
            return rc;
        }
    }
    // the job's own stream does not wait for (or hold up) work on other
    // streams, such as a served tick that is running while this job is
    // admitted
    // The allocation comes from the device's stream-ordered pool: freeing it
    // (cd_job_destroy) does not wait for every stream of the device, which
    // hipFree does, so jobs can come and go while others run.
    hipError_t e = hipStreamCreateWithFlags(&J->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = keep_pool_memory();
    if (e == hipSuccess) e = hipMallocAsync(&J->mem, o.total, J->stream);
    if (e != hipSuccess) {
        if (J->stream) (void)hipStreamDestroy(J->stream);
        delete J;
        return fail(CD_E_NOMEM, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    }
    J->mem_bytes = o.total;
    char *base = (char *)J->mem;
    Upload up;
    cd::DevJob &h = J->h;
    h.L = J->L;
    h.m = J->m;
    h.N = J->N;
    h.P = J->P;
    h.R = J->R;
    h.W = J->W;
    h.den_sl = sel_only ? 1 : (3 * J->L - 4) / 2;
    h.Ls = sel_only ? 0 : cd::row_stride(J->L);
    h.seed = d->seed;
    h.thr_pm = bern_threshold_host(d->Pm);
    h.thr_2pm = bern_threshold_host(d->Pm * 2.0f);
    h.gc3_ref = d->tables.gc3_ref;
    if (!sel_only) {
        {  // per position the first codon of its amino acid and its synonym count
            const int Ls = cd::row_stride(J->L);
            std::vector<uint8_t> pos(2 * (size_t)Ls, 0);
            for (int k = 0; k < Ls; ++k) {
                const int a = k < J->L ? d->aa[k] : -1;
                pos[k] = a >= 0 ? (uint8_t)SYN_START[a] : 0;
                pos[Ls + k] = a >= 0 ? (uint8_t)SYN_NUM[a] : 1;
            }
            up.add(o.pos, pos.data(), pos.size());
        }
        up.add(o.code, t.code, 64);
        up.add(o.lw6, t.lw6, 64);
        up.add(o.cpfx6, t.cpfx6, 64 * 64);
        up.add(o.cps6, t.cps6, 64 * 64);
        h.pos = (const uint8_t *)(base + o.pos);
        h.code = (const uint8_t *)(base + o.code);
        h.lw6 = (const int64_t *)(base + o.lw6);
        h.cpfx6 = (const int64_t *)(base + o.cpfx6);
        h.cps6 = (const float *)(base + o.cps6);
        h.stop6 = t.stop6;
        for (int b = 0; b < 2; ++b) {
            h.genome[b] = (int8_t *)(base + o.genome[b]);
            h.oidx[b] = (int8_t *)(base + o.oidx[b]);
            h.pql[b] = (int32_t *)(base + o.pql[b]);
            h.cst[b] = (cd::CdsStat *)(base + o.cst[b]);
            h.hdp[b] = (int32_t *)(base + o.hdp[b]);
        }
        std::vector<int2> off(J->N);
        for (int c = 0; c < J->N; ++c) off[c] = make_int2(0, c);
        up.add(o.off, off.data(), off.size());
        J->d_off = (int2 *)(base + o.off);
        J->d_slist = (uint32_t *)(base + o.slist);  // the counts in d_sctl start at zero with the cleared allocation
        J->d_sctl = (int32_t *)(base + o.sctl);
        J->lds = cd::design_lds_bytes(J->L, J->m);
        J->stem_lds = cd::stem_lds_bytes(J->L);
        J->stem_parts = cd::stem_parts_of(J->L);
    }
    const auto ref = das_dennis(d->H);
    up.add(o.ref, ref.data(), ref.size());
    h.refpts = (const float *)(base + o.ref);
    h.obj[0] = (float *)(base + o.obj[0]);
    h.obj[1] = (float *)(base + o.obj[1]);
    h.front_of = (int32_t *)(base + o.front);
    h.np = (int32_t *)(base + o.np);
    h.planes = sel_only ? nullptr : (uint2 *)(base + o.planes);
    h.stem_max = sel_only ? nullptr : (int32_t *)(base + o.stem_max);
    h.dom = (uint64_t *)(base + o.dom);
    h.ref_of = (int32_t *)(base + o.ref_of);
    h.dist = (float *)(base + o.dist);
    h.niche = (int32_t *)(base + o.niche);
    h.st = (cd::SelState *)(base + o.st);
    h.acc = (cd::SelAcc *)(base + o.acc);
    h.ranked = (uint64_t *)(base + o.ranked);
    h.selm = (uint64_t *)(base + o.selm);
    h.prio = (uint32_t *)(base + o.prio);
    h.cutl = (uint64_t *)(base + o.cutl);
    h.cutr = (int32_t *)(base + o.cutr);
    up.add(o.self, &h, 1);
    J->d_self = (cd::DevJob *)(base + o.self);
    if (!sel_only) {  // the tuned shapes of this kind of job, if any
        cd_job *self = J;
        cd::Shapes sh;
        if (cdhost::cached_shapes(cdhost::work_signature(&self, 1), &sh)) J->shapes = sh;
    } else {
        J->shapes.dom_rows = cd::DOM_ROWS;
    }
    {
        std::vector<int4> dom;
        std::vector<int2> asc;
        job_lists(J->P, J->W, J->shapes, dom, asc);
        up.add(o.domi, dom.data(), dom.size());
        up.add(o.asc, asc.data(), asc.size());
        J->n_dom = (int)dom.size();
        J->n_asc = (int)asc.size();
    }
    J->d_dom = (int4 *)(base + o.domi);
    J->d_asc = (int2 *)(base + o.asc);
    e = hipMemsetAsync(J->mem, 0, o.total, J->stream);
    for (const auto &pt : up.parts)
        if (e == hipSuccess)
            e = hipMemcpyAsync(base + pt.first, up.bytes.data() + pt.second.first, pt.second.second,
                               hipMemcpyHostToDevice, J->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(J->stream);  // the staged bytes are released below
    if (e == hipSuccess && J->lds > 48 * 1024) e = cd::set_design_lds_limit(J->lds);
    if (e == hipSuccess && cd::stem_list_lds(J->stem_lds) > 48 * 1024)
        e = cd::set_stem_lds_limit(cd::stem_list_lds(J->stem_lds));
    if (e == hipSuccess && cd::niche_lds_bytes(J->P) > 48 * 1024) e = cd::set_niche_lds_limit(cd::niche_lds_bytes(J->P));
    if (e != hipSuccess) {
        (void)hipFreeAsync(J->mem, J->stream);
        (void)hipStreamSynchronize(J->stream);
        (void)hipStreamDestroy(J->stream);
        delete J;
        return fail(CD_E_DEVICE, std::string("job setup: ") + hipGetErrorString(e));
    }
    *out = J;
    return CD_OK;
}

// Evaluate the designs of `items`: mutation (children) and objectives 0-4,
// the stem of every changed CDS, objective 5.
// Initialisation and evaluation of the designs of `items` (offspring
// generations go through enqueue_generation).
hipError_t enqueue_designs(const Launch &w, int mode, const int2 *items, int n_items, hipStream_t s) {
    hipError_t e = cd::launch_design(mode, w.jobs, items, n_items, w.lds, s);
    if (e == hipSuccess) e = cd::launch_stem(mode, w.jobs, items, n_items, w.mg, w.stem_parts, w.stem_lds, s);
    if (e == hipSuccess) e = cd::launch_stem_finish(mode, w.jobs, items, n_items, s);
    return e;
}

}  // namespace

namespace cdhost {

hipError_t enqueue_generation(const Launch &w, hipStream_t s, hipEvent_t *ev) {
    // ev (optional): CD_PROFILE_KERNELS + 1 events recorded around the six
    // phases -- offspring (both k_design launch classes), stem list,
    // dominance, fronts, association, niching
    auto mark = [&](int k) { return ev ? hipEventRecord(ev[k], s) : hipSuccess; };
    hipError_t e = mark(0);
    const cd::StemList sl = cd::stem_list_geom(w.n_off, w.mg);
    if (e == hipSuccess) {
        const int ns = w.n_small > 0 && w.n_small < w.n_off ? w.n_small : 0;
        if (ns)
            e = cd::launch_design(cd::MODE_OFFSPRING, w.jobs, w.off, ns, w.lds_small, s, w.stem_list, w.stem_ctl, sl, 0);
        if (e == hipSuccess)
            e = cd::launch_design(cd::MODE_OFFSPRING, w.jobs, w.off + ns, w.n_off - ns, w.lds, s, w.stem_list,
                                  w.stem_ctl, sl, ns, w.sh.design_threads);
    }
    if (e == hipSuccess) e = mark(1);
    if (e == hipSuccess)
        e = cd::launch_stem_list(w.jobs, w.off, w.stem_list, w.stem_ctl, sl, w.sh.stem_waves, w.stem_lds, s);
    if (e == hipSuccess) e = mark(2);
    if (e == hipSuccess)
        e = cd::launch_dominance(w.jobs, w.dom, w.n_dom, true, w.stem_ctl, sl.groups, w.sh.dom_rows, s);
    if (e == hipSuccess) e = mark(3);
    if (e == hipSuccess) e = cd::launch_fronts(w.jobs, w.njobs, w.max_pool, w.sh.fronts_threads, s);
    if (e == hipSuccess) e = mark(4);
    if (e == hipSuccess) e = cd::launch_associate(w.jobs, w.asc, w.n_asc, w.sh.assoc_designs, s);
    if (e == hipSuccess) e = mark(5);
    if (e == hipSuccess) e = cd::launch_niche(w.jobs, w.njobs, w.max_pool, w.max_refs, w.sh.niche_threads, s);
    if (e == hipSuccess) e = mark(6);
    return e;
}

void build_lists(cd_job *const *jobs, int count, const cd::Shapes &sh, WorkLists &wl) {
    wl = WorkLists();
    // children of the jobs whose k_design LDS is at most sh.design_split
    // first (a launch of its own, so that more of them share a CU); within
    // each class the jobs with the most codons per child (L * m) first, so
    // that the longest workgroups start early and short ones fill the end of
    // the launch
    std::vector<int> order(count);
    for (int j = 0; j < count; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (int64_t)jobs[a]->L * jobs[a]->m > (int64_t)jobs[b]->L * jobs[b]->m;
    });
    for (int pass = 0; pass < 2; ++pass)
        for (int j : order) {
            cd_job *J = jobs[j];
            if ((J->lds <= (size_t)sh.design_split) != (pass == 0)) continue;
            for (int c = 0; c < J->N; ++c) wl.off.push_back(make_int2(j, c));
            if (pass == 0) {
                wl.n_small += J->N;
                wl.lds_small = std::max(wl.lds_small, J->lds);
            }
        }
    for (int j = 0; j < count; ++j) {
        cd_job *J = jobs[j];
        wl.descs.push_back(J->h);
        wl.n_cds += J->N * J->m;
        job_lists(J->P, J->W, sh, wl.dom, wl.asc, j);
        wl.lds = std::max(wl.lds, J->lds);
        wl.stem_lds = std::max(wl.stem_lds, J->stem_lds);
        wl.max_pool = std::max(wl.max_pool, J->P);
        wl.max_refs = std::max(wl.max_refs, J->R);
        wl.mg = std::max(wl.mg, J->m);
        wl.stem_parts = std::max(wl.stem_parts, J->stem_parts);
    }
}

hipError_t set_lds_limits(const WorkLists &wl) {
    hipError_t e = hipSuccess;
    if (wl.lds > 48 * 1024) e = cd::set_design_lds_limit(wl.lds);
    if (e == hipSuccess && cd::stem_list_lds(wl.stem_lds) > 48 * 1024)
        e = cd::set_stem_lds_limit(cd::stem_list_lds(wl.stem_lds));
    if (e == hipSuccess && cd::niche_lds_bytes(wl.max_pool) > 48 * 1024)
        e = cd::set_niche_lds_limit(cd::niche_lds_bytes(wl.max_pool));
    return e;
}

}  // namespace cdhost

using cdhost::enqueue_generation;

namespace {

int capture(hipStream_t s, hipGraph_t *g, hipGraphExec_t *ge, const Launch &w, int generations) {
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; k < generations && e == hipSuccess; ++k) e = enqueue_generation(w, s, nullptr);
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &graph);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(CD_E_DEVICE, std::string("capture: ") + hipGetErrorString(e));
    }
    if (e2 != hipSuccess) return fail(CD_E_DEVICE, std::string("end capture: ") + hipGetErrorString(e2));
    e = hipGraphInstantiate(ge, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(CD_E_DEVICE, std::string("graph instantiation: ") + hipGetErrorString(e));
    }
    *g = graph;
    return CD_OK;
}

// Enqueue `generations` generations as graph launches: blocks of
// GRAPH_BLOCK, then single generations (each graph captured on first use).
int step_graphs(StepGraphs &G, hipStream_t s, const Launch &w, int generations) {
    int g = generations;
    if (g >= GRAPH_BLOCK && !G.exec[1]) {
        const int rc = capture(s, &G.graph[1], &G.exec[1], w, GRAPH_BLOCK);
        if (rc != CD_OK) return rc;
    }
    for (; g >= GRAPH_BLOCK; g -= GRAPH_BLOCK) HIPCHK(hipGraphLaunch(G.exec[1], s));
    if (g > 0 && !G.exec[0]) {
        const int rc = capture(s, &G.graph[0], &G.exec[0], w, 1);
        if (rc != CD_OK) return rc;
    }
    for (; g > 0; --g) HIPCHK(hipGraphLaunch(G.exec[0], s));
    return CD_OK;
}

Launch job_launch(cd_job *J) {
    return Launch{J->d_self, 1,          J->d_off,     J->N,          J->d_dom,   J->n_dom,   J->d_asc,
                  J->n_asc,  J->lds,     J->stem_lds,  J->P,          J->R,       J->m,          J->stem_parts,
                  J->d_slist,
                  J->d_sctl, J->N * J->m, 0,           0,             J->shapes};
}

// Wait for the job's own stream and for the streams of the batches that hold it.
hipError_t sync_job(cd_job *J) {
    for (cd_batch *B : J->batches) {
        const hipError_t e = hipStreamSynchronize(B->stream);
        if (e != hipSuccess) return e;
    }
    return hipStreamSynchronize(J->stream);
}

// Run `generations` generations of `w` launching every kernel apart, with
// HIP events on stream s around each phase; kernel_ms receives the mean
// duration of each phase (CD_PROFILE_KERNELS values).
int profile_generations(const Launch &w, hipStream_t s, int generations, float *kernel_ms) {
    constexpr int NK = CD_PROFILE_KERNELS;
    hipEvent_t ev[NK + 1];
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    double acc[NK] = {0};
    for (int g = 0; g < generations; ++g) {
        HIPCHK(cdhost::enqueue_generation(w, s, ev));
        HIPCHK(hipEventSynchronize(ev[NK]));
        for (int k = 0; k < NK; ++k) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
            acc[k] += t;
        }
    }
    for (int k = 0; k < NK; ++k) kernel_ms[k] = generations > 0 ? (float)(acc[k] / generations) : 0.f;
    for (auto &e : ev) HIPCHK(hipEventDestroy(e));
    return CD_OK;
}

// Coordinate search over the launch shapes.  Every candidate runs `gens`
// ordinary generations of the work (all shapes give the same results, so
// the job's trajectory does not depend on the search); a candidate's cost is
// the mean time of the phases its parameter changes, from HIP events around
// each phase.  Parameters are settled one after another, each from the best
// shapes so far; `apply` installs shapes and returns the launch description.
template <class Apply>
int tune_shapes(const cd::Shapes &start, int njobs, int gens, hipStream_t s, Apply apply, cd::Shapes *best,
                int *run) {
    struct Param {
        int cd::Shapes::*field;
        std::vector<int> values;
        unsigned phases;  // bit k: phase k of the generation profile
    };
    const std::vector<Param> params = {
        {&cd::Shapes::design_split, {0, 20 * 1024, 24 * 1024, 29 * 1024}, 0x3u},
        {&cd::Shapes::design_threads, {256, 512}, 0x1u},
        {&cd::Shapes::stem_waves, {5, 6, 7, 8}, 0x2u},
        {&cd::Shapes::dom_rows, {64, 128, 256}, 0x4u},
        {&cd::Shapes::fronts_threads, {256, 512, 1024}, 0x8u},
        {&cd::Shapes::assoc_designs, {4, 8, 16, 32, 64}, 0x10u},
        {&cd::Shapes::niche_threads, {256, 512, 1024}, 0x20u},
    };
    cd::Shapes cur = start;
    *run = 0;
    for (const Param &p : params) {
        if (p.field == &cd::Shapes::design_split && njobs == 1) continue;  // one job is one launch class
        double best_cost = 1e30;
        int best_v = cur.*(p.field);
        for (int v : p.values) {
            cd::Shapes cand = cur;
            cand.*(p.field) = v;
            Launch w{};
            int rc = apply(cand, &w);
            if (rc != CD_OK) return rc;
            float ms[CD_PROFILE_KERNELS];
            rc = profile_generations(w, s, gens, ms);
            if (rc != CD_OK) return rc;
            *run += gens;
            double cost = 0.0;
            for (int k = 0; k < CD_PROFILE_KERNELS; ++k)
                if (p.phases >> k & 1u) cost += ms[k];
            if (cost < best_cost) {
                best_cost = cost;
                best_v = v;
            }
        }
        cur.*(p.field) = best_v;
    }
    Launch w{};
    const int rc = apply(cur, &w);
    if (rc != CD_OK) return rc;
    *best = cur;
    return CD_OK;
}

}  // namespace

namespace cdhost {

namespace {
std::mutex g_tune_mu;
std::map<std::string, cd::Shapes> g_tuned;

int log2_bucket(int64_t x) {
    int b = 0;
    while (x > 1) {
        x >>= 1;
        ++b;
    }
    return b;
}
}  // namespace

std::string work_signature(cd_job *const *jobs, int count) {
    int64_t children = 0, lm = 0, pool = 0;
    for (int j = 0; j < count; ++j) {
        children += jobs[j]->N;
        lm = std::max(lm, (int64_t)jobs[j]->L * jobs[j]->m);
        pool = std::max(pool, (int64_t)jobs[j]->P);
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    return "dev" + std::to_string(dev) + "/jobs" + std::to_string(log2_bucket(count)) + "/children" +
           std::to_string(log2_bucket(children)) + "/codons" + std::to_string(log2_bucket(lm)) + "/pool" +
           std::to_string(log2_bucket(pool));
}

bool cached_shapes(const std::string &sig, cd::Shapes *out) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(sig);
    if (it == g_tuned.end()) return false;
    *out = it->second;
    return true;
}

void store_shapes(const std::string &sig, const cd::Shapes &sh) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[sig] = sh;
}

}  // namespace cdhost

extern "C" {

const char *cd_last_error(void) { return cdhost::g_err.c_str(); }

int64_t cd_job_device_bytes(int32_t L, int32_t m, int32_t N, int32_t H) {
    if (L < 3 || m < 2 || N < 1 || H < 1) return -1;
    const int P = 2 * N, W = (P + 63) / 64;
    int n_dom, n_asc;
    work_counts(P, W, &n_dom, &n_asc);
    return (int64_t)layout(L, m, N, P, (int)n_ref(H), W, n_dom, n_asc, false).total;
}

int cd_job_create(const cd_job_desc *desc, cd_job **out) {
    if (!out) return fail(CD_E_INVALID, "null output handle");
    int rc = validate(desc);
    if (rc != CD_OK) return rc;
    return create_job(desc, false, 0, out);
}

int cd_job_init(cd_job *J) {
    if (!J || J->selection_only) return fail(CD_E_INVALID, "null job");
    // the offspring work list doubles as the (job, design) list of the init launch
    HIPCHK(enqueue_designs(job_launch(J), cd::MODE_INIT, J->d_off, J->N, J->stream));
    HIPCHK(cd::launch_init_state(J->d_self, 1, J->stream));
    J->inited = true;
    return CD_OK;
}

int cd_job_step(cd_job *J, int32_t generations) {
    if (!J || J->selection_only) return fail(CD_E_INVALID, "null job");
    if (!J->inited) return fail(CD_E_STATE, "cd_job_step before cd_job_init");
    if (!J->batches.empty()) return fail(CD_E_STATE, "the job is held by a batch: advance it through the batch");
    if (generations < 0) return fail(CD_E_INVALID, "negative generation count");
    return step_graphs(J->graphs, J->stream, job_launch(J), generations);
}

int cd_job_step_timed(cd_job *J, int32_t generations, float *ms) {
    if (!J) return fail(CD_E_INVALID, "null job");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, J->stream));
    int rc = cd_job_step(J, generations);
    if (rc != CD_OK) return rc;
    HIPCHK(hipEventRecord(b, J->stream));
    HIPCHK(hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    if (ms) *ms = t;
    HIPCHK(hipEventDestroy(a));
    HIPCHK(hipEventDestroy(b));
    return CD_OK;
}

int cd_job_sync(cd_job *J) {
    if (!J) return fail(CD_E_INVALID, "null job");
    HIPCHK(sync_job(J));
    return CD_OK;
}

// Reads of a job's device state go through its own (non-blocking) stream,
// never the legacy null stream, so that they neither wait for nor disturb
// other streams -- for instance a served tick being captured in another
// thread.
// The survivors in ascending pool index: the first N set bits of the
// survivor mask.
static void survivors_of(const std::vector<uint64_t> &mask, int N, int32_t *out) {
    int k = 0;
    for (size_t w = 0; w < mask.size() && k < N; ++w)
        for (uint64_t x = mask[w]; x && k < N; x &= x - 1) out[k++] = (int32_t)(64 * w + __builtin_ctzll(x));
    for (; k < N; ++k) out[k] = -1;
}

static hipError_t read_state(cd_job *J, cd::SelState &st) {
    hipError_t e = sync_job(J);
    if (e == hipSuccess) e = hipMemcpyAsync(&st, J->h.st, sizeof(st), hipMemcpyDeviceToHost, J->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(J->stream);
    return e;
}

int cd_job_generation(cd_job *J, int32_t *generation) {
    if (!J || !generation) return fail(CD_E_INVALID, "null argument");
    cd::SelState st;
    HIPCHK(read_state(J, st));
    *generation = (int32_t)st.gen;
    return CD_OK;
}

int cd_job_read(cd_job *J, int8_t *genome, float *obj, int8_t *idx, int32_t *pql, int32_t *sorted, int32_t *front_of,
                float *state) {
    if (!J || J->selection_only) return fail(CD_E_INVALID, "null job");
    if (!J->inited) return fail(CD_E_STATE, "cd_job_read before cd_job_init");
    cd::SelState st;
    HIPCHK(read_state(J, st));
    const int b = (int)(st.gen & 1u), P = J->P;
    const size_t rows = (size_t)P * J->m, L = J->L, Ls = J->h.Ls;
    hipStream_t s = J->stream;
    std::unique_lock<std::mutex> stage(g_rows.mu, std::defer_lock);
    if (genome) {  // the padded rows in one transfer, compacted below
        stage.lock();
        HIPCHK(g_rows.reserve(rows * Ls));
        HIPCHK(hipMemcpyAsync(g_rows.p, J->h.genome[b], rows * Ls, hipMemcpyDeviceToHost, s));
    }
    if (obj) HIPCHK(hipMemcpyAsync(obj, J->h.obj[b], (size_t)P * 6 * 4, hipMemcpyDeviceToHost, s));
    if (idx) HIPCHK(hipMemcpyAsync(idx, J->h.oidx[b], (size_t)P * 12, hipMemcpyDeviceToHost, s));
    std::vector<int32_t> t(pql ? (size_t)P * 4 : 0);
    if (pql) HIPCHK(hipMemcpyAsync(t.data(), J->h.pql[b], t.size() * 4, hipMemcpyDeviceToHost, s));
    std::vector<uint64_t> mask(sorted ? J->W : 0);
    if (sorted) HIPCHK(hipMemcpyAsync(mask.data(), J->h.selm, (size_t)J->W * 8, hipMemcpyDeviceToHost, s));
    if (front_of) HIPCHK(hipMemcpyAsync(front_of, J->h.front_of, (size_t)P * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (genome)  // CDS rows of L codons from rows of stride Ls
        for (size_t r = 0; r < rows; ++r) std::memcpy(genome + r * L, g_rows.p + r * Ls, L);
    if (sorted) survivors_of(mask, J->N, sorted);
    if (pql)
        for (int i = 0; i < P; ++i)
            for (int k = 0; k < 3; ++k) pql[i * 3 + k] = t[(size_t)i * 4 + k];
    if (state) {
        std::memcpy(state, st.ext, sizeof(st.ext));
        std::memcpy(state + 36, st.ideal, sizeof(st.ideal));
        std::memcpy(state + 42, st.worst, sizeof(st.worst));
        std::memcpy(state + 48, st.nadir, sizeof(st.nadir));
    }
    return CD_OK;
}

int cd_job_profile(cd_job *J, int32_t generations, float *kernel_ms) {
    if (!J || J->selection_only || !kernel_ms) return fail(CD_E_INVALID, "null argument");
    if (!J->inited) return fail(CD_E_STATE, "cd_job_profile before cd_job_init");
    if (!J->batches.empty()) return fail(CD_E_STATE, "the job is held by a batch: advance it through the batch");
    return profile_generations(job_launch(J), J->stream, generations, kernel_ms);
}

int cd_job_cds_scanned(cd_job *J, int64_t *count) {
    if (!J || !count) return fail(CD_E_INVALID, "null argument");
    cd::SelState st;
    HIPCHK(read_state(J, st));
    *count = (int64_t)st.cds_scanned;
    return CD_OK;
}

int cd_job_destroy(cd_job *J) {
    if (!J) return CD_OK;
    if (!J->batches.empty()) return fail(CD_E_STATE, "the job is held by a batch: destroy the batch first");
    if (J->stream) (void)hipStreamSynchronize(J->stream);
    J->graphs.reset();
    if (J->mem) (void)hipFreeAsync(J->mem, J->stream);  // back to the pool, after the job's last work
    if (J->stream) {
        (void)hipStreamSynchronize(J->stream);
        (void)hipStreamDestroy(J->stream);
    }
    delete J;
    return CD_OK;
}

int cd_eval_batch(const cd_job_desc *desc, const int8_t *genome, int32_t count, float *obj, int8_t *idx,
                  int32_t *pql) {
    if (!desc || !genome || count <= 0) return fail(CD_E_INVALID, "bad eval arguments");
    cd_job_desc d = *desc;
    d.N = (count + 1) / 2;
    d.H = 1;
    d.Pm = 0.f;
    int rc = validate(&d);
    if (rc != CD_OK) return rc;
    cd_job *J = nullptr;
    rc = create_job(&d, false, 0, &J);
    if (rc != CD_OK) return rc;
    std::vector<int2> items(count);
    for (int i = 0; i < count; ++i) items[i] = make_int2(0, i);
    int2 *d_items = nullptr;
    hipError_t e = hipMalloc(&d_items, sizeof(int2) * count);
    if (e == hipSuccess) e = hipMemcpy(d_items, items.data(), sizeof(int2) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) {  // CDS rows of L codons into rows of stride Ls (zero pads), one transfer
        const size_t rows = (size_t)count * J->m, L = J->L, Ls = J->h.Ls;
        std::lock_guard<std::mutex> lk(g_rows.mu);
        e = g_rows.reserve(rows * Ls);
        if (e == hipSuccess) {
            for (size_t r = 0; r < rows; ++r) {
                std::memcpy(g_rows.p + r * Ls, genome + r * L, L);
                std::memset(g_rows.p + r * Ls + L, 0, Ls - L);
            }
            e = hipMemcpy(J->h.genome[0], g_rows.p, rows * Ls, hipMemcpyHostToDevice);
        }
    }
    if (e == hipSuccess) e = enqueue_designs(job_launch(J), cd::MODE_EVAL, d_items, count, J->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(J->stream);
    if (e == hipSuccess && obj) e = hipMemcpy(obj, J->h.obj[0], (size_t)count * 24, hipMemcpyDeviceToHost);
    if (e == hipSuccess && idx) e = hipMemcpy(idx, J->h.oidx[0], (size_t)count * 12, hipMemcpyDeviceToHost);
    std::vector<int32_t> t((size_t)count * 4);
    if (e == hipSuccess && pql) e = hipMemcpy(t.data(), J->h.pql[0], t.size() * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && pql)
        for (int i = 0; i < count; ++i)
            for (int k = 0; k < 3; ++k) pql[i * 3 + k] = t[(size_t)i * 4 + k];
    if (d_items) (void)hipFree(d_items);
    cd_job_destroy(J);
    if (e != hipSuccess) return fail(CD_E_DEVICE, std::string("eval: ") + hipGetErrorString(e));
    return CD_OK;
}

static int run_selection(const float *obj, int32_t P, int32_t N, int32_t H, uint32_t gen, uint64_t seed,
                         const float *extremes, bool full, cd_job **out) {
    if (!obj || P <= 0 || N <= 0 || N > P || P > 64 * cd::MAX_W || H < 1)
        return fail(CD_E_INVALID, "bad selection arguments");
    cd_job_desc d{};
    d.N = N;
    d.H = H;
    d.seed = seed;
    cd_job *J = nullptr;
    int rc = create_job(&d, true, P, &J);
    if (rc != CD_OK) return rc;
    cd::SelState st{};
    st.gen = gen;
    if (extremes) std::memcpy(st.ext, extremes, sizeof(st.ext));
    hipError_t e = hipMemcpy(J->h.obj[0], obj, (size_t)P * 24, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(J->h.st, &st, sizeof(st), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = cd::launch_zero_np(J->d_self, 1, J->stream);
    const cd::Shapes &sh = J->shapes;
    if (e == hipSuccess) e = cd::launch_dominance(J->d_self, J->d_dom, J->n_dom, false, nullptr, 0, sh.dom_rows, J->stream);
    if (e == hipSuccess) e = cd::launch_fronts(J->d_self, 1, J->P, sh.fronts_threads, J->stream);
    if (e == hipSuccess && full) e = cd::launch_associate(J->d_self, J->d_asc, J->n_asc, sh.assoc_designs, J->stream);
    if (e == hipSuccess && full) e = cd::launch_niche(J->d_self, 1, J->P, J->R, sh.niche_threads, J->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(J->stream);
    if (e != hipSuccess) {
        cd_job_destroy(J);
        return fail(CD_E_DEVICE, std::string("selection: ") + hipGetErrorString(e));
    }
    *out = J;
    return CD_OK;
}

int cd_nds(const float *obj, int32_t P, int32_t N, int32_t *front_of, int32_t *n_fronts) {
    cd_job *J = nullptr;
    int rc = run_selection(obj, P, N, 1, 0, 0, nullptr, false, &J);
    if (rc != CD_OK) return rc;
    cd::SelState st;
    hipError_t e = hipMemcpy(&st, J->h.st, sizeof(st), hipMemcpyDeviceToHost);
    if (e == hipSuccess && front_of) e = hipMemcpy(front_of, J->h.front_of, (size_t)P * 4, hipMemcpyDeviceToHost);
    if (n_fronts) *n_fronts = st.last + 1;
    cd_job_destroy(J);
    if (e != hipSuccess) return fail(CD_E_DEVICE, std::string("nds: ") + hipGetErrorString(e));
    return CD_OK;
}

int cd_select(const float *obj, int32_t N, int32_t H, uint32_t gen, uint64_t seed, float *extremes, int32_t *sorted,
              int32_t *front_of, float *norm, int32_t *flags) {
    cd_job *J = nullptr;
    int rc = run_selection(obj, 2 * N, N, H, gen, seed, extremes, true, &J);
    if (rc != CD_OK) return rc;
    cd::SelState st;
    hipError_t e = hipMemcpy(&st, J->h.st, sizeof(st), hipMemcpyDeviceToHost);
    std::vector<uint64_t> mask(J->W);
    if (e == hipSuccess && sorted) e = hipMemcpy(mask.data(), J->h.selm, (size_t)J->W * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && sorted) survivors_of(mask, N, sorted);
    if (e == hipSuccess && front_of) e = hipMemcpy(front_of, J->h.front_of, (size_t)2 * N * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
        if (extremes) std::memcpy(extremes, st.ext, sizeof(st.ext));
        if (norm) {
            std::memcpy(norm, st.ideal, 24);
            std::memcpy(norm + 6, st.worst, 24);
            std::memcpy(norm + 12, st.nadir, 24);
        }
        if (flags) {
            flags[0] = st.n_cut;
            flags[1] = st.hyp;
        }
    }
    cd_job_destroy(J);
    if (e != hipSuccess) return fail(CD_E_DEVICE, std::string("select: ") + hipGetErrorString(e));
    return CD_OK;
}

// ---------------------------------------------------------------------------
// Batched form
// ---------------------------------------------------------------------------
int cd_batch_create(cd_job *const *jobs, int32_t count, cd_batch **out) {
    if (!jobs || count <= 0 || !out) return fail(CD_E_INVALID, "bad batch arguments");
    for (int j = 0; j < count; ++j) {
        cd_job *J = jobs[j];
        if (!J || J->selection_only || !J->inited) return fail(CD_E_STATE, "batch jobs must be created and initialised");
        for (int k = 0; k < j; ++k)
            if (jobs[k] == J) return fail(CD_E_INVALID, "a job appears twice in the batch");
    }
    cd_batch *B = new cd_batch();
    for (int j = 0; j < count; ++j) {
        (void)hipStreamSynchronize(jobs[j]->stream);
        B->jobs.push_back(jobs[j]);
    }
    cd::Shapes sh;
    if (cdhost::cached_shapes(cdhost::work_signature(jobs, count), &sh)) B->shapes = sh;
    cdhost::WorkLists wl;
    cdhost::build_lists(jobs, count, B->shapes, wl);
    // room for the lists in any shape: 64-row tiles and 4-design groups give the most items
    size_t cap_dom = 0, cap_asc = 0;
    for (int j = 0; j < count; ++j) {
        int nd = 0, na = 0;
        work_counts(jobs[j]->P, jobs[j]->W, &nd, &na);
        cap_dom += nd;
        cap_asc += na;
    }
    B->lds = wl.lds;
    B->stem_lds = wl.stem_lds;
    B->max_pool = wl.max_pool;
    B->max_refs = wl.max_refs;
    B->mg = wl.mg;
    B->stem_parts = wl.stem_parts;
    B->n_cds = wl.n_cds;
    B->n_small = wl.n_small;
    B->lds_small = wl.lds_small;
    B->n_off = (int)wl.off.size();
    B->n_dom = (int)wl.dom.size();
    B->n_asc = (int)wl.asc.size();
    hipError_t e = hipMalloc(&B->d_jobs, sizeof(cd::DevJob) * count);
    if (e == hipSuccess) e = hipMalloc(&B->d_off, sizeof(int2) * wl.off.size());
    if (e == hipSuccess) e = hipMalloc(&B->d_dom, sizeof(int4) * cap_dom);
    if (e == hipSuccess) e = hipMalloc(&B->d_asc, sizeof(int2) * cap_asc);
    if (e == hipSuccess) e = hipMalloc(&B->d_slist, sizeof(uint32_t) * cd::stem_list_entries((int)wl.off.size(), wl.mg));
    if (e == hipSuccess) e = hipMalloc(&B->d_sctl, sizeof(int32_t) * cd::STEM_MAX_GROUPS);
    if (e == hipSuccess) e = hipMemset(B->d_sctl, 0, sizeof(int32_t) * cd::STEM_MAX_GROUPS);
    if (e == hipSuccess) e = hipMemcpy(B->d_jobs, wl.descs.data(), sizeof(cd::DevJob) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_off, wl.off.data(), sizeof(int2) * wl.off.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_dom, wl.dom.data(), sizeof(int4) * wl.dom.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_asc, wl.asc.data(), sizeof(int2) * wl.asc.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamCreate(&B->stream);
    if (e == hipSuccess) e = cdhost::set_lds_limits(wl);
    if (e != hipSuccess) {
        cd_batch_destroy(B);
        return fail(CD_E_DEVICE, std::string("batch setup: ") + hipGetErrorString(e));
    }
    for (cd_job *J : B->jobs) J->batches.push_back(B);
    *out = B;
    return CD_OK;
}

static Launch batch_launch(cd_batch *B) {
    return Launch{B->d_jobs,    (int)B->jobs.size(), B->d_off,      B->n_off,   B->d_dom,
                  B->n_dom,     B->d_asc,            B->n_asc,      B->lds,     B->stem_lds,
                  B->max_pool,  B->max_refs,         B->mg,         B->stem_parts, B->d_slist, B->d_sctl,
                  B->n_cds,     B->n_small,          B->lds_small,  B->shapes};
}

int cd_batch_profile(cd_batch *B, int32_t generations, float *kernel_ms) {
    if (!B || !kernel_ms) return fail(CD_E_INVALID, "null argument");
    return profile_generations(batch_launch(B), B->stream, generations, kernel_ms);
}

int cd_batch_step(cd_batch *B, int32_t generations) {
    if (!B) return fail(CD_E_INVALID, "null batch");
    if (generations < 0) return fail(CD_E_INVALID, "negative generation count");
    return step_graphs(B->graphs, B->stream, batch_launch(B), generations);
}

int cd_batch_sync(cd_batch *B) {
    if (!B) return fail(CD_E_INVALID, "null batch");
    HIPCHK(hipStreamSynchronize(B->stream));
    return CD_OK;
}

int cd_batch_step_timed(cd_batch *B, int32_t generations, float *ms) {
    if (!B) return fail(CD_E_INVALID, "null batch");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, B->stream));
    int rc = cd_batch_step(B, generations);
    if (rc != CD_OK) return rc;
    HIPCHK(hipEventRecord(b, B->stream));
    HIPCHK(hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    if (ms) *ms = t;
    HIPCHK(hipEventDestroy(a));
    HIPCHK(hipEventDestroy(b));
    return CD_OK;
}

// Install shapes on a batch: its work lists in that shape, the captured
// graph dropped (the next step captures the new launches).
static int batch_apply_shapes(cd_batch *B, const cd::Shapes &sh) {
    cdhost::WorkLists wl;
    cdhost::build_lists(B->jobs.data(), (int)B->jobs.size(), sh, wl);
    HIPCHK(hipStreamSynchronize(B->stream));
    HIPCHK(hipMemcpy(B->d_off, wl.off.data(), sizeof(int2) * wl.off.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(B->d_dom, wl.dom.data(), sizeof(int4) * wl.dom.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(B->d_asc, wl.asc.data(), sizeof(int2) * wl.asc.size(), hipMemcpyHostToDevice));
    B->n_dom = (int)wl.dom.size();
    B->n_asc = (int)wl.asc.size();
    B->n_small = wl.n_small;
    B->lds_small = wl.lds_small;
    B->shapes = sh;
    B->graphs.reset();
    return CD_OK;
}

static int job_apply_shapes(cd_job *J, const cd::Shapes &sh) {
    std::vector<int4> dom;
    std::vector<int2> asc;
    job_lists(J->P, J->W, sh, dom, asc);
    HIPCHK(hipStreamSynchronize(J->stream));
    HIPCHK(hipMemcpy(J->d_dom, dom.data(), sizeof(int4) * dom.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(J->d_asc, asc.data(), sizeof(int2) * asc.size(), hipMemcpyHostToDevice));
    J->n_dom = (int)dom.size();
    J->n_asc = (int)asc.size();
    J->shapes = sh;
    J->graphs.reset();
    return CD_OK;
}

static_assert(sizeof(cd_shapes) == sizeof(cd::Shapes), "cd_shapes mirrors cd::Shapes");
static cd_shapes to_abi(const cd::Shapes &s) {
    cd_shapes o;
    std::memcpy(&o, &s, sizeof(o));
    return o;
}
static cd::Shapes from_abi(const cd_shapes &s) {
    cd::Shapes o;
    std::memcpy(&o, &s, sizeof(o));
    return o;
}

int cd_batch_shapes(cd_batch *B, cd_shapes *get, const cd_shapes *set) {
    if (!B) return fail(CD_E_INVALID, "null batch");
    if (set) {
        const cd::Shapes sh = from_abi(*set);
        if (!cd::shapes_valid(sh)) return fail(CD_E_INVALID, "invalid launch shapes");
        const int rc = batch_apply_shapes(B, sh);
        if (rc != CD_OK) return rc;
    }
    if (get) *get = to_abi(B->shapes);
    return CD_OK;
}

int cd_job_shapes(cd_job *J, cd_shapes *get, const cd_shapes *set) {
    if (!J || J->selection_only) return fail(CD_E_INVALID, "null job");
    if (set) {
        const cd::Shapes sh = from_abi(*set);
        if (!cd::shapes_valid(sh)) return fail(CD_E_INVALID, "invalid launch shapes");
        if (!J->batches.empty()) return fail(CD_E_STATE, "the job is held by a batch");
        const int rc = job_apply_shapes(J, sh);
        if (rc != CD_OK) return rc;
    }
    if (get) *get = to_abi(J->shapes);
    return CD_OK;
}

int cd_batch_tune(cd_batch *B, int32_t gens, cd_shapes *chosen, int32_t *generations_run) {
    if (!B || gens < 1) return fail(CD_E_INVALID, "bad tune arguments");
    cd::Shapes best;
    int run = 0;
    const int rc = tune_shapes(
        B->shapes, (int)B->jobs.size(), gens, B->stream,
        [&](const cd::Shapes &sh, Launch *w) {
            const int r = batch_apply_shapes(B, sh);
            *w = batch_launch(B);
            return r;
        },
        &best, &run);
    if (generations_run) *generations_run = run;
    if (rc != CD_OK) return rc;
    cdhost::store_shapes(cdhost::work_signature(B->jobs.data(), (int)B->jobs.size()), best);
    if (chosen) *chosen = to_abi(best);
    return CD_OK;
}

int cd_job_tune(cd_job *J, int32_t gens, cd_shapes *chosen, int32_t *generations_run) {
    if (!J || J->selection_only || gens < 1) return fail(CD_E_INVALID, "bad tune arguments");
    if (!J->inited) return fail(CD_E_STATE, "cd_job_tune before cd_job_init");
    if (!J->batches.empty()) return fail(CD_E_STATE, "the job is held by a batch: tune the batch");
    cd::Shapes best;
    int run = 0;
    const int rc = tune_shapes(
        J->shapes, 1, gens, J->stream,
        [&](const cd::Shapes &sh, Launch *w) {
            const int r = job_apply_shapes(J, sh);
            *w = job_launch(J);
            return r;
        },
        &best, &run);
    if (generations_run) *generations_run = run;
    if (rc != CD_OK) return rc;
    cd_job *self = J;
    cdhost::store_shapes(cdhost::work_signature(&self, 1), best);
    if (chosen) *chosen = to_abi(best);
    return CD_OK;
}

int cd_batch_destroy(cd_batch *B) {
    if (!B) return CD_OK;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    for (cd_job *J : B->jobs) J->batches.erase(std::remove(J->batches.begin(), J->batches.end(), B), J->batches.end());
    B->graphs.reset();
    if (B->stream) (void)hipStreamDestroy(B->stream);
    if (B->d_jobs) (void)hipFree(B->d_jobs);
    if (B->d_off) (void)hipFree(B->d_off);
    if (B->d_dom) (void)hipFree(B->d_dom);
    if (B->d_asc) (void)hipFree(B->d_asc);
    if (B->d_slist) (void)hipFree(B->d_slist);
    if (B->d_sctl) (void)hipFree(B->d_sctl);
    delete B;
    return CD_OK;
}

}  // extern "C"
