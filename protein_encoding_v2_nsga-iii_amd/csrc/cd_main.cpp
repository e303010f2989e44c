// cd_main.cpp -- command-line driver with the reference program's interface
// (src/main.cu:456-901) on top of the C ABI:
//
//   cd_main FASTA N G m Pm [--H h] [--seed s] [--tables reference_tables.json]
//
// Reads the protein (first line skipped, stop appended, main.cu:497-523),
// derives H from N as main.cu:564-577 does unless --H is given, runs one
// initialisation and G generations on the GPU, and writes Result.txt and
// Normalized_value_quality_computation.txt in the reference's formats
// (main.cu:797-844) before printing the hypervolume, the minimum distance to
// the ideal point and the timings (main.cu:846-863).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "codon_nsga3.h"

namespace {

const char AMINO[] = "ACDEFGHIKLMNPQRSTVWYZ";
const int SYN_START[21] = {0, 4, 6, 8, 10, 12, 16, 18, 21, 23, 29, 30, 32, 36, 38, 44, 50, 54, 58, 59, 61};

int die(const char *msg) {
    std::printf("%s\n", msg);
    return EXIT_FAILURE;
}

// Minimal reader for the table file: "codons": "...", "weights": [...],
// "cps": [...], "gc3_ref": x.
bool read_tables(const std::string &path, std::string &codons, std::vector<float> &w, std::vector<float> &cps,
                 float &gc3) {
    std::ifstream f(path);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string t = ss.str();
    auto key = [&](const char *k) { return t.find(std::string("\"") + k + "\""); };
    size_t p = key("codons");
    if (p == std::string::npos) return false;
    p = t.find('"', t.find(':', p) + 1);
    codons = t.substr(p + 1, t.find('"', p + 1) - p - 1);
    auto array = [&](const char *k, std::vector<float> &out) {
        size_t a = t.find('[', key(k));
        size_t b = t.find(']', a);
        std::stringstream s(t.substr(a + 1, b - a - 1));
        std::string item;
        while (std::getline(s, item, ',')) out.push_back((float)std::strtod(item.c_str(), nullptr));
    };
    array("weights", w);
    array("cps", cps);
    size_t g = key("gc3_ref");
    gc3 = (float)std::strtod(t.c_str() + t.find(':', g) + 1, nullptr);
    return codons.size() == 192 && w.size() == 64 && cps.size() == 61 * 61;
}

long long comb(int n, int k) {
    long long r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 6)
        return die("usage: cd_main FASTA N G m Pm [--H h] [--seed s] [--tables reference_tables.json]");
    const int N = std::atoi(argv[2]), G = std::atoi(argv[3]), m = std::atoi(argv[4]);
    const float Pm = (float)std::atof(argv[5]);
    if (N <= 0 || G < 0 || m <= 1 || Pm < 0.f || Pm > 1.f)  // main.cu:485-489
        return die("Please check the input parameters (N > 0, G >= 0, m > 1, 0 <= Pm <= 1)");
    int H = 0;
    unsigned long long seed = (unsigned long long)std::time(nullptr);
    std::string tables_path;
    {
        std::string exe = argv[0];
        const size_t slash = exe.rfind('/');
        tables_path = (slash == std::string::npos ? std::string(".") : exe.substr(0, slash)) + "/data/reference_tables.json";
    }
    for (int i = 6; i + 1 < argc; i += 2) {
        if (!std::strcmp(argv[i], "--H")) H = std::atoi(argv[i + 1]);
        else if (!std::strcmp(argv[i], "--seed")) seed = std::strtoull(argv[i + 1], nullptr, 10);
        else if (!std::strcmp(argv[i], "--tables")) tables_path = argv[i + 1];
    }
    // FASTA: skip the header line, keep every other byte except newlines, the
    // final byte becomes the stop 'Z' (main.cu:497-523)
    std::ifstream fa(argv[1], std::ios::binary);
    if (!fa) return die("Opening Protein FASTA format file is failed..");
    std::string line, body;
    std::getline(fa, line);
    char c;
    while (fa.get(c))
        if (c != '\n') body.push_back(c);
    std::vector<int8_t> aa;
    for (char r : body) {
        const char *p = std::strchr(AMINO, r);
        if (!p || r == 'Z' || r == '\0') return die("FindAminoIndex Function failure");
        aa.push_back((int8_t)(p - AMINO));
    }
    aa.push_back(20);
    const int L = (int)aa.size();
    if (H <= 0) {  // main.cu:564-577: the largest p with C(5+p, p) <= N; at least 1 (the
                   // reference's 0 for N < 6 is a single all-zero reference direction)
        int p = 1;
        while (comb(5 + p, p) <= N) ++p;
        H = p - 1 < 1 ? 1 : p - 1;
    }
    std::printf("Reference point divisions (H) : %d, reference points : %lld\n", H, comb(H + 5, 5));
    std::string codons;
    std::vector<float> w, cps;
    float gc3 = 0.381f;
    if (!read_tables(tables_path, codons, w, cps, gc3)) return die("cannot read the codon tables");

    cd_job_desc d{};
    d.L = L;
    d.m = m;
    d.N = N;
    d.H = H;
    d.Pm = Pm;
    d.seed = seed;
    d.aa = aa.data();
    d.tables.codons = codons.c_str();
    d.tables.weights = w.data();
    d.tables.cps = cps.data();
    d.tables.gc3_ref = gc3;
    cd_job *job = nullptr;
    if (cd_job_create(&d, &job) != CD_OK) return die(cd_last_error());
    std::printf("Device memory usage : %lld bytes\n", (long long)cd_job_device_bytes(L, m, N, H));
    auto t0 = std::chrono::steady_clock::now();
    if (cd_job_init(job) != CD_OK || cd_job_sync(job) != CD_OK) return die(cd_last_error());
    auto t1 = std::chrono::steady_clock::now();
    if (cd_job_step(job, G) != CD_OK || cd_job_sync(job) != CD_OK) return die(cd_last_error());
    auto t2 = std::chrono::steady_clock::now();
    const double init_s = std::chrono::duration<double>(t1 - t0).count();
    const double gen_s = std::chrono::duration<double>(t2 - t1).count();

    const int P = 2 * N;
    std::vector<int8_t> genome((size_t)P * m * L), idx((size_t)P * 12);
    std::vector<float> obj((size_t)P * 6);
    if (cd_job_read(job, genome.data(), obj.data(), idx.data(), nullptr, nullptr, nullptr, nullptr) != CD_OK)
        return die(cd_last_error());
    cd_job_destroy(job);

    // Result.txt, main.cu:789-818
    FILE *fp = std::fopen("Result.txt", "w");
    for (int i = 0; i < P; ++i) {
        std::fprintf(fp, "%d Solutions\n", i + 1);
        for (int j = 0; j < m; ++j) {
            std::fprintf(fp, "%d CDS : ", j + 1);
            for (int k = 0; k < L; ++k) {
                const int g = SYN_START[aa[k]] + genome[((size_t)i * m + j) * L + k];
                std::fprintf(fp, "%.3s", codons.c_str() + 3 * g);
            }
            std::fprintf(fp, "\n");
        }
        const float *o = &obj[(size_t)i * 6];
        const int8_t *x = &idx[(size_t)i * 12];
        std::fprintf(fp, "\n %d Solution\n", i + 1);
        std::fprintf(fp, "%d\tmCAI : %f\n", x[0], -o[0]);
        std::fprintf(fp, "%d\tmCBP : %f\n", x[2], -o[1]);
        std::fprintf(fp, "%d\tmHSC : %f\n", x[4], -o[2]);
        std::fprintf(fp, "%d %d\tmHD : %f\n", x[6], x[7], -o[3]);
        std::fprintf(fp, "%d\tMGC : %f\n", x[8], o[4]);
        std::fprintf(fp, "%d\tMSL : %f\n", x[10], o[5]);
    }
    std::fclose(fp);
    // Normalized_value_quality_computation.txt, main.cu:820-844
    std::vector<float> z((size_t)P * 6);
    cd_normalize(obj.data(), P, z.data());
    fp = std::fopen("Normalized_value_quality_computation.txt", "w");
    for (int i = 0; i < P; ++i) {
        const float *v = &z[(size_t)i * 6];
        std::fprintf(fp, "%f %f %f %f %f %f\n", v[0], v[1], v[2], v[3], v[4], v[5]);
    }
    std::fclose(fp);
    std::vector<double> zd(z.begin(), z.end());
    const double ref[6] = {1, 1, 1, 1, 1, 1};
    double hv = 0.0;
    float md = 0.f, sp = 0.f;
    cd_hypervolume(zd.data(), P, 6, ref, &hv);
    cd_min_distance_spread(z.data(), P, &md, &sp);
    std::printf("\n\nHypervolume : %f\n", hv);
    std::printf("Minimum Distance to Ideal Point : %f\n", md);
    std::printf("Initialization time : %f seconds\n", init_s);
    std::printf("Generation cycles time : %f seconds\n", gen_s);
    std::printf("Total time : %f seconds\n", init_s + gen_s);
    return EXIT_SUCCESS;
}
