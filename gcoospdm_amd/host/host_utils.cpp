// host_utils.cpp -- see host_utils.h.
#include "host_utils.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace gcoo_host {

long long random_init_ref(float *data, size_t size, float sparsity)
{
    const int num = 1000000;
    const float cut = sparsity * num;
    long long nnz = 0;
    for (size_t i = 0; i < size; ++i) {
        const int r = rand();
        if ((float)(r % num) < cut) {
            data[i] = 0.0f;
        } else {
            data[i] = r / (float)RAND_MAX;
            ++nnz;
        }
    }
    return nnz;
}

namespace {
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

uint64_t keep_threshold(double sparsity)
{
    if (sparsity <= 0.0) return 0;
    if (sparsity >= 1.0) return 1ull << 40;
    return (uint64_t)std::floor(sparsity * 1099511627776.0);  // 2^40
}

long long fill_counter(float *data, size_t size, uint64_t seed, uint64_t stream_id, float scale,
                       uint64_t keep_thr, bool sparse)
{
    const uint64_t key = mix64(seed * kGamma + stream_id);
    long long nnz = 0;
    for (size_t i = 0; i < size; ++i) {
        const uint64_t h = mix64(key + (uint64_t)(i + 1) * kGamma);
        const int num = 2 * (int)(h >> 40) - (1 << 24) + 1;
        float v = (float)num * 0x1p-24f * scale;
        if (sparse && (h & ((1ull << 40) - 1)) < keep_thr) v = 0.0f;
        data[i] = v;
        nnz += v != 0.0f;
    }
    return nnz;
}

double rel_l2_vs_fp64(const float *S, const float *D, const float *C, int M, int K, int N,
                      int threads)
{
    threads = std::max(1, threads);
    std::vector<double> num(threads, 0.0), den(threads, 0.0);
    auto work = [&](int t) {
        std::vector<double> acc(N);
        for (int i = t; i < M; i += threads) {
            std::fill(acc.begin(), acc.end(), 0.0);
            const float *srow = S + (size_t)i * K;
            for (int k = 0; k < K; ++k) {
                const double a = srow[k];
                if (a == 0.0) continue;
                const float *drow = D + (size_t)k * N;
                for (int j = 0; j < N; ++j) acc[j] += a * (double)drow[j];
            }
            const float *crow = C + (size_t)i * N;
            for (int j = 0; j < N; ++j) {
                const double ref = (double)(float)acc[j];
                const double d = (double)crow[j] - ref;
                num[t] += d * d;
                den[t] += ref * ref;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    double n = 0, d = 0;
    for (int t = 0; t < threads; ++t) {
        n += num[t];
        d += den[t];
    }
    if (d == 0.0) return n == 0.0 ? 0.0 : INFINITY;
    return std::sqrt(n / d);
}

const char *Flags::find(const char *name) const
{
    const size_t len = std::strlen(name);
    for (int i = 1; i < argc_; ++i) {
        const char *a = argv_[i];
        while (*a == '-') ++a;
        if (std::strncmp(a, name, len) == 0 && (a[len] == '=' || a[len] == '\0'))
            return a + len;
    }
    return nullptr;
}

bool Flags::has(const char *name) const { return find(name) != nullptr; }

long long Flags::get_int(const char *name, long long def) const
{
    const char *v = find(name);
    if (!v || *v != '=') return def;
    return std::strtoll(v + 1, nullptr, 10);
}

double Flags::get_float(const char *name, double def) const
{
    const char *v = find(name);
    if (!v || *v != '=') return def;
    return std::strtod(v + 1, nullptr);
}

std::string Flags::get_str(const char *name, const std::string &def) const
{
    const char *v = find(name);
    if (!v || *v != '=') return def;
    return std::string(v + 1);
}

bool read_matrix_market(const std::string &path, MtxMatrix &out, std::string &err)
{
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) {
        err = "cannot find file: " + path;
        return false;
    }
    std::vector<char> line(4096);
    auto next = [&]() { return std::fgets(line.data(), (int)line.size(), fp) != nullptr; };
    auto fail = [&](const std::string &why) {
        std::fclose(fp);
        err = path + ": " + why;
        return false;
    };
    if (!next()) return fail("empty file");
    const bool general = std::strstr(line.data(), "general") != nullptr;
    bool more;
    do {
        more = next();
    } while (more && line[0] == '%');
    if (!more) return fail("no size line");
    long long r = 0, c = 0, n = 0;
    if (std::sscanf(line.data(), "%lld %lld %lld", &r, &c, &n) != 3 || r < 0 || c < 0 ||
        r > 2147483647ll || c > 2147483647ll)
        return fail("bad size line");
    out.nRows = (int)r;
    out.nCols = (int)c;
    out.nnz = 0;
    out.dense.assign((size_t)r * (size_t)c, 0.0f);
    while (next()) {
        char *ch = line.data();
        if (std::strspn(ch, " \t\r\n") == std::strlen(ch)) continue;  // blank line
        const long long row = std::atoll(ch) - 1;
        ch = std::strchr(ch, ' ');
        if (!ch) return fail("entry line without a column index");
        ++ch;
        const long long col = std::atoll(ch) - 1;
        ch = std::strchr(ch, ' ');
        float val = (float)std::rand();  // drawn for every entry, used when none is given
        if (ch) val = (float)std::atof(ch + 1);
        if (row < 0 || row >= r || col < 0 || col >= c)
            return fail("entry (" + std::to_string(row + 1) + ", " + std::to_string(col + 1) +
                        ") outside the matrix");
        out.dense[(size_t)row * (size_t)c + (size_t)col] = val;
        ++out.nnz;
        if (!general) {
            if (col >= r || row >= c) return fail("mirrored entry outside the matrix");
            out.dense[(size_t)col * (size_t)c + (size_t)row] = val;
            if (col != row) ++out.nnz;
        }
    }
    std::fclose(fp);
    return true;
}

}  // namespace gcoo_host
