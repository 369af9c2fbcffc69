// serve.cpp -- `gcoo_serve`: the serving workload driver.
//
// Builds `-jobs` resident operand sets (S converted to GCOO once, D dense,
// both in HBM; M, K log-uniform in [256, 8192], N in {32, 64, 128, 256},
// sparsity uniform in [0.95, 0.999], FP32 and BF16 forms of D), then drives
// the library's continuous-batching server (gcoo_server_*):
//   1. saturated capacity of each mode: every job submitted at once;
//   2. Poisson arrivals at each -loads fraction of that mode's own capacity,
//      `-requests` requests drawn uniformly over the jobs, in two modes:
//        batched  continuous batching (-slots streams, -max_batch per batch)
//        plain    a FIFO of plain multiply launches on one stream, one
//                 request per launch, up to 64 launches in flight
// and prints one JSON line per (dtype, mode, load) with p50 / p99 latency
// (submit -> completion), jobs/s and GFLOP/s (2 * nnz * N per job), plus the
// relative Frobenius error of a sample of results against FP64 products of
// the FP32 operands computed on the host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "gcoo_spdm.h"
#include "host_utils.h"

using gcoo_host::Flags;

#define CK(expr)                                                                       \
    do {                                                                               \
        int rc_ = (expr);                                                              \
        if (rc_ != GCOO_OK) {                                                          \
            std::fprintf(stderr, "%s failed (%d): %s\n", #expr, rc_, gcoo_last_error()); \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

namespace {

struct Job {
    int M, K, N;
    double s;
    uint64_t seed;
    gcoo_handle h = nullptr;
    float *D = nullptr;
    uint16_t *D16 = nullptr;
    long long nnz = 0;
    int opF32 = -1, opBF16 = -1;
};

// Tile lookup in the autotuner's JSON table (gcoospdm_amd/tuning.json):
// batch.<dtype>."n<i>r<j>" = {"ms": ..., "rows": R, "vec": V}.
struct TuneTable {
    std::string text;
    void load(const std::string &path)
    {
        std::ifstream f(path);
        if (f) {
            std::stringstream ss;
            ss << f.rdbuf();
            text = ss.str();
        }
    }
    static int nclass(int N) { return N <= 32 ? 0 : N <= 64 ? 1 : N <= 128 ? 2 : 3; }
    static int rclass(double per) { return per < 8 ? 0 : per < 32 ? 1 : per < 128 ? 2 : 3; }
    static int field(const std::string &t, size_t from, size_t to, const char *name)
    {
        const size_t p = t.find(name, from);
        if (p == std::string::npos || p > to) return 0;
        return std::atoi(t.c_str() + p + std::strlen(name));
    }
    void tile(bool bf16, int N, double per, int *rows, int *vec) const
    {
        *rows = *vec = 0;
        const size_t b = text.find("\"batch\"");
        if (b == std::string::npos) return;
        const size_t d = text.find(bf16 ? "\"bf16\"" : "\"f32\"", b);
        if (d == std::string::npos) return;
        char key[32];
        std::snprintf(key, sizeof(key), "\"n%dr%d\"", nclass(N), rclass(per));
        const size_t k = text.find(key, d);
        if (k == std::string::npos) return;
        const size_t end = text.find('}', k);
        *rows = field(text, k, end, "\"rows\": ");
        *vec = field(text, k, end, "\"vec\": ");
    }
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct RunResult {
    gcoo_server_stats_t st;
    double gflop;
};

// Submit `ops` (one request each) with the given arrival gaps (seconds,
// cumulative schedule); a collector thread waits for and releases results.
RunResult drive(gcoo_server_t srv, const std::vector<int> &ops, const std::vector<double> &gflop,
                const std::vector<double> &gaps)
{
    CK(gcoo_server_reset_stats(srv));
    std::vector<uint64_t> ids(ops.size(), 0);
    std::vector<char> submitted(ops.size(), 0);
    std::thread collector([&] {
        for (size_t i = 0; i < ops.size(); ++i) {
            while (!__atomic_load_n(&submitted[i], __ATOMIC_ACQUIRE))
                std::this_thread::sleep_for(std::chrono::microseconds(5));
            CK(gcoo_server_wait(srv, ids[i], nullptr));
            CK(gcoo_server_release(srv, ids[i]));
        }
    });
    const auto t0 = std::chrono::steady_clock::now();
    double t = 0;
    for (size_t i = 0; i < ops.size(); ++i) {
        t += gaps[i];
        if (gaps[i] > 0)
            std::this_thread::sleep_until(t0 + std::chrono::duration_cast<std::chrono::nanoseconds>(
                                                   std::chrono::duration<double>(t)));
        CK(gcoo_server_submit(srv, ops[i], &ids[i]));
        __atomic_store_n(&submitted[i], 1, __ATOMIC_RELEASE);
    }
    collector.join();
    RunResult r;
    CK(gcoo_server_stats(srv, &r.st));
    r.gflop = 0;
    for (int op : ops) r.gflop += gflop[op];
    return r;
}

}  // namespace

int main(int argc, char **argv)
{
    Flags f(argc, argv);
    const int nJobs = (int)f.get_int("jobs", 512);
    const int nReq = (int)f.get_int("requests", 20000);
    const uint64_t master = (uint64_t)f.get_int("seed", 2024);
    const size_t budget = (size_t)f.get_int("budget_mb", 2048) << 20;
    // 2 streams of batches of at most 64 jobs: at fixed arrival rates of
    // 0.8 / 0.95 x 230 k jobs/s (BF16) the p99 was 0.67-0.89 ms, against
    // 0.5-6 ms with 4 streams and up to 24 ms with 8 (more streams than the
    // process's 4 hardware queues, and concurrent batches that split the GPU,
    // made completions stall); larger batches lengthen the tail
    const int slots = (int)f.get_int("slots", 2);
    const int maxBatch = (int)f.get_int("max_batch", 64);
    // -rate_f32= / -rate_bf16=: the batched server's loads relative to this
    // rate (jobs/s) instead of its measured saturated capacity
    const double rateF32 = (double)f.get_int("rate_f32", 0);
    const double rateBF16 = (double)f.get_int("rate_bf16", 0);
    const int nCheck = (int)f.get_int("check", 16);
    const std::string loadsArg = f.get_str("loads", "0.5,0.8,0.95");
    const std::string dtypes = f.get_str("dtype", "both");
    const std::string table = f.get_str("table", "");
    std::vector<double> loads;
    {
        std::stringstream ss(loadsArg);
        std::string tok;
        while (std::getline(ss, tok, ',')) loads.push_back(std::atof(tok.c_str()));
    }
    TuneTable tune;
    if (!table.empty()) tune.load(table);

    // ---- resident operands ----
    std::mt19937_64 rng(master);
    std::uniform_real_distribution<double> ulog(std::log(256.0), std::log(8192.0));
    std::uniform_real_distribution<double> us(0.95, 0.999);
    const int Ns[4] = {32, 64, 128, 256};
    std::vector<Job> jobs(nJobs);
    for (int i = 0; i < nJobs; ++i) {
        Job &j = jobs[i];
        j.M = (int)std::lround(std::exp(ulog(rng)));
        j.K = (int)std::lround(std::exp(ulog(rng)));
        j.N = Ns[rng() % 4];
        j.s = us(rng);
        j.seed = master ^ (uint64_t)i;
    }
    const bool doF32 = dtypes != "bf16", doBF16 = dtypes != "f32";
    gcoo_server_config cfgB{budget, slots, maxBatch};
    gcoo_server_config cfgP{budget, 1, 1};
    gcoo_server_t batched = nullptr, plain = nullptr;
    CK(gcoo_server_create(&cfgB, &batched));
    CK(gcoo_server_create(&cfgP, &plain));
    std::vector<int> opsF32B, opsF32P, opsBF16B, opsBF16P;
    std::vector<double> gflop;
    float *Sbuf = nullptr;
    size_t SbufBytes = 0;
    for (Job &j : jobs) {
        const size_t sb = (size_t)j.M * j.K * 4;
        if (sb > SbufBytes) {
            if (Sbuf) CK(gcoo_dev_free(Sbuf));
            CK(gcoo_dev_malloc((void **)&Sbuf, sb));
            SbufBytes = sb;
        }
        CK(gcoo_fill_sparse_async(Sbuf, (int64_t)j.M * j.K, j.seed, 1, 1.0f,
                                  gcoo_host::keep_threshold(j.s), nullptr));
        CK(gcoo_create(j.M, j.K, 4, &j.h));
        CK(gcoo_convert_dense(j.h, Sbuf, nullptr));
        int64_t nnz = 0;
        CK(gcoo_get_info(j.h, nullptr, nullptr, nullptr, nullptr, &nnz));
        j.nnz = nnz;
        CK(gcoo_dev_malloc((void **)&j.D, (size_t)j.K * j.N * 4));
        CK(gcoo_fill_uniform_async(j.D, (int64_t)j.K * j.N, j.seed, 0, 1.0f, nullptr));
        CK(gcoo_dev_malloc((void **)&j.D16, (size_t)j.K * j.N * 2));
        CK(gcoo_f32_to_bf16_async(j.D, j.D16, (int64_t)j.K * j.N, nullptr));
        CK(gcoo_device_sync());
        const double per = j.K * (1.0 - j.s);
        int rows, vec;
        tune.tile(false, j.N, per, &rows, &vec);
        int op;
        CK(gcoo_server_register(batched, j.h, j.D, 0, j.N, rows, vec, &op));
        opsF32B.push_back(op);
        CK(gcoo_server_register(plain, j.h, j.D, 0, j.N, rows, vec, &op));
        opsF32P.push_back(op);
        tune.tile(true, j.N, per, &rows, &vec);
        CK(gcoo_server_register(batched, j.h, j.D16, 1, j.N, rows, vec, &op));
        opsBF16B.push_back(op);
        CK(gcoo_server_register(plain, j.h, j.D16, 1, j.N, rows, vec, &op));
        opsBF16P.push_back(op);
        gflop.push_back(2.0 * (double)j.nnz * j.N / 1e9);
    }
    if (Sbuf) CK(gcoo_dev_free(Sbuf));
    // op ids are (job * 2 + dtype) in both servers; gflop per op id
    std::vector<double> gflopOp(2 * nJobs);
    for (int i = 0; i < nJobs; ++i) gflopOp[2 * i] = gflopOp[2 * i + 1] = gflop[i];

    // ---- accuracy of a sample of results (through the batched server) ----
    auto check = [&](bool bf16) {
        double worst = 0, sum = 0;
        int n = 0;
        for (int c = 0; c < nCheck && c < nJobs; ++c) {
            const int i = (int)((long long)c * nJobs / std::max(1, nCheck));
            Job &j = jobs[i];
            uint64_t id;
            CK(gcoo_server_submit(batched, bf16 ? opsBF16B[i] : opsF32B[i], &id));
            CK(gcoo_server_wait(batched, id, nullptr));
            const float *dC;
            int M, N;
            CK(gcoo_server_result(batched, id, &dC, &M, &N));
            std::vector<float> C((size_t)M * N), S((size_t)j.M * j.K), D((size_t)j.K * j.N);
            CK(gcoo_memcpy_d2h(C.data(), dC, C.size() * 4, nullptr));
            CK(gcoo_server_release(batched, id));
            gcoo_host::fill_counter(S.data(), S.size(), j.seed, 1, 1.0f,
                                    gcoo_host::keep_threshold(j.s), true);
            gcoo_host::fill_counter(D.data(), D.size(), j.seed, 0, 1.0f, 0, false);
            const double e = gcoo_host::rel_l2_vs_fp64(S.data(), D.data(), C.data(), j.M, j.K,
                                                       j.N, 8);
            worst = std::max(worst, e);
            sum += e;
            ++n;
        }
        std::printf("{\"kind\": \"accuracy\", \"dtype\": \"%s\", \"sampled_jobs\": %d, "
                    "\"max_rel_frobenius\": %.3e, \"mean_rel_frobenius\": %.3e}\n",
                    bf16 ? "bf16" : "f32", n, worst, n ? sum / n : 0.0);
        std::fflush(stdout);
    };

    std::mt19937_64 arr(master + 7);
    for (int pass = 0; pass < 2; ++pass) {
        const bool bf16 = pass == 1;
        if ((bf16 && !doBF16) || (!bf16 && !doF32)) continue;
        const std::vector<int> &opsB = bf16 ? opsBF16B : opsF32B;
        const std::vector<int> &opsP = bf16 ? opsBF16P : opsF32P;
        check(bf16);
        // 1. saturated capacity of each mode: one request per job, all at once
        double cap[2];
        for (int mode = 0; mode < 2; ++mode) {
            gcoo_server_t srv = mode == 0 ? batched : plain;
            const std::vector<int> &ops = mode == 0 ? opsB : opsP;
            std::vector<int> all(ops.begin(), ops.end());
            std::vector<double> zero(all.size(), 0.0);
            drive(srv, all, gflopOp, zero);  // warm-up
            // the median of five drives (one drive takes a few ms, so one
            // slow batch would move a single reading)
            std::vector<RunResult> sats;
            for (int t = 0; t < 5; ++t) sats.push_back(drive(srv, all, gflopOp, zero));
            std::sort(sats.begin(), sats.end(), [](const RunResult &a, const RunResult &b) {
                return a.st.wall_s < b.st.wall_s;
            });
            const RunResult sat = sats[2];
            cap[mode] = sat.st.completed / sat.st.wall_s;
            const double fixedRate = bf16 ? rateBF16 : rateF32;
            const double measured = cap[mode];
            if (mode == 0 && fixedRate > 0) cap[mode] = fixedRate;
            std::printf("{\"kind\": \"saturated\", \"dtype\": \"%s\", \"mode\": \"%s\", \"jobs\": %llu, "
                        "\"wall_ms\": %.3f, \"jobs_per_s\": %.1f, \"gflops\": %.1f, \"batches\": %llu, "
                        "\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"peak_bytes\": %zu, \"budget_bytes\": %zu}\n",
                        bf16 ? "bf16" : "f32", mode == 0 ? "batched" : "plain",
                        (unsigned long long)sat.st.completed, sat.st.wall_s * 1e3, measured,
                        sat.gflop / sat.st.wall_s, (unsigned long long)sat.st.batches,
                        sat.st.p50_ms, sat.st.p99_ms, sat.st.peak_bytes, sat.st.budget_bytes);
            std::fflush(stdout);
        }
        // 2. Poisson arrivals, each mode at the given fraction of its own
        // capacity (the same request sequence, different rates)
        for (double load : loads) {
            std::uniform_int_distribution<int> pick(0, nJobs - 1);
            std::exponential_distribution<double> unit(1.0);
            std::vector<int> idx(nReq);
            std::vector<double> unitGaps(nReq);
            for (int r = 0; r < nReq; ++r) {
                idx[r] = pick(arr);
                unitGaps[r] = unit(arr);
            }
            for (int mode = 0; mode < 2; ++mode) {
                gcoo_server_t srv = mode == 0 ? batched : plain;
                const std::vector<int> &ops = mode == 0 ? opsB : opsP;
                const double lambda = load * cap[mode];
                std::vector<double> gaps(nReq);
                for (int r = 0; r < nReq; ++r) gaps[r] = unitGaps[r] / lambda;
                std::vector<int> seq(nReq);
                for (int r = 0; r < nReq; ++r) seq[r] = ops[idx[r]];
                // one saturated drive right before, so the run starts at the
                // clock a busy GPU runs at
                {
                    std::vector<int> all(ops.begin(), ops.end());
                    drive(srv, all, gflopOp, std::vector<double>(all.size(), 0.0));
                }
                const double t0 = now_s();
                RunResult rr = drive(srv, seq, gflopOp, gaps);
                const double wall = now_s() - t0;
                std::printf("{\"kind\": \"poisson\", \"dtype\": \"%s\", \"mode\": \"%s\", "
                            "\"load\": %.2f, \"arrival_rate\": %.1f, \"requests\": %llu, "
                            "\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"mean_ms\": %.3f, \"max_ms\": %.3f, "
                            "\"jobs_per_s\": %.1f, \"gflops\": %.1f, \"batches\": %llu, "
                            "\"peak_bytes\": %zu, \"budget_bytes\": %zu, \"wall_s\": %.3f}\n",
                            bf16 ? "bf16" : "f32", mode == 0 ? "batched" : "plain", load, lambda,
                            (unsigned long long)rr.st.completed, rr.st.p50_ms, rr.st.p99_ms,
                            rr.st.mean_ms, rr.st.max_ms, rr.st.completed / rr.st.wall_s,
                            rr.gflop / rr.st.wall_s, (unsigned long long)rr.st.batches,
                            rr.st.peak_bytes, rr.st.budget_bytes, wall);
                std::fflush(stdout);
            }
        }
    }
    CK(gcoo_server_destroy(batched));
    CK(gcoo_server_destroy(plain));
    for (Job &j : jobs) {
        gcoo_destroy(j.h);
        gcoo_dev_free(j.D);
        gcoo_dev_free(j.D16);
    }
    return 0;
}
