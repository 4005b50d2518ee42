// Autotuner: picks the kernel variant per (size, dtype) by timing every
// compiled candidate on the device, and caches the choice on disk.
//
// The reference has a fixed size -> kernel table (FFTDescriptor,
// Sources/MetalFFT/Internal/MetalContext.swift:48-67, plus the radix-8 switch
// for batched N=4096 in Sources/MetalFFT/FFT/MetalFFT.swift:163-165); this
// replaces it with measurement.
//
// Cache: text file, first line "mifft-autotune 6 <arch>", then one
// "<kind> <dtype> <n> <variant-name> <ns-per-transform>" line per entry.
// Location: $MIFFT_AUTOTUNE_CACHE, else $XDG_CACHE_HOME/mifft/autotune.txt,
// else $HOME/.cache/mifft/autotune.txt.  A header mismatch (other version or
// arch) discards the file.  MIFFT_AUTOTUNE=0 skips timing and uses the
// built-in defaults; MIFFT_AUTOTUNE=force re-times and rewrites entries;
// MIFFT_AUTOTUNE_LOG=1 prints every candidate's time to stderr.
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace mifft {

namespace {
const char* kHeader = "mifft-autotune 6 gfx950";

std::mutex g_tune_mutex;
bool g_loaded = false;
std::map<std::string, std::string> g_table;  // key -> variant name

std::string cache_path() {
    if (const char* p = std::getenv("MIFFT_AUTOTUNE_CACHE")) return p;
    std::string base;
    if (const char* x = std::getenv("XDG_CACHE_HOME")) base = x;
    else if (const char* h = std::getenv("HOME")) base = std::string(h) + "/.cache";
    else return "";
    return base + "/mifft/autotune.txt";
}

std::string mode() {
    const char* m = std::getenv("MIFFT_AUTOTUNE");
    return m ? m : "1";
}

void load_locked() {
    if (g_loaded) return;
    g_loaded = true;
    std::ifstream f(cache_path());
    if (!f) return;
    std::string line;
    if (!std::getline(f, line) || line != kHeader) return;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::string kind, dtype, n, name;
        if (ss >> kind >> dtype >> n >> name) g_table[kind + " " + dtype + " " + n] = name;
    }
}

void save_locked(const std::string& key, const std::string& name, double ns) {
    std::string path = cache_path();
    if (path.empty()) return;
    std::string dir = path.substr(0, path.find_last_of('/'));
    // mkdir -p of at most two levels ($HOME/.cache/mifft)
    std::string parent = dir.substr(0, dir.find_last_of('/'));
    ::mkdir(parent.c_str(), 0755);
    ::mkdir(dir.c_str(), 0755);
    std::map<std::string, std::string> lines;
    {
        std::ifstream f(path);
        std::string line;
        if (f && std::getline(f, line) && line == kHeader) {
            while (std::getline(f, line)) {
                std::istringstream ss(line);
                std::string kind, dtype, n;
                if (ss >> kind >> dtype >> n) lines[kind + " " + dtype + " " + n] = line;
            }
        }
    }
    char buf[64];
    std::snprintf(buf, sizeof buf, " %.1f", ns);
    lines[key] = key + " " + name + buf;
    std::string tmp = path + ".tmp" + std::to_string(::getpid());
    {
        std::ofstream o(tmp);
        if (!o) return;
        o << kHeader << "\n";
        for (auto& kv : lines) o << kv.second << "\n";
    }
    std::rename(tmp.c_str(), path.c_str());
}

// Built-in choice when not tuning: radix-16 passes everywhere.
const StockhamVariant* default_variant(int n) {
    if (const StockhamVariant* v = find_stockham(n, 16)) return v;
    return find_stockham(n, 8);
}

const StockhamVariant* by_name(const std::string& name) {
    int c = 0;
    const StockhamVariant* v = stockham_variants(&c);
    for (int i = 0; i < c; ++i)
        if (name == v[i].name) return &v[i];
    return nullptr;
}

// Times each candidate on ~32 MiB of data; returns the fastest.
// Timing input: uniform values in [-1, 1) from an integer hash (data with
// real bit activity, as transforms see; zeros would flatter every kernel).
__global__ void k_fill_noise(float2* __restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    a[i] = make_float2((float)(h & 0xffff) / 32768.f - 1.f, (float)(h >> 16) / 32768.f - 1.f);
}

// Timed as launches of 2^24 / n transforms (128 MiB in, 128 MiB out)
// rotating over four disjoint input/output sets (1 GiB in all, four times
// the 256 MiB Infinity Cache), as a pipeline over fresh batches runs: with
// only two sets, kernels whose loads allocate in that cache were timed up to
// 15 % faster than they run on fresh data.  Candidates are timed round-robin
// (five rounds of eight launches each, after eight warm-up launches) and
// each keeps its best round, so a drift of the clocks during tuning does not
// favour whichever ran first.  A stateless one-shot variant ("_swz_nt_pt",
// "_half") within 3 % of the fastest is kept instead: persistent and DMA-staged
// variants trade a per-launch prologue for overlap, so their time per
// transform moves more with batch size and launch mode.
const StockhamVariant* time_variants(int n, hipStream_t stream, const float2* tw, double* best_ns) {
    constexpr int kSets = 4, kRounds = 5, kReps = 8;
    const int64_t batch = std::max<int64_t>(1, (int64_t(1) << 24) / n);  // per launch
    float2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, sizeof(float2) * n * kSets * batch) != hipSuccess ||
        hipMalloc(&b, sizeof(float2) * n * kSets * batch) != hipSuccess) {
        (void)hipGetLastError();
        if (a) (void)hipFree(a);
        return nullptr;
    }
    const int64_t total = (int64_t)n * kSets * batch;
    hipLaunchKernelGGL(k_fill_noise, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, a,
                       total);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int c = 0;
    const StockhamVariant* all = stockham_variants(&c);
    const int64_t set = (int64_t)n * batch;
    auto run = [&](const StockhamVariant* v, int count) {
        bool good = true;
        for (int r = 0; r < count && good; ++r)
            good = v->launch(a + (r % kSets) * set, b + (r % kSets) * set, tw, batch, false, stream) ==
                   hipSuccess;
        return good;
    };
    std::vector<const StockhamVariant*> cand;
    for (int i = 0; i < c; ++i)
        if (all[i].n == n && run(&all[i], kReps)) cand.push_back(&all[i]);
    std::vector<double> t(cand.size(), 1e30);
    for (int round = 0; round < kRounds; ++round) {
        for (size_t q = 0; q < cand.size(); ++q) {
            (void)hipEventRecord(e0, stream);
            run(cand[q], kReps);
            (void)hipEventRecord(e1, stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            t[q] = std::min(t[q], 1e6 * ms / kReps / batch);
        }
    }
    if (const char* lg = std::getenv("MIFFT_AUTOTUNE_LOG"); lg && *lg && *lg != '0')
        for (size_t q = 0; q < cand.size(); ++q)
            std::fprintf(stderr, "mifft autotune n=%d %s %.2f ns/transform\n", n, cand[q]->name, t[q]);
    const StockhamVariant* best = nullptr;
    const StockhamVariant* stateless = nullptr;
    double best_t = 1e30, stateless_t = 1e30;
    for (size_t q = 0; q < cand.size(); ++q) {
        if (t[q] < best_t) {
            best_t = t[q];
            best = cand[q];
        }
        if ((std::strstr(cand[q]->name, "_swz_nt_pt") || std::strstr(cand[q]->name, "_half")) &&
            t[q] < stateless_t) {
            stateless_t = t[q];
            stateless = cand[q];
        }
    }
    if (stateless && stateless_t <= 1.03 * best_t) {
        best = stateless;
        best_t = stateless_t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipGetLastError();
    *best_ns = best_t;
    return best;
}
}  // namespace

const StockhamVariant* choose_variant(int n, hipStream_t stream, const float2* tw) {
    const std::string m = mode();
    if (m == "0") return default_variant(n);
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    load_locked();
    const std::string key = "stockham c64 " + std::to_string(n);
    if (m != "force") {
        auto it = g_table.find(key);
        if (it != g_table.end())
            if (const StockhamVariant* v = by_name(it->second))
                if (v->n == n) return v;
    }
    double ns = 0;
    const StockhamVariant* v = time_variants(n, stream, tw, &ns);
    if (!v) return default_variant(n);
    g_table[key] = v->name;
    save_locked(key, v->name, ns);
    return v;
}

}  // namespace mifft

extern "C" {

// Lists compiled kernel variants: writes up to `cap` names of variants for
// length n into `names` and returns the total count.
int mifft_kernel_variants(int64_t n, const char** names, int cap) {
    int c = 0, k = 0;
    const mifft::StockhamVariant* v = mifft::stockham_variants(&c);
    for (int i = 0; i < c; ++i) {
        if (v[i].n != n) continue;
        if (names && k < cap) names[k] = v[i].name;
        ++k;
    }
    return k;
}

mifft_status mifft_plan_select_kernel(mifft_plan* plan, const char* name) {
    if (!plan || !name) return mifft::fail(MIFFT_ERR_INVALID_ARGUMENT, "null argument");
    const mifft::StockhamVariant* v = mifft::by_name(name);
    if (!v || v->n != plan->n)
        return mifft::fail(MIFFT_ERR_KERNEL_NOT_FOUND, std::string("HIP kernel not found: ") + name);
    plan->variant = v;
    return MIFFT_OK;
}

}  // extern "C"
