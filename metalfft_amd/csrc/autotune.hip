// Autotuner: picks the kernel variant per (size, dtype) by timing every
// compiled candidate on the device, and caches the choice on disk.
//
// The reference has a fixed size -> kernel table (FFTDescriptor,
// Sources/MetalFFT/Internal/MetalContext.swift:48-67, plus the radix-8 switch
// for batched N=4096 in Sources/MetalFFT/FFT/MetalFFT.swift:163-165); this
// replaces it with measurement.
//
// Choices are kept in memory for the life of the process.  A disk cache is
// opt-in: only when $MIFFT_AUTOTUNE_CACHE names a file is it read and
// written (text, first line "mifft-autotune 8 <arch>", then one
// "<kind> <precision> <io> <n> <variant-name> <ns-per-transform>" line per
// entry, e.g. "stockham fp32 c64 4096 ..."; a
// header mismatch discards it); without it the library writes nothing.
// MIFFT_AUTOTUNE=0 skips timing and uses the built-in defaults;
// MIFFT_AUTOTUNE=force re-times and rewrites entries; MIFFT_AUTOTUNE_LOG=1
// prints every candidate's time to stderr.
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "fourstep.hpp"
#include "mfma_fft.hpp"
#include "runtime.hpp"

namespace mifft {

namespace {
const char* kHeader = "mifft-autotune 8 gfx950";

std::mutex g_tune_mutex;
bool g_loaded = false;
std::map<std::string, std::string> g_table;  // key -> variant name

std::string cache_path() {
    const char* p = std::getenv("MIFFT_AUTOTUNE_CACHE");
    return p ? p : "";
}

std::string mode() {
    const char* m = std::getenv("MIFFT_AUTOTUNE");
    return m ? m : "1";
}

void load_locked() {
    if (g_loaded) return;
    g_loaded = true;
    const std::string path = cache_path();
    if (path.empty()) return;
    std::ifstream f(path);
    if (!f) return;
    std::string line;
    if (!std::getline(f, line) || line != kHeader) return;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::string kind, prec, io, n, name;
        if (ss >> kind >> prec >> io >> n >> name) g_table[kind + " " + prec + " " + io + " " + n] = name;
    }
}

void save_locked(const std::string& key, const std::string& name, double ns) {
    std::string path = cache_path();
    if (path.empty()) return;
    std::map<std::string, std::string> lines;
    {
        std::ifstream f(path);
        std::string line;
        if (f && std::getline(f, line) && line == kHeader) {
            while (std::getline(f, line)) {
                std::istringstream ss(line);
                std::string kind, prec, io, n;
                if (ss >> kind >> prec >> io >> n) lines[kind + " " + prec + " " + io + " " + n] = line;
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

// Timing input: uniform values in [-1, 1) from an integer hash (data with
// real bit activity, as transforms see; zeros would flatter every kernel).
__device__ __forceinline__ float2 noise2(int64_t i) {
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    return make_float2((float)(h & 0xffff) / 32768.f - 1.f, (float)(h >> 16) / 32768.f - 1.f);
}
__global__ void k_fill_noise(float2* __restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = noise2(i);
}
typedef _Float16 h2t __attribute__((ext_vector_type(2)));
__global__ void k_fill_noise_h(h2t* __restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 v = noise2(i);
    h2t o = {(_Float16)v.x, (_Float16)v.y};
    a[i] = o;
}

bool tune_log() {
    const char* lg = std::getenv("MIFFT_AUTOTUNE_LOG");
    return lg && *lg && *lg != '0';
}

// Times the candidates of length n (elements of elem_bytes in HBM) and
// returns the index of the one to use (-1 if none ran).
// Timed as launches of 2^24 / n transforms (128 MiB in, 128 MiB out for
// complex64) rotating over four disjoint input/output sets (1 GiB in all,
// four times the 256 MiB Infinity Cache), as a pipeline over fresh batches
// runs: with only two sets, kernels whose loads allocate in that cache were
// timed up to 15 % faster than they run on fresh data.  Candidates are timed
// round-robin (five rounds of eight launches each, after eight warm-up
// launches) and each keeps its best round, so a drift of the clocks during
// tuning does not favour whichever ran first.  A preferred (stateless,
// one-shot) candidate within 3 % of the fastest is kept instead: persistent
// and DMA-staged forms trade a per-launch prologue for overlap, so their
// time per transform moves more with batch size and launch mode.
int time_candidates(int n, size_t elem_bytes, const std::vector<Candidate>& cand, hipStream_t stream,
                    const char* what, double* best_ns) {
    constexpr int kSets = 4, kRounds = 5, kReps = 8;
    const int64_t batch = std::max<int64_t>(1, (int64_t(1) << 24) / n);  // per launch
    const int64_t set = (int64_t)n * batch, total = set * kSets;
    char *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, elem_bytes * total) != hipSuccess || hipMalloc(&b, elem_bytes * total) != hipSuccess) {
        (void)hipGetLastError();
        if (a) (void)hipFree(a);
        return -1;
    }
    const unsigned grid = (unsigned)((total + 255) / 256);
    if (elem_bytes == 4)
        hipLaunchKernelGGL(k_fill_noise_h, dim3(grid), dim3(256), 0, stream, reinterpret_cast<h2t*>(a), total);
    else
        hipLaunchKernelGGL(k_fill_noise, dim3(grid), dim3(256), 0, stream, reinterpret_cast<float2*>(a), total);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    auto run = [&](const Candidate& c, int count) {
        bool good = true;
        for (int r = 0; r < count && good; ++r)
            good = c.run(a + (r % kSets) * set * elem_bytes, b + (r % kSets) * set * elem_bytes, batch) ==
                   hipSuccess;
        return good;
    };
    std::vector<int> ok;
    for (size_t q = 0; q < cand.size(); ++q)
        if (run(cand[q], kReps)) ok.push_back((int)q);
    (void)hipGetLastError();
    std::vector<double> t(cand.size(), 1e30);
    for (int round = 0; round < kRounds; ++round) {
        for (int q : ok) {
            (void)hipEventRecord(e0, stream);
            run(cand[q], kReps);
            (void)hipEventRecord(e1, stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            t[q] = std::min(t[q], 1e6 * ms / kReps / batch);
        }
    }
    if (tune_log())
        for (int q : ok) std::fprintf(stderr, "mifft autotune %s n=%d %s %.2f ns/transform\n", what, n, cand[q].name, t[q]);
    int best = -1, pref = -1;
    for (int q : ok) {
        if (best < 0 || t[q] < t[best]) best = q;
        if (cand[q].prefer && (pref < 0 || t[q] < t[pref])) pref = q;
    }
    if (pref >= 0 && t[pref] <= 1.03 * t[best]) best = pref;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipGetLastError();
    if (best >= 0) *best_ns = t[best];
    return best;
}

}  // namespace

// The table entry for `key`, or the timed choice among `cand` (recorded).
// Returns the candidate index, or -1 (caller falls back to its default).
int tuned_choice(const std::string& key, int n, size_t elem_bytes, const std::vector<Candidate>& cand,
                 hipStream_t stream, const char* what) {
    const std::string m = mode();
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    load_locked();
    if (m != "force") {
        auto it = g_table.find(key);
        if (it != g_table.end())
            for (size_t q = 0; q < cand.size(); ++q)
                if (it->second == cand[q].name) return (int)q;
    }
    double ns = 0;
    const int q = time_candidates(n, elem_bytes, cand, stream, what, &ns);
    if (q < 0) return -1;
    g_table[key] = cand[q].name;
    save_locked(key, cand[q].name, ns);
    return q;
}

bool tuning_enabled() { return mode() != "0"; }

const StockhamVariant* choose_variant(int n, hipStream_t stream, const float2* tw) {
    if (mode() == "0") return default_variant(n);
    int c = 0;
    const StockhamVariant* all = stockham_variants(&c);
    std::vector<const StockhamVariant*> vs;
    std::vector<Candidate> cand;
    for (int i = 0; i < c; ++i) {
        if (all[i].n != n) continue;
        const StockhamVariant* v = &all[i];
        vs.push_back(v);
        cand.push_back({v->name,
                        [=](const void* in, void* out, int64_t batch) {
                            return v->launch(static_cast<const float2*>(in), static_cast<float2*>(out), tw,
                                             batch, false, stream);
                        },
                        std::strstr(v->name, "_swz_nt_pt") || std::strstr(v->name, "_half")});
    }
    const int q = tuned_choice("stockham fp32 c64 " + std::to_string(n), n, 8, cand, stream, "stockham");
    return q < 0 ? default_variant(n) : vs[q];
}

int choose_mfma(int precision, int io, hipStream_t stream, const float2* tw) {
    static const char* prec_names[5] = {"fp32", "fp16", "bf16", "fp16split", "bf16split"};
    if (mode() == "0" || precision < 1 || precision > 4) return kMfmaStaged;
    std::vector<Candidate> cand;
    for (int form = 0; form < kMfmaForms; ++form)
        cand.push_back({mfma_form_name(precision, io, form),
                        [=](const void* in, void* out, int64_t batch) {
                            return launch_fft_mfma(4096, precision, io, form, in, out, tw, batch, false, 1.0f,
                                                   stream);
                        },
                        form == kMfmaOneShot || form == kMfmaOneShot64});
    const std::string key = std::string("mfma ") + prec_names[precision] + (io ? " c32h" : " c64") + " 4096";
    const int q = tuned_choice(key, 4096, io ? 4 : 8, cand, stream, "mfma");
    return q < 0 ? kMfmaStaged : q;
}

void choose_long(int n, hipStream_t stream, const float2* tw, const float2* fs_tw,
                 const StockhamVariant** variant, const FourStepVariant** fourstep) {
    *variant = nullptr;
    *fourstep = fs_tw ? find_fourstep(n) : nullptr;
    if (!*fourstep && tw) *variant = default_variant(n);
    if (mode() == "0") return;
    std::vector<Candidate> cand;
    std::vector<const StockhamVariant*> vs;
    std::vector<const FourStepVariant*> fs;
    int c = 0;
    if (tw) {
        const StockhamVariant* all = stockham_variants(&c);
        for (int i = 0; i < c; ++i) {
            if (all[i].n != n) continue;
            const StockhamVariant* v = &all[i];
            cand.push_back({v->name,
                            [=](const void* in, void* out, int64_t batch) {
                                return v->launch(static_cast<const float2*>(in), static_cast<float2*>(out),
                                                 tw, batch, false, stream);
                            },
                            true});
            vs.push_back(v);
            fs.push_back(nullptr);
        }
    }
    if (fs_tw) {
        const FourStepVariant* all = fourstep_variants(&c);
        for (int i = 0; i < c; ++i) {
            if (all[i].n != n) continue;
            const FourStepVariant* v = &all[i];
            // out of place: the output buffer holds the intermediate
            cand.push_back({v->name,
                            [=](const void* in, void* out, int64_t batch) {
                                return v->launch(static_cast<const float2*>(in), static_cast<float2*>(out),
                                                 static_cast<float2*>(out), fs_tw, batch, false, stream);
                            },
                            true});
            vs.push_back(nullptr);
            fs.push_back(v);
        }
    }
    const int q = tuned_choice("long fp32 c64 " + std::to_string(n), n, 8, cand, stream, "long");
    if (q < 0) return;
    *variant = vs[q];
    *fourstep = fs[q];
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
    const mifft::FourStepVariant* f = mifft::fourstep_variants(&c);
    for (int i = 0; i < c; ++i) {
        if (f[i].n != n) continue;
        if (names && k < cap) names[k] = f[i].name;
        ++k;
    }
    return k;
}

mifft_status mifft_plan_select_kernel(mifft_plan* plan, const char* name) {
    if (!plan || !name) return mifft::fail(MIFFT_ERR_INVALID_ARGUMENT, "null argument");
    if (plan->precision != MIFFT_PREC_FP32) {
        for (int form = 0; form < mifft::kMfmaForms; ++form)
            if (std::strcmp(mifft::mfma_form_name(plan->precision, plan->io, form), name) == 0) {
                plan->mfma_form = form;
                return MIFFT_OK;
            }
    }
    if (plan->precision == MIFFT_PREC_FP32) {
        if (plan->fs_twiddles) {
            int c = 0;
            const mifft::FourStepVariant* f = mifft::fourstep_variants(&c);
            for (int i = 0; i < c; ++i)
                if (f[i].n == plan->n && std::strcmp(f[i].name, name) == 0) {
                    plan->fourstep = &f[i];
                    plan->variant = nullptr;
                    return MIFFT_OK;
                }
        }
        if (plan->twiddles) {
            const mifft::StockhamVariant* v = mifft::by_name(name);
            if (v && v->n == plan->n) {
                plan->variant = v;
                plan->fourstep = nullptr;
                return MIFFT_OK;
            }
        }
    }
    return mifft::fail(MIFFT_ERR_KERNEL_NOT_FOUND, std::string("HIP kernel not found: ") + name);
}

}  // extern "C"
