// LDS-resident Stockham autosort FFT templates for N = 64 ... 16384 (gfx950).
//
// Each transform is worked by T = N/E threads that keep E complex values in
// registers:
//
//   pass 0   : v[m] = x[t + T*m] (coalesced global load) -> DFT-R     -> LDS
//   pass p   : LDS read  v[m] = a[t + T*m] -> twiddle -> DFT-R        -> LDS
//   last     : LDS read                     -> twiddle -> DFT-R, leaving
//              v[m] = X[t + T*m] in registers (coalesced global store)
//
// Butterfly b of a radix-R pass at stride s reads x_j = a[b + j*N/R],
// multiplies by W_N^{j*pos*N/(R*s)} (pos = b mod s), runs a DFT-R and writes
// y_j to a'[(b/s)*R*s + pos + j*s]: the self-sorting Stockham DIT pass of the
// reference's radix-4 and radix-8 kernels
// (Sources/MetalFFT/Resources/fft_multisize.metal:92-111,
//  Sources/MetalFFT/Resources/fft_4096_batched.metal:108-203), here for
// E in {4, 8, 16}, with twiddles read from one FP32 table generated in FP64
// instead of per-butterfly sincos.
//
// The inverse follows MetalFFT.inverse (Sources/MetalFFT/FFT/MetalFFT.swift:56-65):
// conjugate on load, forward transform, conjugate and multiply by 1/N on store.
#pragma once
#include "fft_device.hpp"

namespace mifft {

// ------------------------------------------------------------- LDS layouts
// Padded: one spare complex per 16 (N + N/16 entries).
struct LdsPad {
    static constexpr bool kXor = false;
    static constexpr int len(int n) { return n + n / 16; }
    static __device__ __forceinline__ int idx(int i) { return i + (i >> 4); }
};
// XOR swizzle (exactly N entries): i ^ ((i >> 4) & 31) flips bits 0-4 by
// bits 4-8, a bijection on [0, N) for N >= 32.  Radix-16 scatter writes
// (lane stride 16 or 256 complex) and contiguous reads then hit 32 distinct
// 8-byte slots per 32 lanes.
struct LdsSwz {
    static constexpr bool kXor = true;
    static constexpr int len(int n) { return n; }
    static __device__ __forceinline__ int idx(int i) { return i ^ ((i >> 4) & 31); }
};

// Reads v[m] = buf[L::idx(t + T*m)], m < E.  For the swizzle and T >= 32 a
// power of two with t < T: (t + T*m) >> 4 = (t >> 4) + (T/16)*m without
// carries between the terms, so idx(t + T*m) = (idx(t) ^ c(m)) + T*m with
// c(m) = ((T/16)*m) & 31, which repeats with period 512/T (T < 512) or is 0:
// one or a few base addresses and the T*m terms as immediate offsets,
// instead of the swizzle arithmetic per element.
template <int E, int T, class L>
__device__ __forceinline__ void lds_read_column(float2* v, const float2* buf, int t) {
    if constexpr (sizeof(L) > 0 && L::kXor && T >= 32 && (T & (T - 1)) == 0) {
        constexpr int P = T >= 512 ? 1 : 512 / T;
        const int i0 = L::idx(t);
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = buf[(i0 ^ (((T / 16) * (m % P)) & 31)) + T * m];
    } else if constexpr (!L::kXor && T % 16 == 0) {
        // padded layout: idx(t + T*m) = idx(t) + (T + T/16)*m
        const int i0 = L::idx(t);
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = buf[i0 + (T + T / 16) * m];
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = buf[L::idx(t + T * m)];
    }
}

// Writes the R outputs of a butterfly to buf[L::idx(base + j*SD)], j < R,
// base = (a multiple of R*SD) + pos, pos < SD.  For the swizzle,
// idx(base + j*SD) = idx(base) ^ idx(j*SD): the bits j*SD sets are zero in
// base, and the swizzle's source bits (4-8) of base + j*SD are those of base
// plus those of j*SD without carries, so each output costs one XOR with a
// constant instead of the swizzle arithmetic.
template <int R, int SD, int STRIDE, class L>
__device__ __forceinline__ void lds_write_butterfly(float2* buf, int base, const float2* y) {
    if constexpr (L::kXor) {
        const int i0 = L::idx(base);
#pragma unroll
        for (int j = 0; j < R; ++j) buf[i0 ^ L::idx(j * SD)] = y[STRIDE * j];
    } else if constexpr (!L::kXor && SD % 16 == 0) {
        // padded layout: idx(base + j*SD) = idx(base) + j*(SD + SD/16)
        const int i0 = L::idx(base);
#pragma unroll
        for (int j = 0; j < R; ++j) buf[i0 + j * (SD + SD / 16)] = y[STRIDE * j];
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) buf[L::idx(base + j * SD)] = y[STRIDE * j];
    }
}

// Workgroup barrier ordering LDS only: global loads issued before it stay in
// flight (a full __syncthreads also waits for them).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Barrier that orders this workgroup's DS (LDS) instructions only: each wave
// drains its own LDS queue (lgkmcnt) and meets the others at s_barrier; the
// asm memory clobber keeps the compiler from moving memory accesses across.
// Unlike the fence form it does not wait for LDS-DMA transfers in flight
// (they count in vmcnt), which is the point in the DMA streaming kernel,
// where each DMA target is waited for explicitly before it is read.
__device__ __forceinline__ void ds_barrier() {
    __asm__ __volatile__("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// An opaque copy of an index (empty asm): arithmetic derived from it in one
// phase or loop iteration (LDS and global addresses) is recomputed there,
// instead of being shared with another phase or hoisted out of a loop and
// held live in registers through everything in between.
__device__ __forceinline__ int fresh(int i) {
    __asm__("" : "+v"(i));
    return i;
}

// ------------------------------------------------------- global I/O helpers
typedef float f2v __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ float2 gload(const float2* p) {
    if constexpr (NT) {
        f2v x = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
        return make_float2(x.x, x.y);
    } else {
        return *p;
    }
}

template <bool NT>
__device__ __forceinline__ void gstore(float2* p, float2 v) {
    if constexpr (NT) {
        f2v x = {v.x, v.y};
        __builtin_nontemporal_store(x, reinterpret_cast<f2v*>(p));
    } else {
        *p = v;
    }
}

// Cache policies of global stores and loads, P: 0 non-temporal (streaming),
// 1 the default policy, 2 system-coherent stores (sc0 sc1; loads as 1).
template <int P>
__device__ __forceinline__ float2 gload_p(const float2* p) {
    return gload<P == 0>(p);
}

template <int P>
__device__ __forceinline__ void gstore_p(float2* p, float2 v) {
    if constexpr (P == 2) {
        const unsigned long long u =
            ((unsigned long long)__float_as_uint(v.y) << 32) | __float_as_uint(v.x);
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    } else {
        gstore<P == 0>(p, v);
    }
}

// ------------------------------------------------------------ twiddles
// Number of table entries a radix-R butterfly reads: W^k, W^2k, W^4k, W^8k
// for R = 16 (log2 R of them).
constexpr int twiddle_bases(int r) { return r >= 16 ? 4 : r >= 8 ? 3 : r >= 4 ? 2 : r >= 2 ? 1 : 0; }

// w[j] = W^{j*k} for j < R from b[q] = W^{(2^q)*k}; every other power is a
// product of at most three of them, so it carries at most three FP32
// roundings beyond the table's own (~1e-7 relative).
template <int R>
__device__ __forceinline__ void powers_from_bases(float2 (&w)[R], const float2* b) {
    w[0] = make_float2(1.f, 0.f);
    if constexpr (R >= 2) w[1] = b[0];
    if constexpr (R >= 4) {
        w[2] = b[1];
        w[3] = cmul(w[1], w[2]);
    }
    if constexpr (R >= 8) {
        w[4] = b[2];
        w[5] = cmul(w[1], w[4]);
        w[6] = cmul(w[2], w[4]);
        w[7] = cmul(w[3], w[4]);
    }
    if constexpr (R >= 16) {
        w[8] = b[3];
#pragma unroll
        for (int j = 9; j < 16; ++j) w[j] = cmul(w[j - 8], w[8]);
    }
}

// w[j] = W^{j*k} for j < R from the table entries W^k, W^2k, W^4k, W^8k
// (log2 R loads instead of R - 1).
template <int R>
__device__ __forceinline__ void twiddle_powers(float2 (&w)[R], const float2* __restrict__ tw, int k) {
    float2 b[4];
#pragma unroll
    for (int q = 0; q < twiddle_bases(R); ++q) b[q] = tw[k << q];
    powers_from_bases<R>(w, b);
}

// Row-packed per-pass tables, stored right after the N-entry table W_N^k:
// for pass p >= 1 (radix R, stride s, k = pos*N/(R*s)) row pos < s holds
//   ROW == 0 (bases):  W^k, W^{2k}, W^{4k}, W^{8k}      (4 entries)
//   ROW == 1 (powers): W^{jk}, j = 0 .. R-1              (R entries)
// each entry the FP64 value of W_N^{m mod N} rounded once, so a thread gets
// its butterfly's twiddles with contiguous 16-byte loads (one or two rows of
// 32/128 bytes) instead of log2 R scattered table reads.
template <int N, int E, int ROW>
struct PassTable {
    using S = Schedule<N, E>;
    static constexpr int width(int p) { return ROW == 0 ? 4 : S::radix(p); }
    static constexpr int offset(int p) {
        return p <= 1 ? 0 : offset(p - 1) + S::stride(p - 1) * width(p - 1);
    }
    static constexpr int size = offset(S::passes);
};

// Reads n2 float4 = 2*n2 complex entries starting at row (16-byte aligned).
template <int n2>
__device__ __forceinline__ void load_row(float2* w, const float2* row) {
    const float4* r4 = reinterpret_cast<const float4*>(row);
#pragma unroll
    for (int q = 0; q < n2; ++q) {
        const float4 a = r4[q];
        w[2 * q] = make_float2(a.x, a.y);
        w[2 * q + 1] = make_float2(a.z, a.w);
    }
}

// Fills PassTable<N, E, ROW> at tab (= tw + N): column c of row pos of pass
// p is W_N^{m mod N}, m = k << c (bases) or c*k (powers), evaluated in FP64
// with sincospi and rounded once, as the N-entry table is (bases rows equal
// tw[k << c] bit for bit).
template <int N, int E, int ROW>
__global__ void k_pass_table(float2* __restrict__ tab) {
    using S = Schedule<N, E>;
    using PTab = PassTable<N, E, ROW>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= PTab::size) return;
    int p = 1;
    while (p + 1 < S::passes && i >= PTab::offset(p + 1)) ++p;
    const int w = PTab::width(p), rel = i - PTab::offset(p);
    const int pos = rel / w, c = rel % w;
    const int64_t k = (int64_t)pos * (N / (S::radix(p) * S::stride(p)));
    const int64_t m = (ROW == 0 ? (k << c) : (int64_t)c * k) % N;
    double sn, cs;
    sincospi(-2.0 * (double)m / (double)N, &sn, &cs);
    tab[i] = make_float2((float)cs, (float)sn);
}

// ------------------------------------------------------------ the passes
// Runs pass P and the following ones on v (register-resident column t).
// SYNC0: put a barrier before the first LDS write too (needed when the LDS
// buffer was read by a previous transform in the same workgroup).
// TW4: derive twiddle powers from four table loads (twiddle_powers) instead
// of loading all R - 1.
// PRE: twiddles of passes >= 1 come from `pre` (registers): 1 = all R - 1
// powers per butterfly (PreTwiddles), 2 = the log2 R table entries per
// butterfly, powers derived at use (PreBases); or from the row-packed pass
// tables after the N-entry table: 3 = bases rows (powers derived),
// 4 = powers rows (PassTable).
// DSB: use ds_barrier (LDS-only ordering) instead of the fence barrier.
template <int N, int E, class L, int P, bool SYNC0, bool TW4 = true, bool DSB = false,
          int PRE = 0>
__device__ __forceinline__ void stockham_regs(float2 (&v)[E], float2* buf, int t,
                                              const float2* __restrict__ tw,
                                              const float2 (*pre)[E] = nullptr) {
    using S = Schedule<N, E>;
    constexpr int R = S::radix(P);
    constexpr int SD = S::stride(P);
    constexpr int B = E / R;  // butterflies per thread in this pass
    constexpr int T = S::T;
    constexpr int TWS = N / (R * SD);
    constexpr bool LAST = (P == S::passes - 1);

#pragma unroll
    for (int i = 0; i < B; ++i) {
        float2 x[R];
#pragma unroll
        for (int j = 0; j < R; ++j) x[j] = v[i + B * j];
        // twiddles x[j] *= w[j] (j >= 1) and the DFT-R (dft_tw: at R = 16
        // the products fold into the butterfly's first adds)
        if constexpr (SD > 1) {
            const int pos = (t + i * T) & (SD - 1);
            if constexpr (PRE == 1) {  // register twiddles of pass P (PreTwiddles)
                dft_tw<R>(x, &pre[P - 1][i * R]);
            } else if constexpr (PRE == 2) {  // register table entries (PreBases)
                float2 w[R];
                powers_from_bases<R>(w, &pre[P - 1][i * twiddle_bases(R)]);
                dft_tw<R>(x, w);
            } else if constexpr (PRE == 3) {  // packed bases row, powers derived
                constexpr int off = N + PassTable<N, E, 0>::offset(P);
                float2 b[4];
                load_row<(twiddle_bases(R) + 1) / 2>(b, tw + off + pos * 4);
                float2 w[R];
                powers_from_bases<R>(w, b);
                dft_tw<R>(x, w);
            } else if constexpr (PRE == 4) {  // packed powers row
                constexpr int off = N + PassTable<N, E, 1>::offset(P);
                float2 w[R];
                load_row<R / 2>(w, tw + off + pos * R);
                dft_tw<R>(x, w);
            } else if constexpr (TW4) {
                float2 w[R];
                twiddle_powers<R>(w, tw, pos * TWS);
                dft_tw<R>(x, w);
            } else {
                float2 w[R];
#pragma unroll
                for (int j = 1; j < R; ++j) w[j] = tw[j * pos * TWS];
                dft_tw<R>(x, w);
            }
        } else {
            dft<R>(x);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) v[i + B * j] = x[j];
    }

    // Last pass: R * SD == N, so y_j of butterfly b = t + i*T lands at
    // b + j*SD = t + T*(i + B*j), i.e. v[m] already holds X[t + T*m].
    if constexpr (!LAST) {
        if constexpr (P > 0 || SYNC0) {  // this pass's reads are done
            if constexpr (DSB) ds_barrier(); else lds_barrier();
        }
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int b = t + i * T;
            const int pos = b & (SD - 1);
            const int base = (b / SD) * (R * SD) + pos;
            lds_write_butterfly<R, SD, B, L>(buf, base, &v[i]);
        }
        if constexpr (DSB) ds_barrier(); else lds_barrier();
        lds_read_column<E, T, L>(v, buf, t);
        stockham_regs<N, E, L, P + 1, SYNC0, TW4, DSB, PRE>(v, buf, t, tw, pre);
    }
}

template <int N, int E, bool INV, bool NT>
__device__ __forceinline__ void load_column(float2 (&v)[E], const float2* src, int t, bool active) {
    constexpr int T = N / E;
    if (active) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            float2 x = gload<NT>(src + t + T * m);
            v[m] = INV ? conjf2(x) : x;
        }
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = make_float2(0.f, 0.f);
    }
}

template <int N, int E, bool INV, bool NT>
__device__ __forceinline__ void store_column(const float2 (&v)[E], float2* dst, int t) {
    constexpr int T = N / E;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        float2 y = v[m];
        if constexpr (INV) {
            constexpr float inv_n = 1.0f / (float)N;
            y = make_float2(y.x * inv_n, -y.y * inv_n);
        }
        gstore<NT>(dst + t + T * m, y);
    }
}

// Register twiddles of passes 1 .. passes-1 for thread t: butterfly i of
// pass p (radix R, B = E/R butterflies per thread) needs W_N^{j*pos*N/(R*s)},
// j < R, pos = (t + i*T) mod s, stored at w[p-1][i*R + j] (B*R = E entries
// per pass).  They do not depend on the transform, so a persistent
// workgroup loads them once.
template <int N, int E>
struct PreTwiddles {
    using S = Schedule<N, E>;
    static constexpr int NP = S::passes - 1;
    float2 w[NP > 0 ? NP : 1][E];
    template <int p = 1>
    __device__ __forceinline__ void load(const float2* __restrict__ tw, int t) {
        if constexpr (p <= NP) {
            constexpr int r = S::radix(p), b = E / r, sd = S::stride(p);
            constexpr int tws = N / (r * sd);
#pragma unroll
            for (int i = 0; i < b; ++i) {
                const int pos = (t + i * S::T) & (sd - 1);
#pragma unroll
                for (int j = 0; j < r; ++j) w[p - 1][i * r + j] = tw[j * pos * tws];
            }
            load<p + 1>(tw, t);
        }
    }
};

// Register table entries of passes 1 .. passes-1 for thread t: butterfly i
// of pass p needs W^{(2^q) k}, q < log2 R, k = pos*N/(R*s), stored at
// w[p-1][i*log2(R) + q] (at most E entries per pass); powers_from_bases
// expands them at use.  Loaded with the transform's data, their L2 latency
// leaves the pass-to-pass critical path.
template <int N, int E>
struct PreBases {
    using S = Schedule<N, E>;
    static constexpr int NP = S::passes - 1;
    float2 w[NP > 0 ? NP : 1][E];
    template <int p = 1>
    __device__ __forceinline__ void load(const float2* __restrict__ tw, int t) {
        if constexpr (p <= NP) {
            constexpr int r = S::radix(p), b = E / r, sd = S::stride(p), nb = twiddle_bases(r);
            constexpr int tws = N / (r * sd);
#pragma unroll
            for (int i = 0; i < b; ++i) {
                const int k = ((t + i * S::T) & (sd - 1)) * tws;
#pragma unroll
                for (int q = 0; q < nb; ++q) w[p - 1][i * nb + q] = tw[k << q];
            }
            load<p + 1>(tw, t);
        }
    }
    // the same entries from the bases rows of PassTable<N, E, 0>
    template <int p = 1>
    __device__ __forceinline__ void load_packed(const float2* __restrict__ tw, int t) {
        if constexpr (p <= NP) {
            constexpr int r = S::radix(p), b = E / r, sd = S::stride(p), nb = twiddle_bases(r);
            constexpr int off = N + PassTable<N, E, 0>::offset(p);
#pragma unroll
            for (int i = 0; i < b; ++i) {
                const int pos = (t + i * S::T) & (sd - 1);
                float2 row[4];
                load_row<(nb + 1) / 2>(row, tw + off + pos * 4);
#pragma unroll
                for (int q = 0; q < nb; ++q) w[p - 1][i * nb + q] = row[q];
            }
            load_packed<p + 1>(tw, t);
        }
    }
    // Marks the entries as rewritten (empty asm).  Between two transforms
    // of one thread this stops the compiler from keeping the first
    // transform's derived powers (R - 1 per pass) live for reuse by the
    // second; they are cheaper to derive again than to hold.
    template <int p = 1>
    __device__ __forceinline__ void opaque() {
        if constexpr (p <= NP) {
            constexpr int r = S::radix(p), b = E / r, nb = twiddle_bases(r);
#pragma unroll
            for (int i = 0; i < b * nb; ++i)
                __asm__ volatile("" : "+v"(w[p - 1][i].x), "+v"(w[p - 1][i].y));
            opaque<p + 1>();
        }
    }
};

// The plan's twiddle memory holds packed rows for the radix-16 schedule
// only (twiddle_table_len): E = 16 transforms take their table entries from
// registers (PreBases, re-marked before each transform so derived powers
// are not held from one to the next); E < 16 read the N-entry table.
template <int N, int E>
struct RegTwiddles {
    PreBases<N, E> pb;
    __device__ __forceinline__ void load(const float2* __restrict__ tw, int t) {
        if constexpr (E == 16) pb.load_packed(tw, t);
    }
    template <bool SYNC0, class L = LdsPad>
    __device__ __forceinline__ void fft(float2 (&v)[E], float2* buf, int t,
                                        const float2* __restrict__ tw) {
        if constexpr (E == 16) {
            pb.opaque();
            stockham_regs<N, E, L, 0, SYNC0, true, false, 2>(v, buf, t, tw, pb.w);
        } else {
            stockham_regs<N, E, L, 0, SYNC0>(v, buf, t, tw);
        }
    }
};

// One launch = ceil(batch / TPW) workgroups, TPW transforms each.  `in` and
// `out` may alias (in-place): a transform's loads all precede its stores.
// TWM, twiddles of passes >= 1: 0 = table reads at each pass (TW4 or all
// R - 1), 1 = log2 R entries per butterfly loaded right after the
// transform's data (PreBases), 2 = the same from the packed bases rows,
// 3 = packed bases rows at each pass, 4 = packed powers rows at each pass.
// Modes 2-4 need the PassTable rows after the N-entry table.
template <int N, int E, int TPW, bool INV, class L, bool NT, bool TW4 = true, int TWM = 0>
__global__ void __launch_bounds__((N / E) * TPW)
k_stockham(const float2* in, float2* out,
           const float2* __restrict__ tw, int64_t batch) {
    constexpr int T = N / E;
    __shared__ float2 lds[TPW * L::len(N)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    const int64_t tr = (int64_t)blockIdx.x * TPW + local;
    const bool active = tr < batch;
    float2* buf = lds + local * L::len(N);
    float2 v[E];
    // a slot past the batch end transforms the last transform again and
    // stores nothing: no zero-fill path beside the loads
    load_column<N, E, INV, NT>(v, in + (active ? tr : batch - 1) * N, t, true);
    if constexpr (TWM == 1 || TWM == 2) {
        PreBases<N, E> pb;
        if constexpr (TWM == 1) pb.load(tw, t); else pb.load_packed(tw, t);
        stockham_regs<N, E, L, 0, false, TW4, false, 2>(v, buf, t, tw, pb.w);
    } else {
        stockham_regs<N, E, L, 0, false, TW4, false, TWM == 0 ? 0 : TWM>(v, buf, t, tw);
    }
    if (active) store_column<N, E, INV, NT>(v, out + tr * N, t);
}

// Persistent form: a grid sized to the resident workgroups loops over the
// batch; each iteration issues the global loads of its next transform before
// running the LDS passes of the current one, so HBM reads overlap compute.
// WPE: minimum waves per SIMD the register allocation must allow (0 = free).
// TWM as in k_stockham; with 1 or 2 the table entries stay in registers for
// the whole loop (they depend on the thread only).
template <int N, int E, int TPW, bool INV, class L, bool NT, bool TW4 = true, int WPE = 0,
          int TWM = 0>
__global__ void __launch_bounds__((N / E) * TPW)
__attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1, 8)))
k_stockham_stream(const float2* in, float2* out,
                  const float2* __restrict__ tw, int64_t batch) {
    constexpr int T = N / E;
    __shared__ float2 lds[TPW * L::len(N)];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    float2* buf = lds + local * L::len(N);
    const int64_t step = (int64_t)gridDim.x * TPW;
    float2 nxt[E];
    int64_t tr = (int64_t)blockIdx.x * TPW + local;
    load_column<N, E, INV, NT>(nxt, in + tr * N, t, tr < batch);
    PreBases<N, E> pb;
    if constexpr (TWM == 1) pb.load(tw, t);
    if constexpr (TWM == 2) pb.load_packed(tw, t);
    // the loop bound depends on blockIdx only: every thread of a workgroup
    // runs the same iterations, so barriers stay uniform
    for (int64_t base = (int64_t)blockIdx.x * TPW; base < batch; base += step) {
        const int tl = fresh(t);  // per-iteration index arithmetic (not hoisted)
        tr = base + local;
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = nxt[m];
        const int64_t trn = tr + step;
        load_column<N, E, INV, NT>(nxt, in + trn * N, tl, trn < batch);
        if constexpr (TWM == 1 || TWM == 2) {
            pb.opaque();
            stockham_regs<N, E, L, 0, true, TW4, false, 2>(v, buf, tl, tw, pb.w);
        } else {
            stockham_regs<N, E, L, 0, true, TW4, false, TWM>(v, buf, tl, tw);
        }
        if (tr < batch) store_column<N, E, INV, NT>(v, out + tr * N, fresh(t));
    }
}

// Streaming form with LDS DMA (gfx950 global_load_lds_dwordx4): the next
// transform is copied HBM -> LDS by the memory pipeline, without passing
// through VGPRs, while the current one runs its passes in the other half of
// a double buffer.  Each wave moves 1 KiB blocks; the LDS image is the
// transform in natural order, which pass 0 reads as v[m] = x[t + T*m].
// Copies `count` complex64 (a multiple of 128) HBM -> LDS with
// global_load_lds_dwordx4, 1 KiB per wave instruction.
__device__ __forceinline__ void dma_transform(float2* lds_dst, const float2* src, int count, int tid,
                                              int nthreads) {
    constexpr int kBlk = 128;  // float2 per 1 KiB block: 64 lanes x 16 bytes
    const int waves = nthreads >> 6, w = tid >> 6, lane = tid & 63;
    const int blocks = count / kBlk;
    for (int b = w; b < blocks; b += waves) {
        // LDS byte address of the block (wave-uniform) -> M0; each lane's
        // 16 bytes land at M0 + 16 * lane.  Issued from asm so the compiler
        // does not treat every later LDS access as a hazard on this copy;
        // the caller waits (vmcnt) before it reads the block.
        const uint32_t ldsa = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(
            (__attribute__((address_space(3))) char*)(lds_dst + b * kBlk)));
        const float2* g = src + b * kBlk + 2 * lane;
        __asm__ __volatile__("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off"
                             :
                             : "v"(g), "s"(ldsa)
                             : "memory", "m0");
    }
}

// TPW transforms per workgroup iteration, DMA-double-buffered: one LDS half
// receives the next TPW transforms (contiguous in HBM, TPW*N*8 bytes, whole
// 1 KiB blocks for N >= 128) while the other runs the passes.
template <int N, int E, int TPW, bool INV>
__global__ void __launch_bounds__((N / E) * TPW)
k_stockham_dma(const float2* in, float2* out, const float2* __restrict__ tw, int64_t batch) {
    static_assert(N >= 128, "DMA moves whole 1 KiB blocks");
    constexpr int T = N / E;
    constexpr int NT_ = T * TPW;
    __shared__ float2 lds[2][TPW * N];
    const int local = threadIdx.x / T;
    const int t = threadIdx.x % T;
    const int64_t step = (int64_t)gridDim.x * TPW;
    PreTwiddles<N, E> pw;
    pw.load(tw, t);  // issued before any DMA: their wait does not cover a prefetch
    int cur = 0;
    int64_t base = (int64_t)blockIdx.x * TPW;
    if (base < batch) {
        const int cnt = (int)(batch - base < TPW ? batch - base : TPW);
        dma_transform(lds[0], in + base * N, cnt * N, threadIdx.x, NT_);
    }
    for (; base < batch; base += step) {  // bound uniform per workgroup
        __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA (and earlier stores) done
        ds_barrier();                   // every wave's DMA landed; previous batch finished
        const int64_t tr = base + local;
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const float2 x = lds[cur][local * N + t + T * m];
            v[m] = INV ? conjf2(x) : x;
        }
        const int64_t nb = base + step;
        if (nb < batch) {
            const int cnt = (int)(batch - nb < TPW ? batch - nb : TPW);
            dma_transform(lds[cur ^ 1], in + nb * N, cnt * N, threadIdx.x, NT_);
        }
        stockham_regs<N, E, LdsSwz, 0, true, false, true, 1>(v, lds[cur] + local * N, t, tw,
                                                                  pw.w);
        if (tr < batch) store_column<N, E, INV, true>(v, out + tr * N, t);
        cur ^= 1;
    }
}

// ------------------------------------------------- half-LDS exchange form
// E = 32 values per thread (two radix-16 butterflies per pass) and an LDS
// buffer of N/2 entries: every exchange goes through LDS in two halves.
// Butterfly b of a pass that is not the last writes all R outputs into one
// half of the transform, [0, N/2) iff b < N/(2R): its outputs lie in
// [(b/s)*R*s, (b/s + 1)*R*s) and R*s divides N/2.  With b = t + i*T that is
// i < B/2 for every thread, so half 0 is written by butterflies i < B/2 and
// read back as v[m], m < E/2 (positions t + T*m < N/2), then half 1 likewise.
// Four barriers per exchange instead of two, for half the LDS: at N = 16384
// two transforms per CU (64 KiB each) instead of one.
//
// Bases rows (PassTable<N, E, 0>) of butterfly i of pass p, into w[4i..4i+3].
template <int N, int E, int P>
__device__ __forceinline__ void half_bases(float2 (&w)[E / 16 * 4], const float2* __restrict__ tw,
                                           int t) {
    using S = Schedule<N, E>;
    if constexpr (P < S::passes) {
        constexpr int SD = S::stride(P);
        constexpr int off = N + PassTable<N, E, 0>::offset(P);
        static_assert(S::radix(P) == 16, "twiddled passes are radix 16");
#pragma unroll
        for (int i = 0; i < E / 16; ++i) {
            const int pos = (t + i * S::T) & (SD - 1);
            load_row<2>(&w[4 * i], tw + off + pos * 4);
        }
    }
}

// Pass P on v (w: bases rows of pass P, loaded during the previous
// exchange), then the exchange and the following passes.
template <int N, int E, class L, int P>
__device__ __forceinline__ void stockham_half(float2 (&v)[E], float2* buf, int t,
                                              const float2* __restrict__ tw,
                                              const float2 (&w)[E / 16 * 4]) {
    using S = Schedule<N, E>;
    constexpr int R = S::radix(P);
    constexpr int SD = S::stride(P);
    constexpr int B = E / R;
    constexpr int T = S::T;
    constexpr bool LAST = (P == S::passes - 1);
    static_assert((E == 32 || E == 16) && (B % 2 == 0 || B == 1), "butterflies per thread and pass");

#pragma unroll
    for (int i = 0; i < B; ++i) {
        float2 x[R];
#pragma unroll
        for (int j = 0; j < R; ++j) x[j] = v[i + B * j];
        if constexpr (P > 0) {
            float2 wp[R];
            powers_from_bases<R>(wp, &w[4 * i]);
            dft_tw<R>(x, wp);
        } else {
            dft<R>(x);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) v[i + B * j] = x[j];
    }

    if constexpr (!LAST) {
        float2 wn[E / 16 * 4];
        half_bases<N, E, P + 1>(wn, tw, t);  // in flight through the exchange
        float2 u[E];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (P > 0 || h > 0) lds_barrier();  // the other half's reads are done
            if constexpr (B >= 2) {
#pragma unroll
                for (int i = h * B / 2; i < (h + 1) * B / 2; ++i) {
                    const int b = t + i * T;
                    const int pos = b & (SD - 1);
                    const int base = (b / SD) * (R * SD) + pos - h * (N / 2);
                    lds_write_butterfly<R, SD, B, L>(buf, base, &v[i]);
                }
            } else if ((t < T / 2) == (h == 0)) {
                // one butterfly per thread (b = t): half 0 iff t < N/(2R) = T/2,
                // whole waves on either side
                const int pos = t & (SD - 1);
                const int base = (t / SD) * (R * SD) + pos - h * (N / 2);
                lds_write_butterfly<R, SD, B, L>(buf, base, &v[0]);
            }
            lds_barrier();
            lds_read_column<E / 2, T, L>(&u[h * E / 2], buf, t);  // positions h*N/2 + t + T*m
        }
        stockham_half<N, E, L, P + 1>(u, buf, t, tw, wn);
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = u[m];
    }
}

// One transform src -> dst by the N/32 threads t of a workgroup through
// lds (N/2 entries, swizzled); the caller's workgroup barriers nothing else.
template <int N, bool INV, bool NT>
__device__ __forceinline__ void half_transform(const float2* src, float2* dst,
                                               const float2* __restrict__ tw, float2* lds, int t,
                                               bool active) {
    constexpr int E = 32;
    float2 v[E];
    load_column<N, E, INV, NT>(v, src, t, active);
    float2 w[E / 16 * 4];
    half_bases<N, E, 1>(w, tw, t);  // pass 1's rows, in flight during pass 0
    stockham_half<N, E, LdsSwz, 0>(v, lds, t, tw, w);
    if (active) store_column<N, E, INV, NT>(v, dst, t);
}

// The same exchange with 16 values per thread (N/16 threads, one radix-16
// butterfly per thread and pass): twice the waves per transform, for the
// same N/2 entries of LDS.
template <int N, bool INV, bool NT>
__device__ __forceinline__ void half16_transform(const float2* src, float2* dst,
                                                 const float2* __restrict__ tw, float2* lds, int t,
                                                 bool active) {
    constexpr int E = 16;
    float2 v[E];
    load_column<N, E, INV, NT>(v, src, t, active);
    float2 w[E / 16 * 4];
    half_bases<N, E, 1>(w, tw, t);
    stockham_half<N, E, LdsSwz, 0>(v, lds, t, tw, w);
    if (active) store_column<N, E, INV, NT>(v, dst, t);
}

template <int N, bool INV, bool NT, int WPE>
__global__ void __launch_bounds__(N / 16)
__attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_stockham_half16(const float2* in, float2* out, const float2* __restrict__ tw, int64_t batch) {
    __shared__ float2 lds[N / 2];
    const int64_t tr = blockIdx.x;
    half16_transform<N, INV, NT>(in + tr * N, out + tr * N, tw, lds, threadIdx.x, tr < batch);
}

// One transform per workgroup of N/32 threads, LDS N/2 entries.  WPE: waves
// per SIMD the register allocation must allow (4: two 16384-point or four
// 8192-point transforms per CU).
template <int N, bool INV, bool NT, int WPE>
__global__ void __launch_bounds__(N / 32)
__attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_stockham_half(const float2* in, float2* out, const float2* __restrict__ tw, int64_t batch) {
    __shared__ float2 lds[N / 2];
    const int64_t tr = blockIdx.x;
    half_transform<N, INV, NT>(in + tr * N, out + tr * N, tw, lds, threadIdx.x, tr < batch);
}

// Transforms per workgroup: pack small transforms up to 256 threads.
template <int N, int E>
constexpr int tpw_for() {
    constexpr int T = N / E;
    return T >= 256 ? 1 : 256 / T;
}

}  // namespace mifft
