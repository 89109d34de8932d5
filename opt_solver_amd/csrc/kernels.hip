// HIP kernels of the smfgpu numeric phase (gfx950 / CDNA4).
//
// Layout of a front (see FrontDev): pivot columns in the factor array L,
// trailing contribution block in CB; col_base() gives, for local column c,
// the pointer P with element (r, c) at P[r].
#include "smfgpu_internal.hpp"
#include "mfma_tile.hpp"

#include <cfloat>
#include <type_traits>

namespace smfgpu {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;

__device__ __forceinline__ double *col_base(const DevPtrs &d, const FrontDev &F, int c) {
    if (c < F.p) return d.L + F.lofs + static_cast<int64_t>(c) * F.f;
    return d.CB + F.cbofs + static_cast<int64_t>(c - F.p) * F.q - F.p;
}

// Loads guarded by a condition are written as an unconditional load from a
// valid (clamped) address times a 0 / 1 mask: a select around the load lets
// the compiler move the load into a branch, which then waits for every load
// in flight (one memory latency per element). Front data are finite.
__device__ __forceinline__ double masked(double v, bool keep) { return v * (keep ? 1.0 : 0.0); }

// Workgroups are dealt to the 8 XCDs round-robin by index; this bijection
// gives XCD x the contiguous item range [start_x, start_x + count_x), so
// neighbouring items (which share operands) meet in the same L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_contiguous(int b, int n) {
    const int per = n / kXcds, rem = n % kXcds;
    const int x = b % kXcds, y = b / kXcds;
    return x < rem ? x * (per + 1) + y : rem * (per + 1) + (x - rem) * per + y;
}

// ---------------------------------------------------------------------------
// A1: permuted values. av[slot] = sum of the input entries mapped to slot,
// in input order (map_values_to_permuted, workspace.cpp:129-271), plus
// max |a| and a non-finite flag over the whole input (as a bit pattern max:
// |v| >= 0 orders like its bits, and an Inf/NaN compares above every
// finite value).
__global__ void k_map_values(const double *__restrict__ vals, const int64_t *__restrict__ slot_ptr,
                             const int64_t *__restrict__ slot_src, int64_t nslots,
                             double *__restrict__ av) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < nslots; s += stride) {
        double acc = 0.0;
        for (int64_t k = slot_ptr[s]; k < slot_ptr[s + 1]; ++k) acc += vals[slot_src[k]];
        av[s] = acc;
    }
}

__global__ void k_absmax(const double *__restrict__ v, int64_t n, unsigned long long *out) {
    __shared__ unsigned long long red[kThreads];
    unsigned long long m = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(fabs(v[i])));
        m = b > m ? b : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, red[0]);
}

__global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_map_values(hipStream_t st, const double *vals, int64_t nvals, const int64_t *slot_ptr,
                       const int64_t *slot_src, int64_t nslots, double *av, double *redbuf) {
    auto *rb = reinterpret_cast<unsigned long long *>(redbuf);
    k_fill_u64<<<1, 64, 0, st>>>(rb, 1, 0ull);
    if (nvals > 0) {
        const int64_t g = std::min<int64_t>((nvals + kThreads - 1) / kThreads, 2048);
        k_absmax<<<dim3(static_cast<unsigned>(g)), kThreads, 0, st>>>(vals, nvals, rb);
    }
    if (nslots > 0)
        k_map_values<<<dim3(static_cast<unsigned>(std::min<int64_t>((nslots + kThreads - 1) / kThreads, 1 << 20))),
                       kThreads, 0, st>>>(vals, slot_ptr, slot_src, nslots, av);
}

// ---------------------------------------------------------------------------
// A1: symmetric scaling. The permuted lower slots are reached through the
// full symmetric CSR (row i of it holds the slots (i, j <= i) once each);
// an entry's original row and column are r = max(perm[i], perm[j]),
// c = min(perm[i], perm[j]) (the cleaned matrix is lower in the original
// numbering).

__device__ __forceinline__ unsigned long long block_max_u64(unsigned long long m) {
    __shared__ unsigned long long red[kThreads];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

// S A S (factor_indef.cpp:530-546, apply_scale_to_matrix scaling.cpp:93-101):
// av = a * (s_r * s_c), the unscaled a kept in raw (the refinement's
// residual matrix), max |S A S| as bits into mx.
__global__ void k_scale_slots(int32_t n, const int64_t *__restrict__ fr_ptr, const int32_t *__restrict__ fr_col,
                              const int64_t *__restrict__ fr_slot, const int32_t *__restrict__ perm,
                              const double *__restrict__ s, double *__restrict__ av, double *__restrict__ raw,
                              unsigned long long *mx) {
    unsigned long long m = 0;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t pi = perm[i];
        const double si = s[pi];
        for (int64_t k = fr_ptr[i]; k < fr_ptr[i + 1]; ++k) {
            const int32_t j = fr_col[k];
            if (j > i) continue;
            const int64_t sl = fr_slot[k];
            const double a = av[sl];
            raw[sl] = a;
            const double v = a * (si * s[perm[j]]);
            av[sl] = v;
            const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(fabs(v)));
            m = b > m ? b : m;
        }
    }
    m = block_max_u64(m);
    if (threadIdx.x == 0) atomicMax(mx, m);
}

// red[0] = max|S A S| unless it holds the non-finite mark of the input
__global__ void k_take_scaled_max(unsigned long long *red) {
    if (threadIdx.x == 0 && red[0] < 0x7ff0000000000000ull) red[0] = red[3];
}

// bad = 1 when a scale entry is not finite and positive
// (validate_positive_scale, scaling.cpp:9-19)
__global__ void k_check_scale(const double *__restrict__ s, int32_t n, unsigned long long *bad) {
    unsigned long long m = 0;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = s[i];
        if (!(v > 0.0) || !(v < INFINITY)) m = 1;
    }
    m = block_max_u64(m);
    if (threadIdx.x == 0 && m) atomicMax(bad, m);
}

// One equilibration pass (compute_equilibration_scale, scaling.cpp:64-89):
// row_max_i = max over row i of |a| * s_r * s_c (that product order, the
// previous pass's s), then s_i / sqrt(row_max_i) when row_max_i > 0.
__global__ void k_equil_pass(int32_t n, const int64_t *__restrict__ fr_ptr, const int32_t *__restrict__ fr_col,
                             const int64_t *__restrict__ fr_slot, const int32_t *__restrict__ perm,
                             const double *__restrict__ av, const double *__restrict__ s_in,
                             double *__restrict__ s_out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t pi = perm[i];
    double rm = 0.0;
    for (int64_t k = fr_ptr[i]; k < fr_ptr[i + 1]; ++k) {
        const int32_t pj = perm[fr_col[k]];
        const int32_t r = max(pi, pj), c = min(pi, pj);
        const double v = fabs(av[fr_slot[k]]) * s_in[r] * s_in[c];
        rm = fmax(rm, v);
    }
    s_out[pi] = rm > 0.0 ? s_in[pi] / sqrt(rm) : s_in[pi];
}

__global__ void k_fill_f64(double *p, int64_t n, double v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_scale_values(hipStream_t st, int32_t n, const int64_t *fr_ptr, const int32_t *fr_col,
                         const int64_t *fr_slot, const int32_t *perm, const double *scale, double *av, double *raw,
                         double *redbuf) {
    auto *rb = reinterpret_cast<unsigned long long *>(redbuf);
    k_fill_u64<<<1, 64, 0, st>>>(rb + 3, 1, 0ull);
    if (n > 0) {
        const int g = std::min((n + kThreads - 1) / kThreads, 4096);
        k_scale_slots<<<dim3(g), kThreads, 0, st>>>(n, fr_ptr, fr_col, fr_slot, perm, scale, av, raw, rb + 3);
    }
    k_take_scaled_max<<<1, 64, 0, st>>>(rb);
}

void launch_check_scale(hipStream_t st, const double *scale, int32_t n, double *flag) {
    auto *fb = reinterpret_cast<unsigned long long *>(flag);
    k_fill_u64<<<1, 64, 0, st>>>(fb, 1, 0ull);
    if (n > 0) {
        const int g = std::min((n + kThreads - 1) / kThreads, 1024);
        k_check_scale<<<dim3(g), kThreads, 0, st>>>(scale, n, fb);
    }
}

void launch_equilibrate(hipStream_t st, int32_t n, const int64_t *fr_ptr, const int32_t *fr_col,
                        const int64_t *fr_slot, const int32_t *perm, const double *av, int32_t iters, double *s0,
                        double *s1, double **result) {
    double *cur = s0, *nxt = s1;
    if (n > 0) {
        const unsigned g = static_cast<unsigned>((n + kThreads - 1) / kThreads);
        k_fill_f64<<<dim3(g), kThreads, 0, st>>>(cur, n, 1.0);
        for (int32_t it = 0; it < iters; ++it) {
            k_equil_pass<<<dim3(g), kThreads, 0, st>>>(n, fr_ptr, fr_col, fr_slot, perm, av, cur, nxt);
            std::swap(cur, nxt);
        }
    }
    *result = cur;
}

// ---------------------------------------------------------------------------
// A3: assembly. One workgroup per front column; the column (rows from the
// diagonal down) is built in LDS in row chunks: zero, original entries (CSC
// order), then each child's CB column in child order (the reference's
// summation order per entry, multifrontal_serial.cpp:730-789), and written
// out once, coalesced. The chunk holds the level's longest column up to
// kAsmMaxChunk rows, so nearly every column is one chunk, each child column
// a single coalesced pass. Rows above the diagonal of a pivot column are not
// written: no kernel reads them (the diagonal blocks' upper triangles are
// written by the Cholesky kernel first) and the export zeroes them.
constexpr int kAsmMaxChunk = 16384;

__device__ __forceinline__ int lower_bound_i32(const int32_t *a, int lo, int hi, int32_t key) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Child columns per task whose scatter positions are tracked in LDS
// cursors; further ones (rare) find their chunk start by binary search.
constexpr int kAsmCursors = 128;

__global__ __launch_bounds__(kThreads) void k_assemble(DevPtrs d, const AsmTask *__restrict__ tasks,
                                                       const ChildCol *__restrict__ cc, int chunk) {
    extern __shared__ double buf[];  // chunk + 1: [chunk] is the target of masked-out adds
    __shared__ int cursor[kAsmCursors];
    __shared__ int first_out;
    const AsmTask t = tasks[blockIdx.x];
    const FrontDev F = d.fronts[t.s];
    const int c = t.c;
    const int f = F.f;
    double *colp = col_base(d, F, c);
    const int tid = threadIdx.x;
    double ssq = 0.0, dg = 0.0, offmax = 0.0;  // this thread's share of the column statistics
    // child column jj lands in parent column cm[jj]: its rows ii >= jj go to
    // the increasing rows cm[ii], so the first chunk starts at ii = jj and
    // each chunk continues where the previous one stopped
    for (int e = t.cc0 + tid; e < min(t.cc1, t.cc0 + kAsmCursors); e += kThreads) cursor[e - t.cc0] = cc[e].jj;
    for (int r0 = c; r0 < f; r0 += chunk) {
        const int r1 = min(f, r0 + chunk);
        for (int r = r0 + tid; r < r1; r += kThreads) buf[r - r0] = 0.0;
        __syncthreads();
        if (c < F.p) {
            const int j = F.col0 + c;
            const int k0 = d.pcol_ptr[j], k1 = d.pcol_ptr[j + 1];
            for (int k = k0 + tid; k < k1; k += kThreads) {
                const int lr = d.amap[k];
                const double v = d.av[k];
                buf[(lr >= r0 && lr < r1) ? lr - r0 : chunk] += v;
            }
            __syncthreads();
        }
        if (r0 == c && r1 == f) {
            // the whole column in one chunk: every entry of each child
            // column lands in it (4 loads in flight per thread)
            for (int e = t.cc0; e < t.cc1; ++e) {
                const ChildCol ch = cc[e];
                const FrontDev C = d.fronts[ch.ch];
                const int q = C.q;
                const int32_t *cm = d.cmap + d.cmofs[ch.ch];
                const double *cbcol = d.CB + C.cbofs + static_cast<int64_t>(ch.jj) * q;
                for (int i0 = ch.jj; i0 < q; i0 += 4 * kThreads) {
                    int x[4];
                    double v[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int ii = min(i0 + u * kThreads + tid, q - 1);
                        x[u] = cm[ii];
                        v[u] = cbcol[ii];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) buf[i0 + u * kThreads + tid < q ? x[u] - r0 : chunk] += v[u];
                }
                __syncthreads();
            }
        } else {
            for (int e = t.cc0; e < t.cc1; ++e) {
                const ChildCol ch = cc[e];
                const FrontDev C = d.fronts[ch.ch];
                const int q = C.q;
                const int32_t *cm = d.cmap + d.cmofs[ch.ch];
                const double *cbcol = d.CB + C.cbofs + static_cast<int64_t>(ch.jj) * q;
                const bool tracked = e - t.cc0 < kAsmCursors;
                int cur = tracked ? cursor[e - t.cc0] : lower_bound_i32(cm, ch.jj, q, r0);
                // rounds of 4 x kThreads entries; the entries of this chunk
                // are a prefix (cm increases), its end found by a minimum in
                // LDS
                while (cur < q) {
                    if (tid == 0) first_out = 4 * kThreads;
                    int x[4];
                    double v[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int ii = min(cur + u * kThreads + tid, q - 1);
                        x[u] = cm[ii];
                        v[u] = cbcol[ii];
                    }
                    __syncthreads();
                    int mine = 4 * kThreads;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const bool take = cur + u * kThreads + tid < q && x[u] < r1;
                        buf[take ? x[u] - r0 : chunk] += v[u];
                        if (!take) mine = min(mine, u * kThreads + tid);
                    }
                    if (mine < 4 * kThreads) atomicMin(&first_out, mine);
                    __syncthreads();
                    const int taken = first_out;
                    __syncthreads();
                    cur += taken;
                    if (taken < 4 * kThreads) break;
                }
                if (tracked && tid == 0) cursor[e - t.cc0] = cur;
            }
            __syncthreads();
        }
        for (int r = r0 + tid; r < r1; r += kThreads) {
            const double v = buf[r - r0];
            colp[r] = v;
            ssq += v * v;
            if (r == c) dg = fabs(v);
            else offmax = fmax(offmax, fabs(v));
        }
        __syncthreads();
    }
    if (d.colstat != nullptr && c < F.p) {
        // the selector's column statistics: 2-norm, |diagonal|, largest
        // entry below the diagonal
        __shared__ double red[3][kThreads / 64];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            ssq += __shfl_xor(ssq, o);
            dg = fmax(dg, __shfl_xor(dg, o));
            offmax = fmax(offmax, __shfl_xor(offmax, o));
        }
        if ((tid & 63) == 0) {
            red[0][tid >> 6] = ssq;
            red[1][tid >> 6] = dg;
            red[2][tid >> 6] = offmax;
        }
        __syncthreads();
        if (tid == 0) {
            double a = 0.0, b = 0.0, m = 0.0;
            for (int w = 0; w < kThreads / 64; ++w) {
                a += red[0][w];
                b = fmax(b, red[1][w]);
                m = fmax(m, red[2][w]);
            }
            d.colstat[F.col0 + c] = make_float4(static_cast<float>(sqrt(a)), static_cast<float>(b), static_cast<float>(m), 0.f);
        }
    }
}

// Levels of small fronts (no front of more than wave_rows rows, at most
// kAsmWaveRowsMax): one wavefront per column, four columns per workgroup,
// each column in its own LDS stretch. The bottom levels hold over a million short columns, and one
// workgroup each left their assembly bound by workgroup dispatch. Each entry
// is summed in the order of k_assemble (zero, original entries, children in
// order); a wavefront's LDS accesses take effect in program order, so no
// barrier separates the children.
constexpr int kAsmWaves = kThreads / 64;

__global__ __launch_bounds__(kThreads) void k_assemble_w(DevPtrs d, const AsmTask *__restrict__ tasks,
                                                         int64_t ntasks, const ChildCol *__restrict__ cc,
                                                         int stride) {
    extern __shared__ double bufw[];  // kAsmWaves stretches of stride doubles; the last of each is the junk slot
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t ti = static_cast<int64_t>(blockIdx.x) * kAsmWaves + w;
    if (ti >= ntasks) return;
    double *buf = bufw + w * stride;
    const int junk = stride - 1;
    const AsmTask t = tasks[ti];
    const FrontDev F = d.fronts[t.s];
    const int c = t.c, f = F.f, len = f - c;
    double *colp = col_base(d, F, c);
    for (int r = lane; r < len; r += 64) buf[r] = 0.0;
    if (c < F.p) {
        const int j = F.col0 + c;
        const int k0 = d.pcol_ptr[j], k1 = d.pcol_ptr[j + 1];
        for (int k = k0 + lane; k < k1; k += 64) {
            const int lr = d.amap[k];
            buf[(lr >= c && lr < f) ? lr - c : junk] += d.av[k];
        }
    }
    for (int e = t.cc0; e < t.cc1; ++e) {
        const ChildCol ch = cc[e];
        const FrontDev C = d.fronts[ch.ch];
        const int q = C.q;
        const int32_t *cm = d.cmap + d.cmofs[ch.ch];
        const double *cbcol = d.CB + C.cbofs + static_cast<int64_t>(ch.jj) * q;
        for (int i0 = ch.jj; i0 < q; i0 += 4 * 64) {
            int x[4];
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int ii = min(i0 + u * 64 + lane, q - 1);
                x[u] = cm[ii];
                v[u] = cbcol[ii];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) buf[i0 + u * 64 + lane < q ? x[u] - c : junk] += v[u];
        }
    }
    double ssq = 0.0, dg = 0.0, offmax = 0.0;
    for (int r = lane; r < len; r += 64) {
        const double v = buf[r];
        colp[c + r] = v;
        ssq += v * v;
        if (r == 0) dg = fabs(v);
        else offmax = fmax(offmax, fabs(v));
    }
    if (d.colstat != nullptr && c < F.p) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            ssq += __shfl_xor(ssq, o);
            dg = fmax(dg, __shfl_xor(dg, o));
            offmax = fmax(offmax, __shfl_xor(offmax, o));
        }
        if (lane == 0)
            d.colstat[F.col0 + c] = make_float4(static_cast<float>(sqrt(ssq)), static_cast<float>(dg),
                                                static_cast<float>(offmax), 0.f);
    }
}

void launch_assemble(hipStream_t st, const DevPtrs &d, const AsmTask *tasks, int64_t ntasks,
                     const ChildCol *cc, int32_t fmax, int32_t chunk_cap, int32_t wave_rows) {
    if (ntasks <= 0) return;
    if (fmax <= std::min(wave_rows, kAsmWaveRowsMax)) {
        static bool wattr = false;
        if (!wattr) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_assemble_w),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      kAsmWaves * (kAsmWaveRowsMax + 1) * sizeof(double));
            wattr = true;
        }
        const int stride = std::max(fmax, 1) + 1;
        const int64_t groups = (ntasks + kAsmWaves - 1) / kAsmWaves;
        for (int64_t g0 = 0; g0 < groups; g0 += kMaxLaunchGroups)
            k_assemble_w<<<dim3(static_cast<unsigned>(std::min(kMaxLaunchGroups, groups - g0))), kThreads,
                           kAsmWaves * stride * sizeof(double), st>>>(d, tasks + g0 * kAsmWaves,
                                                                      ntasks - g0 * kAsmWaves, cc, stride);
        return;
    }
    fmax = std::min(fmax, chunk_cap);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_assemble),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (kAsmMaxChunk + 1) * sizeof(double));
        attr = true;
    }
    const int chunk = std::max(1, std::min(fmax, kAsmMaxChunk));
    for (int64_t o = 0; o < ntasks; o += kMaxLaunchGroups)
        k_assemble<<<dim3(static_cast<unsigned>(std::min(kMaxLaunchGroups, ntasks - o))), kThreads,
                     (chunk + 1) * sizeof(double), st>>>(d, tasks + o, cc, chunk);
}

// ---------------------------------------------------------------------------
// A11 (SPD): Cholesky of a kb x kb diagonal block (kb <= 64) in LDS. Writes
// L_kk (lower) and M = inv(L_kk)^T's strict upper part into the block's
// strict upper triangle (M's diagonal is 1 / L_kk's), which the panel TRSM
// and the solves read. A pivot that is not > 0 or not finite records the
// front in flag[0] (smallest supernode index + 1), as dpotrf's info
// (ldlt_kernel.cpp:1246-1289); the factorization then reports
// NotPositiveDefinite.
constexpr int kNb = 64;

// kPotrfWaves wavefronts per diagonal block, each owning kPotrfCols = 64 /
// kPotrfWaves columns: thread (r, h), r the lane and h the wave, holds row r
// of the block's lower triangle in columns [C h, C h + C) (a[i] =
// A[r][C h + i], C = kPotrfCols) and rows [C h, C h + C) of column r of W,
// the identity turned into the unit lower inverse by the same row
// operations (w[i] = W[C h + i][r]). Step j: the wave owning column j puts
// the multipliers g_c = A[c][j] / d_jj, the column A[.][j] and the row
// W[j][.] in LDS (two alternating buffers: one barrier per step); then each
// thread updates A[r][c] -= A[r][j] g_c for c > j and W[i][r] -= g_i W[j][r]
// for i > j in its columns. inv(L_kk) = D^-1/2 W. Rows and columns past kb
// are the identity. The 64 steps are a dependent chain; more waves shorten
// each step's update (2 C FMAs per thread).
constexpr int kPotrfWaves = 8;
constexpr int kPotrfCols = kNb / kPotrfWaves;

__device__ __forceinline__ void potrf_steps(double (&a)[kPotrfCols], double (&w)[kPotrfCols], int r, int h, int kb,
                                            double (*gl)[kNb], double (*cola)[kNb], double (*colw)[kNb], double *rsv,
                                            bool &bad) {
    constexpr int C = kPotrfCols;
    for (int jb = 0; jb < kPotrfWaves; ++jb) {
#pragma unroll
        for (int jj = 0; jj < C; ++jj) {
            const int j = C * jb + jj;
            const int bs = j & 1;
            if (h == jb) {
                const double ajj = a[jj];
                const double djj = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(ajj), j),
                                                    __builtin_amdgcn_readlane(__double2loint(ajj), j));
                // 1 / sqrt(djj): the hardware estimate refined by two
                // Newton steps (y <- y (3 - d y^2) / 2)
                double rs = __builtin_amdgcn_rsq(djj);
                rs = rs * (1.5 - 0.5 * djj * rs * rs);
                rs = rs * (1.5 - 0.5 * djj * rs * rs);
                gl[bs][r] = r > j ? ajj * (rs * rs) : 0.0;
                cola[bs][r] = ajj;
                colw[bs][r] = w[jj];
                if (r == 0) rsv[j] = rs;
                bad = bad || (j < kb && !(djj > 0.0 && djj <= DBL_MAX));
                a[jj] = r > j ? ajj * rs : (r == j ? djj * rs : ajj);
            }
            __syncthreads();
            const double arj = cola[bs][r], wjr = colw[bs][r];
            if (h == jb) {
#pragma unroll
                for (int i = jj + 1; i < C; ++i) {
                    a[i] -= arj * gl[bs][C * jb + i];
                    w[i] -= gl[bs][C * jb + i] * wjr;
                }
            } else if (h > jb) {
#pragma unroll
                for (int i = 0; i < C; ++i) {
                    a[i] -= arj * gl[bs][C * h + i];
                    w[i] -= gl[bs][C * h + i] * wjr;
                }
            }
        }
    }
}

__global__ __launch_bounds__(kNb * kPotrfWaves) void k_potrf(DevPtrs d, const PotrfItem *__restrict__ items) {
    constexpr int C = kPotrfCols;
    __shared__ double gl[2][kNb];
    __shared__ double cola[2][kNb];
    __shared__ double colw[2][kNb];
    __shared__ double rsv[kNb];
    __shared__ int flag_bad;
    const PotrfItem it = items[blockIdx.x];
    const FrontDev F = d.fronts[it.s];
    const int f = F.f, k = it.k, kb = it.kb;
    double *base = d.L + F.lofs;
    const int r = threadIdx.x & 63;
    const int h = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (threadIdx.x == 0) flag_bad = 0;
    bool bad = false;
    double a[C], w[C];
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const int c = C * h + i;
        a[i] = base[static_cast<int64_t>(k + min(c, kb - 1)) * f + k + min(r, kb - 1)];
    }
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const int c = C * h + i;
        const bool in = r < kb && c < kb && r >= c;
        a[i] = masked(a[i], in) + ((!in && r == c) ? 1.0 : 0.0);
        w[i] = (c == r) ? 1.0 : 0.0;
    }
    potrf_steps(a, w, r, h, kb, gl, cola, colw, rsv, bad);
    if (bad) flag_bad = 1;
    __syncthreads();
    if (threadIdx.x == 0 && flag_bad) atomicMin(d.flag, it.s + 1);
    // lower (c <= r): L_kk; strict upper (r < c): M[r][c] = inv(L_kk)[c][r]
    // = W[c][r] / L_cc
    if (r < kb) {
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const int c = C * h + i;
            if (c < kb) base[static_cast<int64_t>(k + c) * f + k + r] = c <= r ? a[i] : w[i] * rsv[c];
        }
    }
}

void launch_potrf(hipStream_t st, const DevPtrs &d, const PotrfItem *it, int64_t n) {
    if (n <= 0) return;
    for (int64_t o = 0; o < n; o += kMaxLaunchGroups)
        k_potrf<<<dim3(static_cast<unsigned>(std::min(kMaxLaunchGroups, n - o))), kNb * kPotrfWaves, 0, st>>>(d, it + o);
}

// Panel TRSM: rows [r0, r0 + 64) of columns [k, k + kb): X = A M, with
// M = inv(L_kk)^T read from the block's strict upper part and diagonal, on
// FP64 MFMA. Computed transposed (X^T = M^T A^T: MFMA A operand = M^T,
// B operand = A^T) so that 16 lanes store 16 consecutive rows of a column.
// Wave w: rows 16 w .. 16 w + 15, four 16-column blocks; block cb needs
// k-steps t < 16 cb + 16 only (M is upper triangular).
__global__ __launch_bounds__(kThreads) void k_trsm(DevPtrs d, const TrsmItem *__restrict__ items) {
    // M in LDS (every wave reads all of it); each wave's 16 rows of A, the
    // MFMA B operand, in registers: 34 KB of LDS a workgroup, four per CU
    __shared__ double Ms[kNb][kNb + 2];
    const TrsmItem it = items[blockIdx.x];
    const FrontDev F = d.fronts[it.s];
    const int f = F.f, k = it.k, kb = it.kb, r0 = it.r0;
    double *base = d.L + F.lofs;
    const int tid = threadIdx.x;
    const int nr = min(kNb, f - r0);
    const int lane = tid & 63, w = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int rr = 16 * w + lr;
    double bv[16];
    {
        const int rc = r0 + min(rr, nr - 1);
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) bv[kk] = base[static_cast<int64_t>(k + min(4 * kk + lk, kb - 1)) * f + rc];
        const int mr = tid & 63, cg = tid >> 6;
        double gv[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) gv[i] = base[static_cast<int64_t>(k + min(cg + 4 * i, kb - 1)) * f + k + min(mr, kb - 1)];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int t = cg + 4 * i;
            // Ms[t = mr][c]: strict upper M, diagonal 1 / L_cc
            const double g = gv[i];
            const bool in = mr < kb && t < kb;
            Ms[mr][t] = (in && mr < t) ? g : ((in && mr == t) ? 1.0 / g : 0.0);
        }
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) bv[kk] = masked(bv[kk], rr < nr && 4 * kk + lk < kb);
    }
    __syncthreads();
    v4d acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
        const int t = 4 * kk + lk;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
            if (kk < 4 * cb + 4) acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ms[t][16 * cb + lr], bv[kk], acc[cb], 0, 0, 0);
    }
    // fronts with reduced-precision Schur updates also get the panel in
    // the precision their tiles read: FP32 (half the bytes), or bf16 in the
    // packed layout (high part, and low part for split BF16)
    const int pc = d.prec != nullptr ? d.prec[it.s] : 0;
    float *b32 = (d.L32 != nullptr && pc == 1) ? d.L32 + F.lofs : nullptr;
    if ((pc == 2 || pc == 3) && d.L16h != nullptr) {
        __syncthreads();  // every wave is done with M: its LDS stages the result
        if (rr < nr) {
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                for (int v = 0; v < 4; ++v) Ms[rr][16 * cb + 4 * v + lk] = acc[cb][v];
        }
        __syncthreads();
        // 8-column groups of one row as 16-byte vectors; columns past kb
        // (the front's last pivot) stay zero
        const bool split = pc == 3 && d.L16l != nullptr;
        for (int e = tid; e < kNb * (kNb / 8); e += kThreads) {
            const int row = e & 63, g = e >> 6;
            if (row >= nr || 8 * g >= kb) continue;
            v8bf_t hi, lo;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const float x = 8 * g + u < kb ? static_cast<float>(Ms[row][8 * g + u]) : 0.f;
                hi[u] = static_cast<__bf16>(x);
                lo[u] = static_cast<__bf16>(x - static_cast<float>(hi[u]));
            }
            const int64_t o = F.l16ofs + static_cast<int64_t>(k / 8 + g) * f * 8 + static_cast<int64_t>(r0 + row) * 8;
            *reinterpret_cast<v8bf_t *>(d.L16h + o) = hi;
            if (split) *reinterpret_cast<v8bf_t *>(d.L16l + o) = lo;
        }
    }
    if (rr < nr) {
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = 16 * cb + 4 * v + lk;
                if (c < kb) {
                    const int64_t o = static_cast<int64_t>(k + c) * f + r0 + rr;
                    base[o] = acc[cb][v];
                    if (b32) b32[o] = static_cast<float>(acc[cb][v]);
                }
            }
    }
}

void launch_trsm(hipStream_t st, const DevPtrs &d, const TrsmItem *it, int64_t n) {
    if (n <= 0) return;
    for (int64_t o = 0; o < n; o += kMaxLaunchGroups)
        k_trsm<<<dim3(static_cast<unsigned>(std::min(kMaxLaunchGroups, n - o))), kThreads, 0, st>>>(d, it + o);
}

// ---------------------------------------------------------------------------
// A10/A11: Schur-complement update on MFMA (tile machinery in mfma_tile.hpp).
// Precision codes of a front's Schur updates (DevPtrs::prec): 0 FP64 MFMA;
// 1 FP32 MFMA on the FP32 panel copy; 2 BF16 MFMA (operands rounded to
// bf16), 3 split BF16 (hi / lo operands, four MFMAs); 1-3 accumulate in
// FP32 over the panel and subtract the result from the FP64 Schur
// complement.
constexpr int kSyrkKc64 = 16;
constexpr int kSyrkKc32 = 32;
constexpr int kSyrkKcBf = 64;
constexpr int kSyrkKcBf3 = 32;

template <int MODE, int kAhead>
__device__ __forceinline__ void syrk_tile(const DevPtrs &d, const SyrkItem &it, unsigned char *smem) {
    typedef typename std::conditional<MODE == 0, double, float>::type E;
    typedef SyrkMfma<E> M;
    const FrontDev F = d.fronts[it.s];
    const int f = F.f;
    const int i0 = it.i0, j0 = it.j0;
    const int j1 = min(it.j1, f);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    const bool active = (i0 + wr * 64 + 63) >= (j0 + wc * 64);
    typename M::acc_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = M::zero();
    if constexpr (MODE == 0) {
        const double *Lb = d.L + F.lofs;
        if (active) tile_loop<double, true, kSyrkKc64, kAhead>(Lb, Lb, f, 0, i0, j0, f, j1, it.t0, it.t1, smem, acc);
        else tile_loop<double, false, kSyrkKc64, kAhead>(Lb, Lb, f, 0, i0, j0, f, j1, it.t0, it.t1, smem, acc);
    } else if constexpr (MODE == 1) {
        const float *Lb = d.L32 + F.lofs;
        if (active) tile_loop<float, true, kSyrkKc32, kAhead>(Lb, Lb, f, 0, i0, j0, f, j1, it.t0, it.t1, smem, acc);
        else tile_loop<float, false, kSyrkKc32, kAhead>(Lb, Lb, f, 0, i0, j0, f, j1, it.t0, it.t1, smem, acc);
    } else {
        constexpr bool kSplit = MODE == 3;
        constexpr int kKc = kSplit ? kSyrkKcBf3 : kSyrkKcBf;
        const uint16_t *Lh = d.L16h + F.l16ofs;
        const uint16_t *Ll = kSplit ? d.L16l + F.l16ofs : Lh;
        if (active) tile_loop_bf16p<kSplit, true, kKc>(Lh, Ll, Lh, Ll, f, i0, j0, f, j1, it.t0, it.t1, smem, acc);
        else tile_loop_bf16p<kSplit, false, kKc>(Lh, Ll, Lh, Ll, f, i0, j0, f, j1, it.t0, it.t1, smem, acc);
    }
    if (!active) return;
    // C -= acc, 16 elements at a time: all 16 loads issued before the
    // stores; elements outside the tile's lower part read and write a junk
    // slot instead, so loads and stores are unconditional
    double *junk = d.junk + threadIdx.x;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        double *dst[4][4];
        double old[4][4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = j0 + wc * 64 + m * 16 + M::row(v, lane);
            double *cb = col_base(d, F, min(j, j1 - 1));
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) {
                const int i = i0 + wr * 64 + nn * 16 + (lane & 15);
                dst[nn][v] = (i < f && j < j1 && i >= j) ? cb + i : junk;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) old[nn][v] = *dst[nn][v];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) *dst[nn][v] = old[nn][v] - static_cast<double>(acc[m][nn][v]);
    }
}

// A10/A11: Schur-complement update on MFMA. Tile 128 x 128 of
// C[i, j] -= sum_{t in [t0, t1)} L[i, t] L[j, t], i >= j, four waves of
// 64 x 64 (4 x 4 MFMA blocks of 16 x 16). The MFMA computes the transposed
// block (A operand = the j rows, B operand = the i rows) so that the 16
// lanes of a row group hold 16 consecutive rows of one column on the store.
// Two kernels share each step's item list: k_syrk64 runs the items of FP64
// fronts, k_syrk_red those of reduced-precision fronts; each returns at
// once on the other's items. Separate kernels keep the FP32-accumulator
// tiles free of the FP64 tile's register count. The FP64 tile stages 16
// panel columns per LDS stage (64 KB for both operands' two stages: two
// workgroups per CU); the FP32 tile 32; the BF16 tile 64, the split-BF16
// tile 32 (hi and lo planes).
template <int kAhead>
__global__ __launch_bounds__(kThreads) void k_syrk64(DevPtrs d, const SyrkItem *__restrict__ items) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[tile_lds_bytes<double, kSyrkKc64>()];
    const SyrkItem it = items[xcd_contiguous(blockIdx.x, gridDim.x)];
    if (d.prec != nullptr && d.prec[it.s] != 0) return;
    syrk_tile<0, kAhead>(d, it, smem);
}

template <int MODE>
struct RedLds {
    static constexpr int bytes = MODE == 1   ? tile_lds_bytes<float, kSyrkKc32>()
                                 : MODE == 2 ? bf_lds_bytes<false, kSyrkKcBf>()
                                             : bf_lds_bytes<true, kSyrkKcBf3>();
};

// one kernel per reduced precision (its own register count); each returns
// at once on the items of fronts of another precision
template <int MODE, int kAhead>
__global__ __launch_bounds__(kThreads) void k_syrk_red(DevPtrs d, const SyrkItem *__restrict__ items) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[RedLds<MODE>::bytes];
    const SyrkItem it = items[xcd_contiguous(blockIdx.x, gridDim.x)];
    if (d.prec == nullptr || d.prec[it.s] != MODE) return;
    syrk_tile<MODE, kAhead>(d, it, smem);
}

// all three reduced precisions in one kernel, for the learned selector's
// runs (the precisions are chosen on the device, so one launch covers
// whichever codes the fronts got): LDS and registers of the largest
template <int MODE>
struct RedMax {
    static constexpr int a = RedLds<1>::bytes, b = RedLds<2>::bytes, c = RedLds<3>::bytes;
    static constexpr int bytes = a > b ? (a > c ? a : c) : (b > c ? b : c);
};
__global__ __launch_bounds__(kThreads) void k_syrk_red_any(DevPtrs d, const SyrkItem *__restrict__ items) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[RedMax<0>::bytes];
    const SyrkItem it = items[xcd_contiguous(blockIdx.x, gridDim.x)];
    const int pc = d.prec == nullptr ? 0 : d.prec[it.s];
    if (pc == 1) syrk_tile<1, 1>(d, it, smem);
    else if (pc == 2) syrk_tile<2, 1>(d, it, smem);
    else if (pc == 3) syrk_tile<3, 1>(d, it, smem);
}

// n items; the reduced kernels cover the last n_red of them (the items of
// fronts that may run below FP64), one launch per precision code in
// red_modes (bit c: code c), or one launch of the combined kernel when
// several codes are possible
void launch_syrk(hipStream_t st, const DevPtrs &d, const SyrkItem *it, int64_t n, int64_t n_red, int ahead,
                 int red_modes) {
    if (n <= 0) return;
    for (int64_t o = 0; o < n; o += kMaxLaunchGroups) {
        const dim3 grid(static_cast<unsigned>(std::min(kMaxLaunchGroups, n - o)));
        if (ahead == 1) k_syrk64<1><<<grid, kThreads, 0, st>>>(d, it + o);
        else k_syrk64<2><<<grid, kThreads, 0, st>>>(d, it + o);
    }
    const SyrkItem *ir = it + (n - n_red);
    for (int64_t o = 0; o < n_red; o += kMaxLaunchGroups) {
        const dim3 grid(static_cast<unsigned>(std::min(kMaxLaunchGroups, n_red - o)));
        if (__builtin_popcount(red_modes & 14) > 1) {
            k_syrk_red_any<<<grid, kThreads, 0, st>>>(d, ir + o);
            continue;
        }
        if (red_modes & 2) {
            if (ahead == 1) k_syrk_red<1, 1><<<grid, kThreads, 0, st>>>(d, ir + o);
            else k_syrk_red<1, 2><<<grid, kThreads, 0, st>>>(d, ir + o);
        }
        // the bf16 tiles keep one register set of loads in flight
        if (red_modes & 4) k_syrk_red<2, 1><<<grid, kThreads, 0, st>>>(d, ir + o);
        if (red_modes & 8) k_syrk_red<3, 1><<<grid, kThreads, 0, st>>>(d, ir + o);
    }
}

void launch_zero_upper(hipStream_t, const DevPtrs &, const PotrfItem *, int64_t) {}

// ---------------------------------------------------------------------------
// A15: solves. y is the permuted right-hand-side block (n x nrhs, ld n).

// With a scale s (original numbering), y_i = b_perm[i] * s_perm[i] and
// b_perm[i] = y_i * s_perm[i]: the scaled gather and scatter of solve_core's
// full job (solve.cpp:1022-1047), which the forward and backward jobs equal
// (apply_scale_to_rhs before the permuting forward sweep, apply_scale_to_
// solution after the permuting backward sweep, scaling.cpp:105-119).
__global__ void k_gather(const int32_t *__restrict__ perm, const double *__restrict__ b, int64_t ldb,
                         double *__restrict__ y, int32_t n, int32_t nrhs, const double *__restrict__ sc) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t o = perm[i];
    if (sc) {
        const double so = sc[o];
        for (int r = 0; r < nrhs; ++r) y[i + static_cast<int64_t>(r) * n] = b[o + r * ldb] * so;
    } else {
        for (int r = 0; r < nrhs; ++r) y[i + static_cast<int64_t>(r) * n] = b[o + r * ldb];
    }
}

__global__ void k_scatter(const int32_t *__restrict__ perm, const double *__restrict__ y, double *__restrict__ b,
                          int64_t ldb, int32_t n, int32_t nrhs, const double *__restrict__ sc) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t o = perm[i];
    if (sc) {
        const double so = sc[o];
        for (int r = 0; r < nrhs; ++r) b[o + r * ldb] = y[i + static_cast<int64_t>(r) * n] * so;
    } else {
        for (int r = 0; r < nrhs; ++r) b[o + r * ldb] = y[i + static_cast<int64_t>(r) * n];
    }
}

void launch_gather_rhs(hipStream_t st, const int32_t *perm, const double *b, int64_t ldb, double *y,
                       int32_t n, int32_t nrhs, const double *scale) {
    if (n <= 0) return;
    k_gather<<<dim3((n + kThreads - 1) / kThreads), kThreads, 0, st>>>(perm, b, ldb, y, n, nrhs, scale);
}

void launch_scatter_rhs(hipStream_t st, const int32_t *perm, const double *y, double *b, int64_t ldb,
                        int32_t n, int32_t nrhs, const double *scale) {
    if (n <= 0) return;
    k_scatter<<<dim3((n + kThreads - 1) / kThreads), kThreads, 0, st>>>(perm, y, b, ldb, n, nrhs, scale);
}

// ---------------------------------------------------------------------------
// A16: refinement. Residual r = b - A x in permuted numbering over the full
// symmetric CSR, products and sums in double-double (TwoProd by fma,
// TwoSum), as the reference's compute_residual_and_backward_error
// (refinement.cpp:39-130) sums in long double, and per column the four
// maxima the backward errors are formed from, reduced on the device:
// stat[4c] = max |r_i|, stat[4c+1] = max_i |r_i| / ((|A||x|)_i + |b_i|)
// (+inf when that denominator is 0 and r_i is not), stat[4c+2] = max |x_i|,
// stat[4c+3] = max |b_i|. Maxima are taken over the bit patterns of the
// non-negative values (an Inf or NaN sorts above every finite value), by one
// atomic max per workgroup; stat is cleared by the caller.
// Knuth's TwoSum: s + e = a + b exactly. FMA contraction is off here and
// in the residual loop: a product fused into these sums (the default under
// -ffp-contract=fast) would break the error-free transformation.
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
#pragma clang fp contract(off)
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ unsigned long long ubits(double v) {
    return static_cast<unsigned long long>(__double_as_longlong(v));
}

// workgroup max of m (kThreads threads) into *out by one atomic
__device__ __forceinline__ void block_max_atomic(unsigned long long m, unsigned long long *out) {
    __shared__ unsigned long long red[kThreads / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long v = __shfl_xor(m, o, 64);
        m = v > m ? v : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = red[0];
        for (int w = 1; w < kThreads / 64; ++w) b = red[w] > b ? red[w] : b;
        atomicMax(out, b);
    }
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_residual(int32_t n, const int64_t *__restrict__ rp,
                                                       const int32_t *__restrict__ ci,
                                                       const int64_t *__restrict__ slot,
                                                       const double *__restrict__ av, const double *__restrict__ x,
                                                       const double *__restrict__ b, double *__restrict__ r,
                                                       unsigned long long *__restrict__ stat, int32_t nrhs) {
#pragma clang fp contract(off)
    for (int c = 0; c < nrhs; ++c) {
        const double *xc = x + static_cast<int64_t>(c) * n;
        unsigned long long m[4] = {0, 0, 0, 0};
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
            const double bi = b[i + static_cast<int64_t>(c) * n];
            double sh = bi, sl = 0.0, absax = 0.0;
            for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
                const double a = av[slot[k]];
                const double xv = xc[ci[k]];
                const double ph = -a * xv;           // rounded product
                const double pl = fma(-a, xv, -ph);  // its exact rounding error
                double s2, e;
                two_sum(sh, ph, s2, e);
                sh = s2;
                sl += e + pl;
                absax += fabs(a) * fabs(xv);
            }
            const double res = sh + sl;
            r[i + static_cast<int64_t>(c) * n] = res;
            const double ar = fabs(res), den = absax + fabs(bi);
            const double om = den > 0.0 ? ar / den : (ar == 0.0 ? 0.0 : INFINITY);
            const unsigned long long v[4] = {ubits(ar), ubits(fabs(om)), ubits(fabs(xc[i])), ubits(fabs(bi))};
#pragma unroll
            for (int k = 0; k < 4; ++k) m[k] = v[k] > m[k] ? v[k] : m[k];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) block_max_atomic(m[k], stat + 4 * c + k);
    }
}

void launch_residual(hipStream_t st, int32_t n, const int64_t *rp, const int32_t *ci, const int64_t *slot,
                     const double *av, const double *x, const double *b, double *r, double *stat,
                     int32_t nrhs) {
    auto *o = reinterpret_cast<unsigned long long *>(stat);
    k_fill_u64<<<dim3((4 * nrhs + 63) / 64), 64, 0, st>>>(o, 4 * nrhs, 0ull);
    if (n <= 0) return;
    const int64_t g = std::min<int64_t>((n + kThreads - 1) / kThreads, 2048);
    k_residual<<<dim3(static_cast<unsigned>(g)), kThreads, 0, st>>>(n, rp, ci, slot, av, x, b, r, o, nrhs);
}

// w = A z (FP64, full symmetric CSR, permuted numbering)
__global__ __launch_bounds__(kThreads) void k_spmv(int32_t n, const int64_t *__restrict__ rp,
                                                   const int32_t *__restrict__ ci, const int64_t *__restrict__ slot,
                                                   const double *__restrict__ av, const double *__restrict__ z,
                                                   double *__restrict__ w) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += av[slot[k]] * z[ci[k]];
    w[i] = s;
}

void launch_spmv(hipStream_t st, int32_t n, const int64_t *rp, const int32_t *ci, const int64_t *slot,
                 const double *av, const double *z, double *w) {
    if (n <= 0) return;
    k_spmv<<<dim3((n + kThreads - 1) / kThreads), kThreads, 0, st>>>(n, rp, ci, slot, av, z, w);
}

// Dot products of w with nvec columns of V (ld n), in two passes with a
// fixed summation order (deterministic): partial sums per workgroup, then
// one workgroup per column sums them; out[i] = sum (acc: out[i] += sum;
// root: out[i] = sqrt of the sum).
__global__ __launch_bounds__(kThreads) void k_dot_part(int64_t n, const double *__restrict__ V, int32_t nvec,
                                                       const double *__restrict__ w, double *__restrict__ part) {
    __shared__ double red[kThreads];
    const int i = blockIdx.y;
    const double *v = V + static_cast<int64_t>(i) * n;
    double s = 0.0;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; k < n;
         k += static_cast<int64_t>(gridDim.x) * kThreads)
        s += v[k] * w[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[static_cast<int64_t>(i) * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kThreads) void k_dot_finish(const double *__restrict__ part, int32_t nblk,
                                                         double *__restrict__ out, int32_t mode) {
    __shared__ double red[kThreads];
    const int i = blockIdx.x;
    double s = 0.0;
    for (int k = threadIdx.x; k < nblk; k += kThreads) s += part[static_cast<int64_t>(i) * nblk + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (mode == 1) out[i] += red[0];
        else if (mode == 2) out[i] = sqrt(red[0]);
        else out[i] = red[0];
    }
}

void launch_dots(hipStream_t st, int64_t n, const double *V, int32_t nvec, const double *w, double *part,
                 double *out, int32_t mode) {
    if (nvec <= 0) return;
    const int nblk = static_cast<int>(std::min<int64_t>(kDotParts, std::max<int64_t>(1, (n + kThreads - 1) / kThreads)));
    k_dot_part<<<dim3(nblk, nvec), kThreads, 0, st>>>(n, V, nvec, w, part);
    k_dot_finish<<<dim3(nvec), kThreads, 0, st>>>(part, nblk, out, mode);
}

// w -= V h (nvec columns, ld n; h on the device)
__global__ __launch_bounds__(kThreads) void k_vsub(int64_t n, const double *__restrict__ V, int32_t nvec,
                                                   const double *__restrict__ h, double *__restrict__ w) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (k >= n) return;
    double s = w[k];
    for (int i = 0; i < nvec; ++i) s -= V[static_cast<int64_t>(i) * n + k] * h[i];
    w[k] = s;
}

void launch_vsub(hipStream_t st, int64_t n, const double *V, int32_t nvec, const double *h, double *w) {
    if (n <= 0 || nvec <= 0) return;
    k_vsub<<<dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)), kThreads, 0, st>>>(n, V, nvec, h, w);
}

// out = w / *d (d on the device)
__global__ __launch_bounds__(kThreads) void k_vdiv(int64_t n, const double *__restrict__ w,
                                                   const double *__restrict__ d, double *__restrict__ out) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (k < n) out[k] = w[k] / d[0];
}

void launch_vdiv(hipStream_t st, int64_t n, const double *w, const double *d, double *out) {
    if (n <= 0) return;
    k_vdiv<<<dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)), kThreads, 0, st>>>(n, w, d, out);
}

// out = x0 + Z y (nvec columns of Z, ld n; y on the device)
__global__ __launch_bounds__(kThreads) void k_lincomb(int64_t n, const double *__restrict__ x0,
                                                      const double *__restrict__ Z, int32_t nvec,
                                                      const double *__restrict__ y, double *__restrict__ out) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (k >= n) return;
    double s = x0[k];
    for (int i = 0; i < nvec; ++i) s += y[i] * Z[static_cast<int64_t>(i) * n + k];
    out[k] = s;
}

void launch_lincomb(hipStream_t st, int64_t n, const double *x0, const double *Z, int32_t nvec, const double *y,
                    double *out) {
    if (n <= 0) return;
    k_lincomb<<<dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)), kThreads, 0, st>>>(n, x0, Z, nvec, y,
                                                                                                out);
}

// row sums of |A| (for ||A||_inf)
__global__ void k_rowabs(int32_t n, const int64_t *__restrict__ rp, const int64_t *__restrict__ slot,
                         const double *__restrict__ av, double *__restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += fabs(av[slot[k]]);
    out[i] = s;
}

void launch_rowabs(hipStream_t st, int32_t n, const int64_t *rp, const int64_t *slot, const double *av,
                   double *out) {
    if (n <= 0) return;
    k_rowabs<<<dim3((n + kThreads - 1) / kThreads), kThreads, 0, st>>>(n, rp, slot, av, out);
}

// x = y + d (y may be x)
__global__ void k_axpy(int64_t n, const double *__restrict__ y, const double *__restrict__ dlt, double *x) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = y[i] + dlt[i];
}

void launch_axpy(hipStream_t st, int64_t n, const double *y, const double *dlt, double *x) {
    if (n <= 0) return;
    k_axpy<<<dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)), kThreads, 0, st>>>(n, y, dlt, x);
}

// max over i of v[i] (v >= 0), into *out (bit-pattern max, out zeroed first)
__global__ void k_maxbits(const double *__restrict__ v, int64_t n, int64_t stride, unsigned long long *out) {
    __shared__ unsigned long long red[kThreads];
    unsigned long long m = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const unsigned long long bb = static_cast<unsigned long long>(__double_as_longlong(v[i * stride]));
        m = bb > m ? bb : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, red[0]);
}

void launch_reduce_max(hipStream_t st, const double *v, int64_t n, double *out) {
    auto *o = reinterpret_cast<unsigned long long *>(out);
    k_fill_u64<<<1, 64, 0, st>>>(o, 1, 0ull);
    if (n <= 0) return;
    const int64_t g = std::min<int64_t>((n + kThreads - 1) / kThreads, 1024);
    k_maxbits<<<dim3(static_cast<unsigned>(g)), kThreads, 0, st>>>(v, n, 1, o);
}

}  // namespace smfgpu
