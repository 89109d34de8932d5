// A15: forward and backward sweeps of the SPD factor, level by level.
//
// A front's pivot columns are solved in blocks of kSB = 128 (two 64-pivot
// diagonal blocks whose inverses the Cholesky kernel stored in their strict
// upper triangles). One launch per step covers every front of a level:
// in step j one workgroup per front (the "triangle" item) applies block
// j-1's update to block j's rows and solves block j, while the other items
// apply block j-1's update to the rest of the front (forward: 64-row chunks;
// backward: 64-column chunks). Kernel boundaries order the steps, no
// workgroup waits for another. Reference: solve.cpp:858-1075 (the sweeps),
// solve_forward.cpp / solve_backward.cpp (per-front dense steps).
//
// y is the permuted right-hand side block, n x nrhs (ld n). Rows of a front
// below its pivots belong to ancestors that sibling fronts of the same level
// also update: those forward updates are atomic adds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "smfgpu_internal.hpp"
#include "smfgpu_ldlt.hpp"

namespace smfgpu {
namespace {

constexpr int kSB = 128;        // pivot columns per step
constexpr int kSD = 64;         // diagonal block (stored inverse)
constexpr int kSThreads = 512;  // eight waves
constexpr int kSeg = 512;       // rows per backward update item

struct TriRegs {
    double a[8], b[8], c[8];  // Dg_a, L_ba, Dg_b of the triangle, element (rl, g + 8 e)
};

// Element (t = rl, c = g + 8 e) of the triangle's three 64 x 64 blocks:
// a: upper part of the first diagonal block (t <= c: inverse rows and the
// diagonal), b: rows 64.. of the block below it, c: second diagonal block.
// Guarded loads are unconditional loads from clamped (in-front) addresses
// times a 0 / 1 mask: a select around a load lets the compiler move the
// load into a branch that waits for every load in flight.
__device__ __forceinline__ double masked(double v, bool keep) { return v * (keep ? 1.0 : 0.0); }

__device__ __forceinline__ void tri_prefetch(TriRegs &R, const double *Lb, int f, int tk, int tkb, int rl, int g) {
    const int na = min(kSD, tkb), nb = tkb - kSD;
    const int ra = tk + min(rl, na - 1);                     // rows of block a
    const int rb = nb > 0 ? tk + kSD + min(rl, nb - 1) : ra;  // rows of block b (a's when there is none)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int c = g + 8 * e;
        const int ca = tk + min(c, na - 1);
        const int cb = nb > 0 ? tk + kSD + min(c, nb - 1) : ca;
        R.a[e] = masked(Lb[static_cast<int64_t>(ca) * f + ra], rl <= c && c < na);
        R.b[e] = masked(Lb[static_cast<int64_t>(ca) * f + rb], rl < nb && c < na);
        R.c[e] = masked(Lb[static_cast<int64_t>(cb) * f + rb], rl <= c && c < nb);
    }
}

__device__ __forceinline__ void put_block(double (*Dg)[kSD + 1], const double (&v)[8], int rl, int g) {
#pragma unroll
    for (int e = 0; e < 8; ++e) Dg[rl][g + 8 * e] = v[e];
}

// z[c] = y[c] / L_cc + sum_{t<c} M[t][c] y[t]  (M[t][c] = inv(L)[c][t])
template <int NR>
__device__ __forceinline__ void tri_fwd(const double (*Dg)[kSD + 1], const double (*yb)[NR], double (*z)[NR], int nb,
                                        int nrhs) {
    const int c = threadIdx.x & 63, r = threadIdx.x >> 6;
    if (r < NR && r < nrhs && c < nb) {
        double s = yb[c][r] / Dg[c][c];
        for (int t = 0; t < c; ++t) s += Dg[t][c] * yb[t][r];
        z[c][r] = s;
    }
}

// x[c] = v[c] / L_cc + sum_{t>c} M[c][t] v[t]
template <int NR>
__device__ __forceinline__ void tri_bwd(const double (*Dg)[kSD + 1], const double (*v)[NR], double (*x)[NR], int nb,
                                        int nrhs) {
    const int c = threadIdx.x & 63, r = threadIdx.x >> 6;
    if (r < NR && r < nrhs && c < nb) {
        double s = v[c][r] / Dg[c][c];
        for (int t = c + 1; t < nb; ++t) s += Dg[c][t] * v[t][r];
        x[c][r] = s;
    }
}

// Forward update of rows [lo, hi): y[row] -= sum_{t < kb} L[row][k + t] zs[t].
// Rows in the triangle [tk, tk + tkb) go to yb (LDS), other own pivot rows
// in place, ancestor rows by atomic add.
template <int NR>
__device__ void fwd_update(const double *Lb, const int32_t *frows, int f, int p, int col0, int k, int kb, int lo,
                           int hi, const double (*zs)[NR], double (*part)[64][NR], double (*yb)[NR], int tk, int tkb,
                           double *y, int n, int nrhs) {
    const int tid = threadIdx.x, rl = tid & 63, g = tid >> 6;
    for (int r0 = lo; r0 < hi; r0 += 64) {
        const int row = r0 + rl;
        const int rowc = min(row, hi - 1);
        double acc[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] = 0.0;
        double l[16];
#pragma unroll
        for (int tt = 0; tt < 16; ++tt)
            l[tt] = Lb[static_cast<int64_t>(k + min(g * 16 + tt, kb - 1)) * f + rowc];
#pragma unroll
        for (int tt = 0; tt < 16; ++tt) {
            const double lv = masked(l[tt], row < hi && g * 16 + tt < kb);
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] += lv * zs[g * 16 + tt][r];
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) part[g][rl][r] = acc[r];
        __syncthreads();
        if (tid < 64 * NR) {
            const int rr = tid & 63, r = tid >> 6;
            const int i = r0 + rr;
            if (i < hi && r < nrhs) {
                double s = 0.0;
#pragma unroll
                for (int g2 = 0; g2 < 8; ++g2) s += part[g2][rr][r];
                if (i >= tk && i < tk + tkb) {
                    yb[i - tk][r] = y[col0 + i + static_cast<int64_t>(r) * n] - s;
                } else if (i < p) {
                    y[col0 + i + static_cast<int64_t>(r) * n] -= s;
                } else {
                    unsafeAtomicAdd(&y[static_cast<int64_t>(frows[i]) + static_cast<int64_t>(r) * n], -s);
                }
            }
        }
        __syncthreads();
    }
}

template <int NR>
struct FwdLds {
    static_assert(NR * 64 <= kSThreads, "one wave per right-hand side in the triangles");
    double Dg[kSD][kSD + 1];
    double zs[kSB][NR];
    double yb[kSB][NR];
    double part[8][64][NR];
};

template <int NR>
__device__ void fwd_item(const DevPtrs &d, const SolveItem &it, double *__restrict__ y, int32_t n, int32_t nrhs,
                         FwdLds<NR> &S) {
    auto &Dg = S.Dg;
    auto &zs = S.zs;
    auto &yb = S.yb;
    auto &part = S.part;
    const FrontDev F = d.fronts[it.s];
    const int f = F.f, p = F.p;
    const double *Lb = d.L + F.lofs;
    const int32_t *frows = d.rows + F.rofs;
    const int tid = threadIdx.x, rl = tid & 63, g = tid >> 6;
    TriRegs R;
    if (it.tkb > 0) tri_prefetch(R, Lb, f, it.tk, it.tkb, rl, g);
    for (int idx = tid; idx < kSB * NR; idx += kSThreads) {
        const int t = idx % kSB, r = idx / kSB;
        zs[t][r] = (t < it.kb && r < nrhs) ? y[F.col0 + it.k + t + static_cast<int64_t>(r) * n] : 0.0;
        if (it.tkb > 0 && it.kb == 0) yb[t][r] = (t < it.tkb && r < nrhs) ? y[F.col0 + it.tk + t + static_cast<int64_t>(r) * n] : 0.0;
    }
    __syncthreads();
    if (it.kb > 0)
        fwd_update<NR>(Lb, frows, f, p, F.col0, it.k, it.kb, it.a0, it.a1, zs, part, yb, it.tk, it.tkb, y, n, nrhs);
    if (it.tkb <= 0) return;
    // the triangle: z_a = inv(L_a) y_a; y_b -= L_ba z_a; z_b = inv(L_b) y_b
    const int na = min(kSD, it.tkb), nb = it.tkb - kSD;
    put_block(Dg, R.a, rl, g);
    __syncthreads();
    tri_fwd<NR>(Dg, yb, zs, na, nrhs);
    __syncthreads();
    if (nb > 0) {
        double acc[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            acc[r] = 0.0;
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[r] += R.b[e] * zs[g + 8 * e][r];
            part[g][rl][r] = acc[r];
        }
        put_block(Dg, R.c, rl, g);
        __syncthreads();
        if (tid < 64 * NR) {
            const int rr = tid & 63, r = tid >> 6;
            if (rr < nb && r < nrhs) {
                double s = 0.0;
#pragma unroll
                for (int g2 = 0; g2 < 8; ++g2) s += part[g2][rr][r];
                yb[kSD + rr][r] -= s;
            }
        }
        __syncthreads();
        tri_fwd<NR>(Dg, yb + kSD, zs + kSD, nb, nrhs);
        __syncthreads();
    }
    // zs now holds z of the whole block (entries past tkb are still zero)
    for (int idx = tid; idx < it.tkb * nrhs; idx += kSThreads) {
        const int t = idx % it.tkb, r = idx / it.tkb;
        y[F.col0 + it.tk + t + static_cast<int64_t>(r) * n] = zs[t][r];
    }
    if (it.flags & 2)  // the rows below the last block, by this workgroup
        fwd_update<NR>(Lb, frows, f, p, F.col0, it.tk, it.tkb, it.tk + it.tkb, f, zs, part, yb, -1, 0, y, n, nrhs);
}

template <int NR>
__global__ __launch_bounds__(kSThreads) void k_fwd_step(DevPtrs d, const SolveItem *__restrict__ items,
                                                       double *__restrict__ y, int32_t n, int32_t nrhs) {
    __shared__ FwdLds<NR> S;
    fwd_item<NR>(d, items[blockIdx.x], y, n, nrhs, S);
}

// Backward update of columns [c0, c1): y[col0 + c] -= sum_{i in [r0, r1)}
// L[i][c] x_i, x_i = y[col0 + i] for own pivot rows, y[rows[i]] else.
// Lanes: 16 rows x 4 columns; wave g owns columns 8 g + {cq, cq + 4} of
// each 64-column tile. Columns of the triangle go to v (LDS).
template <int NR>
__device__ void bwd_update(const double *Lb, const int32_t *frows, int f, int p, int col0, int r0, int r1, int c0,
                           int c1, bool atomic, double (*xs)[NR], double (*v)[NR], int tk, int tkb,
                           double *y, int n, int nrhs) {
    const int tid = threadIdx.x;
    const int lane = tid & 63, g = tid >> 6, rq = lane & 15, cq = lane >> 4;
    const int nr = r1 - r0;
    for (int idx = tid; idx < kSeg * NR; idx += kSThreads) {
        const int i = idx % kSeg, r = idx / kSeg;
        double xv = 0.0;
        if (i < nr && r < nrhs) {
            const int row = r0 + i;
            const int64_t gi = row < p ? static_cast<int64_t>(col0 + row) : static_cast<int64_t>(frows[row]);
            xv = y[gi + static_cast<int64_t>(r) * n];
        }
        xs[i][r] = xv;
    }
    __syncthreads();
    for (int cb = c0; cb < c1; cb += 64) {
        const int ca = cb + 8 * g + cq, cbb = ca + 4;
        const bool va = ca < c1, vb = cbb < c1;
        const double *La = Lb + static_cast<int64_t>(va ? ca : c0) * f + r0;
        const double *Lbb = Lb + static_cast<int64_t>(vb ? cbb : c0) * f + r0;
        double acc0[NR], acc1[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) acc0[r] = acc1[r] = 0.0;
        for (int i0 = 0; i0 < nr; i0 += 128) {
            double la[8], lb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = min(i0 + 16 * u + rq, nr - 1);
                la[u] = La[i];
                lb[u] = Lbb[i];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bool in = i0 + 16 * u + rq < nr;
                la[u] = masked(la[u], va && in);
                lb[u] = masked(lb[u], vb && in);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = min(i0 + 16 * u + rq, kSeg - 1);
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    acc0[r] += la[u] * xs[i][r];
                    acc1[r] += lb[u] * xs[i][r];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
#pragma unroll
            for (int o = 8; o >= 1; o >>= 1) {
                acc0[r] += __shfl_xor(acc0[r], o);
                acc1[r] += __shfl_xor(acc1[r], o);
            }
        }
        if (rq == 0) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int c = h ? cbb : ca;
                if (!(h ? vb : va)) continue;
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    if (r >= nrhs) break;
                    const double s = h ? acc1[r] : acc0[r];
                    double *dst = &y[col0 + c + static_cast<int64_t>(r) * n];
                    if (c >= tk && c < tk + tkb) v[c - tk][r] = *dst - s;
                    else if (atomic) unsafeAtomicAdd(dst, -s);
                    else *dst -= s;
                }
            }
        }
    }
    __syncthreads();
}

template <int NR>
struct BwdLds {
    double Dg[kSD][kSD + 1];
    double xs[kSeg][NR];
    double v[kSB][NR];
    double xo[kSB][NR];
};

template <int NR>
__device__ void bwd_item(const DevPtrs &d, const SolveItem &it, double *__restrict__ y, int32_t n, int32_t nrhs,
                         BwdLds<NR> &S) {
    auto &Dg = S.Dg;
    auto &xs = S.xs;
    auto &v = S.v;
    auto &xo = S.xo;
    const FrontDev F = d.fronts[it.s];
    const int f = F.f, p = F.p;
    const double *Lb = d.L + F.lofs;
    const int32_t *frows = d.rows + F.rofs;
    const int tid = threadIdx.x, rl = tid & 63, g = tid >> 6;
    TriRegs R;
    if (it.tkb > 0) {
        tri_prefetch(R, Lb, f, it.tk, it.tkb, rl, g);
        if (it.kb == 0)
            for (int idx = tid; idx < kSB * NR; idx += kSThreads) {
                const int t = idx % kSB, r = idx / kSB;
                v[t][r] = (t < it.tkb && r < nrhs) ? y[F.col0 + it.tk + t + static_cast<int64_t>(r) * n] : 0.0;
            }
    }
    if (it.kb > 0)
        bwd_update<NR>(Lb, frows, f, p, F.col0, it.k, it.k + it.kb, it.a0, it.a1, (it.flags & 1) != 0, xs, v, it.tk,
                       it.tkb, y, n, nrhs);
    if (it.tkb <= 0) return;
    // the triangle, second diagonal block first:
    // x_b = inv(L_b)^T v_b; v_a -= L_ba^T x_b; x_a = inv(L_a)^T v_a
    const int na = min(kSD, it.tkb), nb = it.tkb - kSD;
    if (nb > 0) {
        put_block(Dg, R.c, rl, g);
        __syncthreads();
        tri_bwd<NR>(Dg, v + kSD, xo + kSD, nb, nrhs);
        __syncthreads();
        put_block(Dg, R.b, rl, g);  // Dg[t][c] = L_ba[t][c]
        __syncthreads();
        {
            const int c = tid & 63, r = tid >> 6;
            if (r < NR && r < nrhs && c < na) {
                double s = 0.0;
                for (int t = 0; t < nb; ++t) s += Dg[t][c] * xo[kSD + t][r];
                v[c][r] -= s;
            }
        }
        __syncthreads();
    }
    put_block(Dg, R.a, rl, g);
    __syncthreads();
    tri_bwd<NR>(Dg, v, xo, na, nrhs);
    __syncthreads();
    for (int idx = tid; idx < it.tkb * nrhs; idx += kSThreads) {
        const int t = idx % it.tkb, r = idx / it.tkb;
        y[F.col0 + it.tk + t + static_cast<int64_t>(r) * n] = xo[t][r];
    }
}

template <int NR>
__global__ __launch_bounds__(kSThreads) void k_bwd_step(DevPtrs d, const SolveItem *__restrict__ items,
                                                       double *__restrict__ y, int32_t n, int32_t nrhs) {
    __shared__ BwdLds<NR> S;
    bwd_item<NR>(d, items[blockIdx.x], y, n, nrhs, S);
}

// ---------------------------------------------------------------------------
// One launch per sweep: a workgroup per item, each taking a ticket (its
// item, in sweep order) when it starts; before an item of step t runs,
// every item of step t - 1 has finished (its counter reached the step's
// item count): the ordering the per-step launches get from kernel
// boundaries, without the launches. Tickets go out in increasing order to
// running workgroups and an item waits only on items with smaller tickets,
// so the waits end however the workgroups are dispatched. There is no loop
// around the item: every wave meets the same barriers in the same order.
// Writes to y are released with a device-scope fence before the counter
// moves and acquired after the wait. A wait that runs out (never expected)
// sets err and lets every later wait through, so the launch always ends.
constexpr int64_t kSweepSpin = int64_t{1} << 18;

__device__ __forceinline__ void sweep_wait(int32_t *cnt, int32_t target, int32_t *err) {
    for (int64_t spin = 0; __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; ++spin) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (spin > kSweepSpin) {
            atomicExch(err, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
}

template <int NR, bool FWD>
__global__ __launch_bounds__(kSThreads) void k_sweep(DevPtrs d, const SolveItem *__restrict__ items,
                                                    const int32_t *__restrict__ step_of,
                                                    const int32_t *__restrict__ step_cnt, int32_t *ctr, int32_t *err,
                                                    double *__restrict__ y, int32_t n, int32_t nrhs) {
    __shared__ typename std::conditional<FWD, FwdLds<NR>, BwdLds<NR>>::type S;
    __shared__ int32_t s_item;
    if (threadIdx.x == 0) {
        const int32_t i = atomicAdd(&ctr[0], 1);  // < the grid size, the item count
        const int32_t t = step_of[i];
        if (t > 0) sweep_wait(&ctr[t], step_cnt[t - 1], err);  // counter of step t - 1
        s_item = i;
    }
    __syncthreads();
    const int32_t i = __builtin_amdgcn_readfirstlane(s_item);
    const SolveItem it = items[i];
    if constexpr (FWD) fwd_item<NR>(d, it, y, n, nrhs, S);
    else bwd_item<NR>(d, it, y, n, nrhs, S);
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(&ctr[1 + step_of[i]], 1);
    }
}

__global__ void k_zero_i32(int32_t *p, int32_t n) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}


// ---------------------------------------------------------------------------
// Sweeps of an LDL^T factor (unit L with D on the stored diagonal; the D^-1
// step runs between them): steps of kLS pivots (a large front's dense
// triangle costs a launch per kLS pivots), each solved inside its triangle
// item as up to four 64-pivot sub-blocks whose unit triangles go through
// LDS. Items: forward triangle {k, kb: the block solved in the step before;
// tk, tkb: this block}, forward chunk {rows [a0, a1), update by [k, k+kb)},
// backward chunk {columns [a0, a1), contributions of rows [k, k+kb); flags
// bit 0: atomic}, backward triangle {k, kb: block solved before (or the
// rows below the pivots); tk, tkb}.
constexpr int kLB = 64;    // sub-block: a unit triangle in LDS, substituted by one wave
constexpr int kLS = 256;   // pivots per step
constexpr int kLThreads = 256;
constexpr int kLSeg = 512;  // rows of a backward tail item
constexpr int32_t kItemTri = 2;
constexpr int32_t kItemFused = 4;  // forward: the triangle, then the rows [a0, a1) below it in the same workgroup

__device__ __forceinline__ double lwave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// sum_{c < kb} L(i, k + c) xs[c] (kb <= kLS) for row i = r0 + (tid & 63)
// (< r1), wave w taking the columns [w q, (w + 1) q), q = ceil(kb / 4),
// in order; the total ((w0 + w1) + (w2 + w3)) lands in part[0][row]. The
// loads go out 16 columns at a time (clamped addresses, masked values), so
// a workgroup waits on a column group's latency, not each column's.
__device__ __forceinline__ void rows_times_block(const double *L, int f, int k, int kb, int r0, int r1,
                                                 const double *xs, double (*part)[kLB]) {
    constexpr int kG = 16;
    const int row = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = min(r0 + row, r1 - 1);
    const int q = (kb + 3) >> 2;
    const int cb = w * q, ce = min(kb, cb + q);
    double acc = 0.0;
    if (r0 + row < r1 && ce > cb) {
        const double *Lc = L + static_cast<int64_t>(k) * f + i;
        for (int c = cb; c < ce; c += kG) {
            double lv[kG];
#pragma unroll
            for (int j = 0; j < kG; ++j) lv[j] = Lc[static_cast<int64_t>(min(c + j, ce - 1)) * f];
#pragma unroll
            for (int j = 0; j < kG; ++j)
                if (c + j < ce) acc += lv[j] * xs[c + j];
        }
    }
    part[w][row] = acc;
    __syncthreads();
    if (w == 0) part[0][row] = (part[0][row] + part[1][row]) + (part[2][row] + part[3][row]);
    __syncthreads();
}

// the unit triangle of sub-block [c0, c0 + nc): T[c][l] = L(c0 + l, c0 + c), l > c
__device__ __forceinline__ void load_unit_triangle(const double *L, int f, int c0, int nc, double (*T)[kLB + 1]) {
    for (int e = threadIdx.x; e < kLB * kLB; e += kLThreads) {
        const int c = e >> 6, l = e & 63;
        T[c][l] = (c < nc && l < nc && l > c) ? L[static_cast<int64_t>(c0 + c) * f + c0 + l] : 0.0;
    }
}

__global__ __launch_bounds__(kLThreads) void k_ldlt_fwd_step(LdltDev d, const SolveItem *__restrict__ items,
                                                            double *__restrict__ y, int32_t n, int32_t nrhs) {
    __shared__ double xs[kLS];  // the block solved in the step before
    __shared__ double xc[kLS];  // this item's block as it is solved
    __shared__ double part[4][kLB];
    __shared__ double part2[4][kLB];
    __shared__ double T[kLB][kLB + 1];
    const SolveItem it = items[blockIdx.x];
    const FrontLd F = d.fronts[it.s];
    const int f = F.f, p = F.pstore;
    const double *L = d.fac + F.lofs;
    const int32_t *rows = d.rows + F.rofs;
    const int tid = threadIdx.x;
    const bool tri = (it.flags & kItemTri) != 0;
    for (int r = 0; r < nrhs; ++r) {
        double *yr = y + static_cast<int64_t>(r) * n;
        for (int t = tid; t < it.kb; t += kLThreads) xs[t] = yr[rows[it.k + t]];
        __syncthreads();
        if (tri) {
            const int nsub = (it.tkb + kLB - 1) / kLB;
            for (int sb = 0; sb < nsub; ++sb) {
                const int c0 = it.tk + sb * kLB, nc = min(kLB, it.tkb - sb * kLB);
                load_unit_triangle(L, f, c0, nc, T);
                // the sub-block's rows: by the block before (when the item
                // carries it), then by this block's solved sub-blocks
                if (it.kb > 0) rows_times_block(L, f, it.k, it.kb, c0, c0 + nc, xs, part);
                if (sb > 0) rows_times_block(L, f, it.tk, sb * kLB, c0, c0 + nc, xc, part2);
                else __syncthreads();  // the triangle in LDS
                if (tid < 64) {
                    const int l = tid;
                    double yv = l < nc ? yr[rows[c0 + l]] : 0.0;
                    if (it.kb > 0) yv -= part[0][l];
                    if (sb > 0) yv -= part2[0][l];
                    for (int c = 0; c < nc; ++c) {
                        const double xv = __shfl(yv, c, 64);
                        if (l > c) yv -= T[c][l] * xv;
                    }
                    if (l < nc) {
                        yr[rows[c0 + l]] = yv;
                        xc[sb * kLB + l] = yv;
                    }
                }
                __syncthreads();
            }
            if (it.flags & kItemFused) {
                // the rows below, with the block just solved (the same sums
                // as the separate row items of the next step)
                for (int rb = it.a0; rb < it.a1; rb += kLB) {
                    const int re = min(rb + kLB, it.a1);
                    rows_times_block(L, f, it.tk, it.tkb, rb, re, xc, part);
                    if (tid < 64 && rb + tid < re) atomicAdd(&yr[rows[rb + tid]], -part[0][tid]);
                    __syncthreads();
                }
            }
        } else {
            rows_times_block(L, f, it.k, it.kb, it.a0, it.a1, xs, part);
            if (tid < 64) {
                const int i = it.a0 + tid;
                if (i < it.a1) {
                    if (i >= p) atomicAdd(&yr[rows[i]], -part[0][tid]);
                    else yr[rows[i]] -= part[0][tid];
                }
            }
        }
        __syncthreads();
    }
}

// Backward contributions of rows [k, k + kb) (values xs) to the columns
// [c0, c1): each wave takes 16 columns at a time, its lanes the rows t =
// lane + 64 m in order, the 16 columns' loads of one m in flight together;
// out(c, sum) from lane 0.
template <class Out>
__device__ __forceinline__ void cols_times_rows(const double *L, int f, int k, int kb, int c0, int c1, const double *xs,
                                                Out &&out) {
    constexpr int kG = 16;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int cb = c0 + w * kG; cb < c1; cb += 4 * kG) {
        double acc[kG];
#pragma unroll
        for (int j = 0; j < kG; ++j) acc[j] = 0.0;
        for (int t0 = 0; t0 < kb; t0 += 64) {
            const int t = t0 + lane;
            const bool tin = t < kb;
            const double xv = tin ? xs[t] : 0.0;
            const int tc = min(t, kb - 1);
            double lv[kG];
#pragma unroll
            for (int j = 0; j < kG; ++j) lv[j] = L[static_cast<int64_t>(min(cb + j, c1 - 1)) * f + k + tc];
#pragma unroll
            for (int j = 0; j < kG; ++j)
                if (tin) acc[j] += lv[j] * xv;
        }
#pragma unroll
        for (int j = 0; j < kG; ++j) {
            const double v = lwave_sum(acc[j]);
            if (lane == 0 && cb + j < c1) out(cb + j, v);
        }
    }
}

__global__ __launch_bounds__(kLThreads) void k_ldlt_bwd_step(LdltDev d, const SolveItem *__restrict__ items,
                                                            double *__restrict__ y, int32_t n, int32_t nrhs) {
    __shared__ double xs[kLSeg];
    __shared__ double xc[kLS];
    __shared__ double upd[kLS];
    __shared__ double upd2[kLB];
    __shared__ double T[kLB][kLB + 1];
    const SolveItem it = items[blockIdx.x];
    const FrontLd F = d.fronts[it.s];
    const int f = F.f;
    const double *L = d.fac + F.lofs;
    const int32_t *rows = d.rows + F.rofs;
    const int tid = threadIdx.x;
    const bool tri = (it.flags & kItemTri) != 0;
    for (int r = 0; r < nrhs; ++r) {
        double *yr = y + static_cast<int64_t>(r) * n;
        for (int t = tid; t < it.kb; t += kLThreads) xs[t] = yr[rows[it.k + t]];
        __syncthreads();
        if (!tri) {
            cols_times_rows(L, f, it.k, it.kb, it.a0, it.a1, xs, [&](int c, double acc) {
                if (it.flags & 1) atomicAdd(&yr[rows[c]], -acc);
                else yr[rows[c]] -= acc;
            });
            __syncthreads();
            continue;
        }
        // the block's columns: contributions of the rows solved before
        // (when the item carries them)
        if (it.kb > 0)
            cols_times_rows(L, f, it.k, it.kb, it.tk, it.tk + it.tkb, xs,
                            [&](int c, double acc) { upd[c - it.tk] = acc; });
        else
            for (int c = tid; c < it.tkb; c += kLThreads) upd[c] = 0.0;
        __syncthreads();
        const int nsub = (it.tkb + kLB - 1) / kLB;
        for (int sb = nsub - 1; sb >= 0; --sb) {
            const int c0 = it.tk + sb * kLB, nc = min(kLB, it.tkb - sb * kLB);
            const int e0 = c0 + nc, e1 = it.tk + it.tkb;  // this block's sub-blocks solved so far
            load_unit_triangle(L, f, c0, nc, T);
            cols_times_rows(L, f, e0, e1 - e0, c0, c0 + nc, xc + (e0 - it.tk),
                            [&](int c, double acc) { upd2[c - c0] = acc; });
            __syncthreads();
            if (tid < 64) {
                const int l = tid;
                double yv = l < nc ? (yr[rows[c0 + l]] - upd[sb * kLB + l]) - (e1 > e0 ? upd2[l] : 0.0) : 0.0;
                for (int c = nc - 1; c >= 0; --c) {
                    const double xv = __shfl(yv, c, 64);
                    if (l < c) yv -= T[l][c] * xv;
                }
                if (l < nc) {
                    yr[rows[c0 + l]] = yv;
                    xc[sb * kLB + l] = yv;
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace

void launch_ldlt_fwd_step(hipStream_t st, const LdltDev &d, const SolveItem *items, int64_t count, double *y,
                          int32_t n, int32_t nrhs) {
    if (count <= 0) return;
    for (int64_t o = 0; o < count; o += kMaxLaunchGroups)
        k_ldlt_fwd_step<<<dim3(static_cast<unsigned>(std::min(kMaxLaunchGroups, count - o))), kLThreads, 0, st>>>(
            d, items + o, y, n, nrhs);
}

void launch_ldlt_bwd_step(hipStream_t st, const LdltDev &d, const SolveItem *items, int64_t count, double *y,
                          int32_t n, int32_t nrhs) {
    if (count <= 0) return;
    for (int64_t o = 0; o < count; o += kMaxLaunchGroups)
        k_ldlt_bwd_step<<<dim3(static_cast<unsigned>(std::min(kMaxLaunchGroups, count - o))), kLThreads, 0, st>>>(
            d, items + o, y, n, nrhs);
}

void launch_sweep(hipStream_t st, const DevPtrs &d, bool forward, const SolveItem *items, const int32_t *step_of,
                  const int32_t *step_cnt, int32_t nitems, int32_t nsteps, int32_t *ctr, int32_t *err, double *y,
                  int32_t n, int32_t nrhs) {
    if (nitems <= 0) return;
    k_zero_i32<<<dim3(std::max(1, std::min(64, (nsteps + 1 + 255) / 256))), 256, 0, st>>>(ctr, nsteps + 1);
    const dim3 grid(static_cast<unsigned>(nitems));  // one workgroup per item
    if (forward) {
        if (nrhs == 1) k_sweep<1, true><<<grid, kSThreads, 0, st>>>(d, items, step_of, step_cnt, ctr, err, y, n, nrhs);
        else if (nrhs <= 4) k_sweep<4, true><<<grid, kSThreads, 0, st>>>(d, items, step_of, step_cnt, ctr, err, y, n, nrhs);
        else k_sweep<8, true><<<grid, kSThreads, 0, st>>>(d, items, step_of, step_cnt, ctr, err, y, n, nrhs);
    } else {
        if (nrhs == 1) k_sweep<1, false><<<grid, kSThreads, 0, st>>>(d, items, step_of, step_cnt, ctr, err, y, n, nrhs);
        else if (nrhs <= 4) k_sweep<4, false><<<grid, kSThreads, 0, st>>>(d, items, step_of, step_cnt, ctr, err, y, n, nrhs);
        else k_sweep<8, false><<<grid, kSThreads, 0, st>>>(d, items, step_of, step_cnt, ctr, err, y, n, nrhs);
    }
}

void launch_fwd_step(hipStream_t st, const DevPtrs &d, const SolveItem *items, int64_t count, double *y, int32_t n,
                     int32_t nrhs) {
    for (int64_t o = 0; o < count; o += kMaxLaunchGroups) {
        const dim3 grid(static_cast<unsigned>(std::min(kMaxLaunchGroups, count - o)));
        if (nrhs == 1) k_fwd_step<1><<<grid, kSThreads, 0, st>>>(d, items + o, y, n, nrhs);
        else if (nrhs <= 4) k_fwd_step<4><<<grid, kSThreads, 0, st>>>(d, items + o, y, n, nrhs);
        else k_fwd_step<8><<<grid, kSThreads, 0, st>>>(d, items + o, y, n, nrhs);
    }
}

void launch_bwd_step(hipStream_t st, const DevPtrs &d, const SolveItem *items, int64_t count, double *y, int32_t n,
                     int32_t nrhs) {
    for (int64_t o = 0; o < count; o += kMaxLaunchGroups) {
        const dim3 grid(static_cast<unsigned>(std::min(kMaxLaunchGroups, count - o)));
        if (nrhs == 1) k_bwd_step<1><<<grid, kSThreads, 0, st>>>(d, items + o, y, n, nrhs);
        else if (nrhs <= 4) k_bwd_step<4><<<grid, kSThreads, 0, st>>>(d, items + o, y, n, nrhs);
        else k_bwd_step<8><<<grid, kSThreads, 0, st>>>(d, items + o, y, n, nrhs);
    }
}

}  // namespace smfgpu
