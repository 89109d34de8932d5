// HIP kernels of the indefinite (LDL^T, threshold 1x1 / 2x2 pivoting)
// numeric phase. Compiled with -ffp-contract=off: every multiply and add
// below rounds separately, as the reference's ldlt_kernel.cpp
// (solver/CMakeLists.txt:632-642), so the unblocked front kernel
// reproduces ldlt_factor_front's pivot sequence and values bit for bit.
#include "smfgpu_ldlt.hpp"
#include "mfma_tile.hpp"

#include <cfloat>
#include <type_traits>

namespace smfgpu {

constexpr int kLT = 256;            // threads per workgroup
constexpr double kWeakPivotRatio = 1e-6;  // ldlt_kernel.hpp:100

// ---------------------------------------------------------------------------
// Workgroup reductions (exact: max and first argmax).
template <int NT>
struct RedScratchT {
    double v[NT];
    int i[NT];
};
using RedScratch = RedScratchT<kLT>;

// (max, first index of the max) over the thread's candidates, reduced
// across the workgroup: ties keep the smaller index, as a sequential scan
// with a strict '>' does.
// One wavefront (every lane active): within each row of 16 lanes the
// partners of four DPP steps (quad swaps of 1 and 2, half-row mirror, row
// mirror) bring every lane the row's result; the four rows' results are
// then read from lanes 0, 16, 32, 48. Max and first-argmax do not depend on
// the order of combination.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL>(static_cast<int>(b));
    const int hi = dpp_i<CTRL>(static_cast<int>(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ double lane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;
__device__ __forceinline__ bool better_first(double b, int ib, double a, int ia) {
    return b > a || (b == a && ib >= 0 && (ia < 0 || ib < ia));
}
template <int CTRL>
__device__ __forceinline__ void maxfirst_step(double &v, int &idx) {
    const double b = dpp_d<CTRL>(v);
    const int ib = dpp_i<CTRL>(idx);
    if (better_first(b, ib, v, idx)) {
        v = b;
        idx = ib;
    }
}
template <int CTRL>
__device__ __forceinline__ void max_step(double &v) {
    const double b = dpp_d<CTRL>(v);
    if (b > v) v = b;
}

template <int NT>
__device__ __forceinline__ void red_maxfirst(RedScratchT<NT> &rs, double v, int idx, double &outv, int &outi) {
    if constexpr (NT == 64) {
        maxfirst_step<kDppXor1>(v, idx);
        maxfirst_step<kDppXor2>(v, idx);
        maxfirst_step<kDppHalfMirror>(v, idx);
        maxfirst_step<kDppMirror>(v, idx);
        double bv = lane_d(v, 0);
        int bi = __builtin_amdgcn_readlane(idx, 0);
        for (int r = 16; r < 64; r += 16) {
            const double c = lane_d(v, r);
            const int ic = __builtin_amdgcn_readlane(idx, r);
            if (better_first(c, ic, bv, bi)) {
                bv = c;
                bi = ic;
            }
        }
        outv = bv;
        outi = bi;
        return;
    }
    const int tid = threadIdx.x;
    rs.v[tid] = v;
    rs.i[tid] = idx;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (tid < w) {
            const double a = rs.v[tid], b = rs.v[tid + w];
            const int ia = rs.i[tid], ib = rs.i[tid + w];
            if (b > a || (b == a && ib >= 0 && (ia < 0 || ib < ia))) {
                rs.v[tid] = b;
                rs.i[tid] = ib;
            }
        }
        __syncthreads();
    }
    outv = rs.v[0];
    outi = rs.i[0];
    __syncthreads();
}

template <int NT>
__device__ __forceinline__ double red_max(RedScratchT<NT> &rs, double v) {
    if constexpr (NT == 64) {
        max_step<kDppXor1>(v);
        max_step<kDppXor2>(v);
        max_step<kDppHalfMirror>(v);
        max_step<kDppMirror>(v);
        double m = lane_d(v, 0);
        for (int r = 16; r < 64; r += 16) {
            const double c = lane_d(v, r);
            if (c > m) m = c;
        }
        return m;
    }
    const int tid = threadIdx.x;
    rs.v[tid] = v;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (tid < w && rs.v[tid + w] > rs.v[tid]) rs.v[tid] = rs.v[tid + w];
        __syncthreads();
    }
    const double out = rs.v[0];
    __syncthreads();
    return out;
}

// View of one front: element (i, j), i >= j, at F[j * ld + i].
struct FV {
    double *F;
    int ld, f;
    __device__ __forceinline__ double &at(int i, int j) const { return F[static_cast<int64_t>(j) * ld + i]; }
    __device__ __forceinline__ double sym(int i, int j) const { return i >= j ? at(i, j) : at(j, i); }
};

// Column sweeps keep kU independent loads in flight per thread: the loads
// of a group are issued from clamped (in-range) addresses before any use,
// and only the in-range elements are used or stored. Updates read and
// write different columns, which the compiler cannot tell apart, so
// without the explicit grouping every element waits out a full memory
// latency.
constexpr int kU = 4;

// f(i, v) for v = ld(i), i = lo, lo + st, ... < hi, in increasing i per
// thread
template <class Ld, class Fn>
__device__ __forceinline__ void sweep(int lo, int hi, int st, Ld &&ld, Fn &&fn) {
    for (int i0 = lo; i0 < hi; i0 += kU * st) {
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) v[u] = ld(min(i0 + u * st, hi - 1));
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (i0 + u * st < hi) fn(i0 + u * st, v[u]);
    }
}

// c[i] = c[i] - a[i] * l (multiply, then subtract), i = lo, lo + st, ... < hi;
// c and a are different columns
__device__ __forceinline__ void col_axpy(double *c, const double *a, double l, int lo, int hi, int st) {
    for (int i0 = lo; i0 < hi; i0 += kU * st) {
        double cv[kU], av[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = min(i0 + u * st, hi - 1);
            cv[u] = c[i];
            av[u] = a[i];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (i0 + u * st < hi) c[i0 + u * st] = __dsub_rn(cv[u], __dmul_rn(av[u], l));
    }
}

// c[i] = c[i] - (a0[i] * l0 + a1[i] * l1)
__device__ __forceinline__ void col_axpy2(double *c, const double *a0, const double *a1, double l0, double l1, int lo,
                                          int hi, int st) {
    for (int i0 = lo; i0 < hi; i0 += kU * st) {
        double cv[kU], x0[kU], x1[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = min(i0 + u * st, hi - 1);
            cv[u] = c[i];
            x0[u] = a0[i];
            x1[u] = a1[i];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (i0 + u * st < hi)
                c[i0 + u * st] = __dsub_rn(cv[u], __dadd_rn(__dmul_rn(x0[u], l0), __dmul_rn(x1[u], l1)));
    }
}

// dst[i] = src[i] (different arrays)
__device__ __forceinline__ void col_copy(double *dst, const double *src, int lo, int hi, int st) {
    for (int i0 = lo; i0 < hi; i0 += kU * st) {
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) v[u] = src[min(i0 + u * st, hi - 1)];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (i0 + u * st < hi) dst[i0 + u * st] = v[u];
    }
}

// exchange a[i] and b[i] (different columns, or a row and a column that
// share no element over the range)
template <class PA, class PB>
__device__ __forceinline__ void swap_range(PA &&pa, PB &&pb, int lo, int hi, int st) {
    for (int i0 = lo; i0 < hi; i0 += kU * st) {
        double x[kU], y[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = min(i0 + u * st, hi - 1);
            x[u] = *pa(i);
            y[u] = *pb(i);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (i0 + u * st < hi) {
                *pa(i0 + u * st) = y[u];
                *pb(i0 + u * st) = x[u];
            }
    }
}

// min_abs: the front's in-order static pivot (static_pivot_in_order, 0 off);
// S: static pivoting of the launch (fronts pick theirs with
// front_static_value)
struct Prm {
    double u, eff_small, small_pivot;
    double min_abs;
    StaticPrm S;
};

// The pivoting parameters of front F: its static value (0: none) and, in
// the in-order mode, min_abs (ldlt_kernel.cpp:1131-1132 and
// multifrontal_serial.cpp:624-629: a fallback front keeps in-order only when
// static pivoting is on overall).
__device__ __forceinline__ double front_prm(const Prm &prm, const FrontLd &F, Prm &fp) {
    fp = prm;
    const double sp = front_static_value(prm.S, F.root, F.nd, F.psym);
    fp.min_abs = (prm.S.sp > 0.0 && (prm.S.flags & kStaticInOrder)) ? sp : 0.0;
    return sp;
}

// static_counts_as_zero (ldlt_kernel.cpp:325-328): a fallback front keeps a
// pivot raised from an exact zero as a zero (static_zero_stays_zero set).
__device__ __forceinline__ bool static_as_zero(int raised, const Prm &prm, const FrontLd &F) {
    if (raised == 0) return false;
    const bool fallback = prm.S.fallback > 0.0 && (F.root != 0 || F.nd > F.psym);
    const bool stays = fallback || (prm.S.flags & kStaticZeroStaysZero);
    return (prm.S.flags & kStaticAsZero) || (raised == 2 && stays);
}

// select_pivot (ldlt_kernel.cpp:127-275) with workgroup reductions. Every
// thread returns the same choice.
template <int NT>
__device__ void select_pivot_wg(const FV &A, int k, int ne, int limit, const Prm &prm, RedScratchT<NT> &rs, int &kind,
                                int &pa, int &pb) {
    const int tid = threadIdx.x;
    const int f = A.f;
    kind = 0;
    pa = pb = -1;
    if (prm.min_abs > 0.0) {
        // in order, no search: column k or nothing (then the static path)
        const double d0 = A.at(k, k);
        if (isfinite(d0) && fabs(d0) >= prm.min_abs) {
            kind = 1;
            pa = k;
        }
        return;
    }
    if (k < limit) {
        const double a0 = fabs(A.at(k, k));
        if (a0 >= prm.eff_small) {
            double m = 0.0;
            const double *ck = &A.at(0, k);
            sweep(k + 1 + tid, f, NT, [&](int i) { return ck[i]; }, [&](int, double v) { m = fmax(m, fabs(v)); });
            const double colmax = red_max(rs, m);
            if (!(prm.u * colmax > a0) && !(a0 < kWeakPivotRatio * colmax)) {
                kind = 1;
                pa = k;
                return;
            }
        }
    }
    for (int cand = k; cand < limit; ++cand) {
        const double a_cand = fabs(A.at(cand, cand));
        // largest entry over the eligible rows [k, ne) other than cand (first
        // index on ties), then the largest of the rest
        // One sweep of the column: each thread keeps its largest entry (first
        // index), the largest of its other entries, and the maximum over
        // rows [ne, f). The thread holding the overall first argmax then
        // contributes its second value to m2, every other thread its first.
        const int ext0 = max(ne, cand + 1);
        double bv = -1.0, b2 = 0.0, mel = 0.0;
        int bi = -1;
        const auto top2 = [&](int i, double x) {
            const double v = fabs(x);
            if (v > bv) {
                b2 = fmax(b2, bv);
                bv = v;
                bi = i;
            } else {
                b2 = fmax(b2, v);
            }
        };
        // rows [k, cand) of the candidate's row, then rows (cand, ne) and
        // [ext0, f) of its column: each thread still meets its indices in
        // increasing order
        const double *cc = &A.at(0, cand);
        const int k1 = k + tid;
        sweep(k1, min(cand, ne), NT, [&](int i) { return A.at(cand, i); }, top2);
        {
            // first index of this thread's sequence past cand
            const int c1 = cand + 1 + (((k1 - (cand + 1)) % NT) + NT) % NT;
            sweep(c1, ne, NT, [&](int i) { return cc[i]; }, top2);
            const int e1 = ext0 + (((k1 - ext0) % NT) + NT) % NT;
            sweep(e1, f, NT, [&](int i) { return cc[i]; }, [&](int, double x) { mel = fmax(mel, fabs(x)); });
        }
        double m1;
        int r_best;
        red_maxfirst(rs, bv, bi, m1, r_best);
        if (!(m1 > 0.0)) {
            m1 = 0.0;
            r_best = -1;
        }
        const double m2 = red_max(rs, (r_best >= 0 && bi == r_best) ? b2 : fmax(bv, 0.0));
        const double m_ext = red_max(rs, mel);
        bool weak = false;
        double weak_growth = 0.0;
        if (a_cand >= prm.eff_small && !(prm.u * m1 > a_cand)) {
            if (!(prm.u * m_ext > a_cand)) {
                const double colmax = fmax(m1, m_ext);
                if (!(a_cand < kWeakPivotRatio * colmax)) {
                    kind = 1;
                    pa = cand;
                    return;
                }
                weak = true;
                weak_growth = colmax / a_cand;
            }
        }
        if (r_best < k || r_best >= limit || m1 <= 0.0) {
            if (weak) {
                kind = 1;
                pa = cand;
                return;
            }
            continue;
        }
        // partner column: max excluding r_best, and excluding cand too
        double o2 = 0.0, o1 = 0.0;
        const auto part = [&](int i, double x) {
            const double v = fabs(x);
            o2 = fmax(o2, v);
            if (i != cand) o1 = fmax(o1, v);
        };
        const double *cr = &A.at(0, r_best);
        sweep(k + tid, r_best, NT, [&](int i) { return A.at(r_best, i); }, part);
        sweep(r_best + 1 + (((k + tid - (r_best + 1)) % NT) + NT) % NT, f, NT, [&](int i) { return cr[i]; }, part);
        const double om2 = red_max(rs, o2);
        const double off1 = red_max(rs, o1);
        // two_by_two_ok (ldlt_kernel.cpp:97-122)
        bool ok2 = false;
        double growth2 = 0.0;
        if (k + 1 < limit) {
            const double d00 = A.at(cand, cand);
            const double d10 = A.sym(r_best, cand);
            const double d11 = A.at(r_best, r_best);
            double s2 = fabs(d00);
            if (s2 < fabs(d10)) s2 = fabs(d10);
            if (s2 < fabs(d11)) s2 = fabs(d11);
            if (s2 >= prm.eff_small && isfinite(s2)) {
                const double a = d00 / s2, b = d10 / s2, c = d11 / s2;
                const double det_s = __dsub_rn(__dmul_rn(a, c), __dmul_rn(b, b));
                if (fabs(det_s) > prm.small_pivot) {
                    const double off0 = fmax(m2, m_ext);
                    ok2 = true;
                    if (off0 > 0.0 || off1 > 0.0) {
                        const double bound0 = __dadd_rn(__dmul_rn(fabs(c), off0), __dmul_rn(fabs(b), off1)) / s2;
                        const double bound1 = __dadd_rn(__dmul_rn(fabs(b), off0), __dmul_rn(fabs(a), off1)) / s2;
                        const double bm = bound0 > bound1 ? bound0 : bound1;
                        if (prm.u * bm > fabs(det_s)) ok2 = false;
                        else growth2 = bm / fabs(det_s);
                    }
                }
            }
        }
        if (ok2 && (!weak || growth2 < weak_growth)) {
            kind = 2;
            pa = cand;
            pb = r_best;
            return;
        }
        if (weak) {
            kind = 1;
            pa = cand;
            return;
        }
        const double a_r = fabs(A.at(r_best, r_best));
        if (a_r >= prm.eff_small && (om2 == 0.0 || a_r >= prm.u * om2)) {
            kind = 1;
            pa = r_best;
            return;
        }
    }
}

// Partner rows of a stalled first-pass panel (ldlt_kernel.cpp:1052-1075):
// for each failed column pos of [k, pl) in order, the first row of [pl, ne)
// holding the column's largest entry, kept when that entry exceeds the
// column's largest entry inside the panel and the row is no earlier
// column's partner. (column, partner) row ids go to order (at most 64
// pairs, nb / 2 when smaller); returns the number of ids written. Every
// thread returns the same count.
template <int NT>
__device__ int partner_pairs_wg(const FV &A, int k, int pl, int ne, int nb, const int32_t *rows, int32_t *order,
                                RedScratchT<NT> &rs) {
    const int tid = threadIdx.x;
    int np = 0;
    const int max_pairs = min(nb / 2, 64);
    for (int pos = k; pos < pl && np / 2 < max_pairs; ++pos) {
        double inl = 0.0;
        for (int i = k + tid; i < pl; i += NT)
            if (i != pos) inl = fmax(inl, fabs(A.sym(i, pos)));
        const double in_max = red_max(rs, inl);
        double bv = 0.0;
        int bi = -1;
        const double *cp = &A.at(0, pos);
        sweep(pl + tid, ne, NT, [&](int i) { return cp[i]; }, [&](int i, double x) {
            const double v = fabs(x);
            if (v > bv) {
                bv = v;
                bi = i;
            }
        });
        double out_max;
        int r;
        red_maxfirst(rs, bv, bi, out_max, r);
        if (r < 0 || !(out_max > in_max)) continue;
        bool taken = false;
        for (int t = 1; t < np; t += 2) taken = taken || order[t] == rows[r];
        if (taken) continue;
        if (tid == 0) {
            order[np] = rows[pos];
            order[np + 1] = rows[r];
        }
        np += 2;
        __syncthreads();
    }
    return np;
}

// Lanes per column when the workgroup sweeps the lower part of several
// columns of up to `len` rows: log2 of a power of two covering len (at most
// the workgroup), so NT >> lg columns are in flight at a time.
template <int NT = kLT>
__device__ __forceinline__ int lg_rows(int len) {
    int lg = 0;
    while ((1 << lg) < len && (1 << lg) < NT) ++lg;
    return lg;
}

// sym_swap (ldlt_kernel.cpp:333-345): positions p < q. Columns [c_lo, p)
// swap rows p and q (c_lo = 0: all earlier columns).
template <int NT = kLT>
__device__ void sym_swap_wg(const FV &A, int p, int q, int32_t *rows, double *W, int wld, int w_cols, int c_lo) {
    if (p == q) return;
    if (p > q) {
        const int t = p;
        p = q;
        q = t;
    }
    const int tid = threadIdx.x;
    const int f = A.f;
    __syncthreads();
    if (tid == 0) {
        const double t = A.at(p, p);
        A.at(p, p) = A.at(q, q);
        A.at(q, q) = t;
        const int32_t tr = rows[p];
        rows[p] = rows[q];
        rows[q] = tr;
    }
    // rows p and q of the columns before p; column p and row q between
    // them; columns p and q below q
    swap_range([&](int m) { return &A.at(p, m); }, [&](int m) { return &A.at(q, m); }, c_lo + tid, p, NT);
    swap_range([&](int m) { return &A.at(m, p); }, [&](int m) { return &A.at(q, m); }, p + 1 + tid, q, NT);
    swap_range([&](int m) { return &A.at(m, p); }, [&](int m) { return &A.at(m, q); }, q + 1 + tid, f, NT);
    swap_range([&](int idx) { return &W[static_cast<int64_t>(idx) * wld + p]; },
               [&](int idx) { return &W[static_cast<int64_t>(idx) * wld + q]; }, tid, w_cols, NT);
    __syncthreads();
}

// Brings the rows order[0..np) to positions k, k + 1, ... in turn with
// sym_swap (ldlt_kernel.cpp:1076-1082): each row id is unique and, when
// its turn comes, lies in [k + t, ne); the workgroup finds it in one
// parallel sweep instead of a sequential scan.
template <int NT>
__device__ void place_rows_wg(const FV &A, int k, int ne, int np, const int32_t *order, int32_t *rows,
                              RedScratchT<NT> &rs) {
    const int tid = threadIdx.x;
    for (int t = 0; t < np; ++t) {
        const int32_t want = order[t];
        int mine = -1;
        for (int i0 = k + t + tid; i0 < ne; i0 += kU * NT) {
            int32_t v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) v[u] = rows[min(i0 + u * NT, ne - 1)];
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (i0 + u * NT < ne && v[u] == want) mine = i0 + u * NT;
        }
        double hit;
        int q;
        red_maxfirst(rs, mine >= 0 ? 1.0 : 0.0, mine, hit, q);
        if (q >= k + t) sym_swap_wg<NT>(A, k + t, q, rows, nullptr, 0, 0, 0);
    }
}

// eliminate_1x1 (ldlt_kernel.cpp:350-432): columns (k, ucol_end) get the
// rank-1 update c_j[i] -= c_k[i] * (c_k[j] / d) (multiply, then subtract),
// the unscaled column is saved in W column w_col, then c_k is scaled.
template <int NT = kLT>
__device__ void elim1_wg(const FV &A, int k, int ucol_end, double *W, int wld, int w_col) {
    const int tid = threadIdx.x;
    const int f = A.f;
    const double inv_d = 1.0 / A.at(k, k);
    double *ck = &A.at(0, k);
    if (W) col_copy(W + static_cast<int64_t>(w_col) * wld, ck, k + 1 + tid, f, NT);
    if (ucol_end > k + 1) {
        const int lg = lg_rows<NT>(f - (k + 1));
        for (int jb = k + 1; jb < ucol_end; jb += NT >> lg) {
            const int j = jb + (tid >> lg);
            if (j >= ucol_end) continue;
            const double lj = __dmul_rn(A.at(j, k), inv_d);
            if (lj == 0.0) continue;
            col_axpy(&A.at(0, j), ck, lj, j + (tid & ((1 << lg) - 1)), f, 1 << lg);
        }
    }
    __syncthreads();
    for (int i = k + 1 + tid; i < f; i += NT) ck[i] = __dmul_rn(ck[i], inv_d);
    __syncthreads();
}

// eliminate_2x2 (ldlt_kernel.cpp:436-545)
template <int NT = kLT>
__device__ void elim2_wg(const FV &A, int k, int ucol_end, double *W, int wld, int w_col) {
    const int tid = threadIdx.x;
    const int f = A.f;
    const double d00 = A.at(k, k), d10 = A.at(k + 1, k), d11 = A.at(k + 1, k + 1);
    double s = fabs(d00);
    if (s < fabs(d10)) s = fabs(d10);
    if (s < fabs(d11)) s = fabs(d11);
    const double a = d00 / s, b = d10 / s, c = d11 / s;
    const double inv = 1.0 / __dmul_rn(__dsub_rn(__dmul_rn(a, c), __dmul_rn(b, b)), s);
    const double i00 = __dmul_rn(c, inv), i10 = __dmul_rn(-b, inv), i11 = __dmul_rn(a, inv);
    double *ck0 = &A.at(0, k), *ck1 = &A.at(0, k + 1);
    if (W) {
        col_copy(W + static_cast<int64_t>(w_col) * wld, ck0, k + 2 + tid, f, NT);
        col_copy(W + static_cast<int64_t>(w_col + 1) * wld, ck1, k + 2 + tid, f, NT);
    }
    if (ucol_end > k + 2) {
        const int lg = lg_rows<NT>(f - (k + 2));
        for (int jb = k + 2; jb < ucol_end; jb += NT >> lg) {
            const int j = jb + (tid >> lg);
            if (j >= ucol_end) continue;
            const double a0 = A.at(j, k), a1 = A.at(j, k + 1);
            const double l0 = __dadd_rn(__dmul_rn(i00, a0), __dmul_rn(i10, a1));
            const double l1 = __dadd_rn(__dmul_rn(i10, a0), __dmul_rn(i11, a1));
            if (l0 == 0.0 && l1 == 0.0) continue;
            col_axpy2(&A.at(0, j), ck0, ck1, l0, l1, j + (tid & ((1 << lg) - 1)), f, 1 << lg);
        }
    }
    __syncthreads();
    for (int i = k + 2 + tid; i < f; i += NT) {
        const double a0 = A.at(i, k), a1 = A.at(i, k + 1);
        A.at(i, k) = __dadd_rn(__dmul_rn(i00, a0), __dmul_rn(i10, a1));
        A.at(i, k + 1) = __dadd_rn(__dmul_rn(i10, a0), __dmul_rn(i11, a1));
    }
    __syncthreads();
}

__device__ __forceinline__ void inertia_1x1(double d, double small, int *cnt) {
    if (!isfinite(d) || fabs(d) < small) ++cnt[2];
    else if (d > 0.0) ++cnt[0];
    else ++cnt[1];
}

// add_inertia_2x2 (dense_indef.cpp:52-88) with scaled_det_2x2 (pivoting.cpp:18-28)
__device__ void inertia_2x2(double d00, double d10, double d11, double small, int *cnt) {
    if (d10 == 0.0) {
        inertia_1x1(d00, small, cnt);
        inertia_1x1(d11, small, cnt);
        return;
    }
    double s = fabs(d00);
    if (s < fabs(d10)) s = fabs(d10);
    if (s < fabs(d11)) s = fabs(d11);
    double det_s = 0.0;
    if (isfinite(s) && s != 0.0) {
        const double a = d00 / s, b = d10 / s, c = d11 / s;
        det_s = __dsub_rn(__dmul_rn(a, c), __dmul_rn(b, b));
    }
    if (s == 0.0 || !isfinite(det_s)) {
        cnt[2] += 2;
        return;
    }
    const double tr_s = __dadd_rn(d00 / s, d11 / s);
    if (det_s < -small) {
        ++cnt[0];
        ++cnt[1];
    } else if (fabs(det_s) <= small) {
        ++cnt[2];
        if (tr_s > small) ++cnt[0];
        else if (tr_s < -small) ++cnt[1];
        else ++cnt[2];
    } else {
        if (tr_s > 0.0) cnt[0] += 2;
        else if (tr_s < 0.0) cnt[1] += 2;
        else cnt[2] += 2;
    }
}

// force_static_pivot (ldlt_kernel.cpp:285-320): the column of [k, ne)
// whose diagonal is largest relative to its column (first on ties; column k
// in the in-order mode), |d| raised to sp keeping its sign when smaller.
// raised: 0 as it is, 1 from a nonzero |d|, 2 from an exact zero (below
// eff_small or not finite). Every thread returns the same column.
template <int NT>
__device__ int force_static_wg(const FV &A, int k, int ne, double sp, const Prm &prm, RedScratchT<NT> &rs,
                               int &raised, int &nstatic, double &growth) {
    const int tid = threadIdx.x;
    const int f = A.f;
    int best = k;
    double best_ratio = -1.0, best_colmax = 0.0;
    const int search_end = prm.min_abs > 0.0 ? k : ne;
    for (int cand = k; cand < search_end; ++cand) {
        double m = 0.0;
        const double *cc = &A.at(0, cand);
        sweep(k + tid, cand, NT, [&](int i) { return A.at(cand, i); }, [&](int, double v) { m = fmax(m, fabs(v)); });
        sweep(cand + 1 + tid, f, NT, [&](int i) { return cc[i]; }, [&](int, double v) { m = fmax(m, fabs(v)); });
        const double colmax = red_max(rs, m);
        const double a = fabs(A.at(cand, cand));
        const double ratio = colmax > 0.0 ? a / colmax : a;
        if (ratio > best_ratio) {
            best_ratio = ratio;
            best = cand;
            best_colmax = colmax;
        }
    }
    __syncthreads();
    double dv = A.at(best, best);
    raised = 0;
    if (!(fabs(dv) >= sp)) {
        raised = (isfinite(dv) && fabs(dv) >= prm.eff_small) ? 1 : 2;
        dv = dv < 0.0 ? -sp : sp;
        ++nstatic;
        __syncthreads();
        if (tid == 0) A.at(best, best) = dv;
    }
    if (best_colmax > 0.0 && fabs(dv) > 0.0) growth = fmax(growth, best_colmax / fabs(dv));
    __syncthreads();
    return best;
}

// apply_pivot (ldlt_kernel.cpp:665-698). Thread 0 keeps the counters.
template <int NT = kLT>
__device__ void apply_pivot_wg(const FV &A, int kind, int pa, int pb, int &k, int32_t *rows, int8_t *tags, double *W,
                               int wld, int w_cols, int ucol_end, const Prm &prm, int *cnt, double &flops, int c_lo,
                               bool as_zero = false) {
    const int f = A.f;
    if (kind == 1) {
        sym_swap_wg<NT>(A, k, pa, rows, W, wld, w_cols, c_lo);
        if (threadIdx.x == 0) {
            if (as_zero) ++cnt[2];
            else inertia_1x1(A.at(k, k), prm.eff_small, cnt);
            tags[k] = 1;
        }
        elim1_wg<NT>(A, k, ucol_end, W, wld, w_cols);
        const double rem = static_cast<double>(f - (k + 1));
        flops += rem + rem * (rem + 1.0);
        k += 1;
        return;
    }
    int partner = pb;
    if (pa != k) {
        sym_swap_wg<NT>(A, k, pa, rows, W, wld, w_cols, c_lo);
        if (partner == k) partner = pa;
    }
    if (partner != k + 1) sym_swap_wg<NT>(A, k + 1, partner, rows, W, wld, w_cols, c_lo);
    if (threadIdx.x == 0) {
        inertia_2x2(A.at(k, k), A.at(k + 1, k), A.at(k + 1, k + 1), prm.small_pivot, cnt);
        tags[k] = 2;
        tags[k + 1] = -1;
    }
    __syncthreads();
    elim2_wg<NT>(A, k, ucol_end, W, wld, w_cols);
    const double rem = static_cast<double>(f - (k + 2));
    flops += 6.0 * rem + 2.0 * rem * (rem + 1.0);
    k += 2;
}

// ---------------------------------------------------------------------------
// Unblocked fronts (f < blocked_min_front): the whole front in LDS, one
// workgroup, ldlt_factor_front's unblocked loop (ldlt_kernel.cpp:977-993).
constexpr int kSmallMax = 95;

__global__ __launch_bounds__(kLT) void k_ldlt_small(LdltDev d, const int32_t *__restrict__ flist, Prm prm) {
    __shared__ double Fs[kSmallMax * kSmallMax];
    __shared__ RedScratch rs;
    __shared__ int cnt[3];
    const int s = flist[blockIdx.x];
    FrontLd &Fd = d.fronts[s];
    const int f = Fd.f, ne = Fd.ne;
    double *G = d.arena + Fd.fofs;
    int32_t *rows = d.rows + Fd.rofs;
    int8_t *tags = d.tags + Fd.rofs;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < f * f; idx += kLT) Fs[idx] = G[idx];
    if (tid < 3) cnt[tid] = 0;
    __syncthreads();
    FV A{Fs, f, f};
    int k = 0;
    double flops = 0.0;
    Prm fp;
    const double sp = front_prm(prm, Fd, fp);
    int nstatic = 0;
    double growth = 0.0;
    while (k < ne) {
        int kind, pa, pb;
        select_pivot_wg(A, k, ne, ne, fp, rs, kind, pa, pb);
        bool az = false;
        if (kind == 0) {
            if (!(sp > 0.0)) break;
            int raised;
            pa = force_static_wg(A, k, ne, sp, fp, rs, raised, nstatic, growth);
            kind = 1;
            az = static_as_zero(raised, fp, Fd);
        }
        apply_pivot_wg(A, kind, pa, pb, k, rows, tags, nullptr, 0, 0, f, fp, cnt, flops, 0, az);
    }
    __syncthreads();
    for (int idx = tid; idx < f * f; idx += kLT) G[idx] = Fs[idx];
    if (tid == 0) {
        Fd.p = k;
        Fd.npos = cnt[0];
        Fd.nneg = cnt[1];
        Fd.nzero = cnt[2];
        Fd.flops = flops;
        Fd.nstatic = nstatic;
        Fd.growth = growth;
        Fd.stalls = Fd.pulls = Fd.rotated = Fd.searches = 0;  // unblocked: no panels
    }
}

void launch_ldlt_small(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double u, double eff_small,
                       double small_pivot, const StaticPrm &S) {
    if (nf <= 0) return;
    Prm prm{u, eff_small, small_pivot, 0.0, S};
    for (int64_t o = 0; o < nf; o += kMaxLaunchGroups)
        k_ldlt_small<<<dim3(static_cast<unsigned>(std::min<int64_t>(kMaxLaunchGroups, nf - o))), kLT, 0, st>>>(
            d, flist + o, prm);
}

// ---------------------------------------------------------------------------
// Blocked fronts (f >= blocked_min_front), one workgroup per front in global
// memory: ldlt_factor_front's blocked path (ldlt_kernel.cpp:1028-1139):
// first pass over nb-wide panels with partner pulls and rotation of failed
// columns, second pass with a full search on a stall, deferred panel
// updates, final Schur update. W (f x nb) in the front's workspace.
__device__ void deferred_update_wg(const FV &A, int p0, int p1, int c0, int e, const double *W, int wld) {
    const int jb = p1 - p0;
    const int m = e - c0;
    if (jb <= 0 || m <= 0) return;
    const int f = A.f;
    const int tid = threadIdx.x;
    const int lg = lg_rows(f - c0);
    const int st = 1 << lg;
    for (int cb = c0; cb < e; cb += kLT >> lg) {
        const int j = cb + (tid >> lg);
        if (j >= e) continue;
        double *cj = &A.at(0, j);
        // kU rows at a time, each summed over t in order
        for (int i0 = j + (tid & (st - 1)); i0 < f; i0 += kU * st) {
            double sacc[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) sacc[u] = 0.0;
            for (int t = 0; t < jb; ++t) {
                const double w = W[static_cast<int64_t>(t) * wld + j];
                const double *ct = &A.at(0, p0 + t);
                double v[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) v[u] = ct[min(i0 + u * st, f - 1)];
#pragma unroll
                for (int u = 0; u < kU; ++u) sacc[u] += v[u] * w;
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (i0 + u * st < f) cj[i0 + u * st] -= sacc[u];
        }
    }
    __syncthreads();
}

__device__ void schur_update_wg(const FV &A, int p, int e, const int8_t *tags) {
    const int f = A.f;
    const int q = f - e;
    if (p <= 0 || q <= 0) return;
    const int64_t tot = static_cast<int64_t>(q) * q;
    for (int64_t idx = threadIdx.x; idx < tot; idx += kLT) {
        const int jj = static_cast<int>(idx / q), ii = static_cast<int>(idx % q);
        if (ii < jj) continue;
        const int i = e + ii, j = e + jj;
        double sacc = 0.0;
        for (int c = 0; c < p; ++c) {
            const int8_t tg = tags[c];
            double ld;
            if (tg == 1) ld = A.at(j, c) * A.at(c, c);
            else if (tg == 2) ld = A.at(j, c) * A.at(c, c) + A.at(j, c + 1) * A.at(c + 1, c);
            else if (tg == -1) ld = A.at(j, c - 1) * A.at(c, c - 1) + A.at(j, c) * A.at(c, c);
            else ld = 0.0;
            sacc += A.at(i, c) * ld;
        }
        A.at(i, j) -= sacc;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kLT) void k_ldlt_blocked(LdltDev d, const int32_t *__restrict__ flist, Prm prm,
                                                     int32_t nb) {
    __shared__ RedScratch rs;
    __shared__ int cnt[3];
    __shared__ int32_t order[128];
    const int s = flist[blockIdx.x];
    FrontLd &Fd = d.fronts[s];
    const int f = Fd.f, ne = Fd.ne;
    int32_t *rows = d.rows + Fd.rofs;
    int8_t *tags = d.tags + Fd.rofs;
    double *W = d.work + Fd.wofs;
    const int wld = f;
    const int tid = threadIdx.x;
    FV A{d.arena + Fd.fofs, f, f};
    if (tid < 3) cnt[tid] = 0;
    __syncthreads();
    int k = 0;
    double flops = 0.0;
    Prm fp;
    const double sp = front_prm(prm, Fd, fp);
    const Prm &prm_f = fp;
    int nstatic = 0;
    double growth = 0.0;
    int n_stall = 0, n_pull = 0, n_rot = 0, n_search = 0;
    // the first pass with stall handling is for threshold pivoting only
    // (ldlt_kernel.cpp:1028); static fronts go straight to the second
    if (!(sp > 0.0) && !(fp.min_abs > 0.0)) {
        int live = ne, k_paired = -1;
        while (k < live) {
            const int ps = k;
            const int pl = min(live, ps + nb);
            bool stalled = false;
            while (k < pl) {
                int kind, pa, pb;
                select_pivot_wg(A, k, ne, pl, prm_f, rs, kind, pa, pb);
                if (kind == 0) {
                    stalled = true;
                    break;
                }
                apply_pivot_wg(A, kind, pa, pb, k, rows, tags, W, wld, k - ps, pl, prm_f, cnt, flops, 0);
            }
            deferred_update_wg(A, ps, k, pl, ne, W, wld);
            n_stall += stalled ? 1 : 0;
            if (stalled && k != k_paired) {
                // (failed column, partner) row-id pairs, distinct partners
                const int np = partner_pairs_wg(A, k, pl, ne, nb, rows, order, rs);
                if (np > 0) {
                    n_pull += np / 2;
                    place_rows_wg(A, k, ne, np, order, rows, rs);
                    live = max(live, k + np);
                    k_paired = k;
                    continue;
                }
            }
            if (stalled) {
                const int m = pl - k;
                const int mv = min(m, live - pl);
                for (int t = 0; t < mv; ++t) sym_swap_wg(A, k + t, live - mv + t, rows, nullptr, 0, 0, 0);
                live -= m;
                n_rot += m;
            }
        }
    }
    while (k < ne) {
        const int ps = k;
        const int pl = min(ne, ps + nb);
        bool stalled = false;
        while (k < pl) {
            int kind, pa, pb;
            select_pivot_wg(A, k, ne, pl, prm_f, rs, kind, pa, pb);
            if (kind == 0) {
                stalled = true;
                break;
            }
            apply_pivot_wg(A, kind, pa, pb, k, rows, tags, W, wld, k - ps, pl, prm_f, cnt, flops, 0);
        }
        deferred_update_wg(A, ps, k, pl, ne, W, wld);
        if (stalled && k < ne) {
            ++n_search;
            const int ss = k;
            int kind, pa, pb;
            select_pivot_wg(A, k, ne, ne, prm_f, rs, kind, pa, pb);
            bool az = false;
            if (kind == 0) {
                if (!(sp > 0.0)) break;
                int raised;
                pa = force_static_wg(A, k, ne, sp, prm_f, rs, raised, nstatic, growth);
                kind = 1;
                az = static_as_zero(raised, prm_f, Fd);
            }
            apply_pivot_wg(A, kind, pa, pb, k, rows, tags, W, wld, 0, ss, prm_f, cnt, flops, 0, az);
            deferred_update_wg(A, ss, k, k, ne, W, wld);
        }
    }
    schur_update_wg(A, k, ne, tags);
    if (tid == 0) {
        Fd.p = k;
        Fd.npos = cnt[0];
        Fd.nneg = cnt[1];
        Fd.nzero = cnt[2];
        Fd.flops = flops;
        Fd.nstatic = nstatic;
        Fd.growth = growth;
        Fd.stalls = n_stall;
        Fd.pulls = n_pull;
        Fd.rotated = n_rot;
        Fd.searches = n_search;
    }
}

void launch_ldlt_blocked(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double u,
                         double eff_small, double small_pivot, int32_t nb, const StaticPrm &S) {
    if (nf <= 0) return;
    Prm prm{u, eff_small, small_pivot, 0.0, S};
    for (int64_t o = 0; o < nf; o += kMaxLaunchGroups)
        k_ldlt_blocked<<<dim3(static_cast<unsigned>(std::min<int64_t>(kMaxLaunchGroups, nf - o))), kLT, 0, st>>>(
            d, flist + o, prm, nb);
}

// ---------------------------------------------------------------------------
// Large fronts: ldlt_factor_front's blocked path (ldlt_kernel.cpp:1028-1139)
// split over kernels. The panel kernel runs the pivot logic of one panel
// (select_pivot / apply_pivot on the panel's columns, the unscaled columns
// kept in W); the update kernel applies the deferred updates on MFMA; the
// stall kernel does the partner pulls or the rotation of a stalled panel.
// The deferred update of a panel (columns [pl, ne), rows [pl, f),
// ldlt_kernel.cpp:705-761) is applied at once to the outer block's columns
// and held back for the columns past it until a flush; sums differ from the
// reference's only in order.

// Tile `tile` of the lower part (row block >= column block) of nrb x ncb
// tile blocks, in 8 x 8 super-blocks: super-columns left to right, each
// from its diagonal super-block down, tiles column-major inside a
// super-block. Consecutive tiles (one XCD's run) then share 8 row and 8
// column operand panels, which stay in its L2.
__device__ __forceinline__ void super_block_tile(int nrb, int ncb, int tile, int &ib, int &jb) {
    int sj = 0, rem = tile;
    for (;; ++sj) {
        const int w = min(8, ncb - 8 * sj), h = min(8, nrb - 8 * sj);
        // diagonal super-block: column jj holds rows jj .. h-1
        const int diag = w * h - w * (w - 1) / 2;
        const int total = diag + w * (nrb - 8 * sj - h);
        if (rem < total || 8 * (sj + 1) >= ncb) {
            if (rem < diag) {
                int jj = 0;
                while (rem >= h - jj) {
                    rem -= h - jj;
                    ++jj;
                }
                ib = 8 * sj + jj + rem;
                jb = 8 * sj + jj;
                return;
            }
            rem -= diag;
            const int below = nrb - 8 * sj - h;  // rows under the diagonal super-block
            const int nf = below / 8, lr = below % 8;
            const int base = 8 * sj + h;
            if (rem < nf * 8 * w) {
                const int b = rem / (8 * w), q = rem % (8 * w);
                ib = base + 8 * b + q % 8;
                jb = 8 * sj + q / 8;
            } else {
                const int q = rem - nf * 8 * w;
                ib = base + 8 * nf + q % lr;
                jb = 8 * sj + q / lr;
            }
            return;
        }
        rem -= total;
    }
}

// MFMA tile of C[i][j] -= sum_t A[i][t] * B[j][t - boff] over the lower
// part of rows [c0, f) x columns [c0, c1) of front F (c1 - c0 <= f - c0);
// tile `tile` of the region in super-block order. MODE 0: FP64 MFMA;
// 1: FP32 MFMA on the FP64 operands (rounded as they enter LDS); 2: BF16
// MFMA (operands rounded to bf16); 3: split BF16 (hi / lo operands, four
// MFMAs). Modes 1-3 accumulate in FP32 over the panel and subtract the
// result from the FP64 entries.
constexpr int kBigKcBf = 32;  // panel columns per LDS stage of the BF16 tiles (FP64 operands in registers)
template <int MODE>
__host__ __device__ constexpr int big_lds_bytes() {
    return MODE <= 1 ? 4 * 16 * kT * static_cast<int>(sizeof(double)) : bf_lds_bytes<MODE == 3, kBigKcBf>();
}
template <int MODE>
__device__ __forceinline__ void big_tile(const LdltDev &d, const FrontLd &F, const double *Bm, int boff, int c0, int c1, int t0,
                         int t1, int tile, unsigned char *smem) {
    typedef typename std::conditional<MODE == 0, double, float>::type E;
    typedef SyrkMfma<E> M;
    const int f = F.f;
    int ib, jb;
    super_block_tile((f - c0 + kT - 1) / kT, (c1 - c0 + kT - 1) / kT, tile, ib, jb);
    const int i0 = c0 + ib * kT, j0 = c0 + jb * kT;
    double *G = d.arena + F.fofs;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w >> 1, wc = w & 1;
    const bool active = (i0 + wr * 64 + 63) >= (j0 + wc * 64);
    typename M::acc_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = M::zero();
    if constexpr (MODE <= 1) {
        if (active) tile_loop<E, true, 16, 2, double>(G, Bm, f, boff, i0, j0, f, c1, t0, t1, smem, acc);
        else tile_loop<E, false, 16, 2, double>(G, Bm, f, boff, i0, j0, f, c1, t0, t1, smem, acc);
    } else {
        if (active) tile_loop_bf16<MODE == 3, true, kBigKcBf, double>(G, Bm, f, boff, i0, j0, f, c1, t0, t1, smem, acc);
        else tile_loop_bf16<MODE == 3, false, kBigKcBf, double>(G, Bm, f, boff, i0, j0, f, c1, t0, t1, smem, acc);
    }
    if (!active) return;
    double *junk = d.junk + threadIdx.x;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        double *dst[4][4];
        double old[4][4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = j0 + wc * 64 + m * 16 + M::row(v, lane);
            double *col = G + static_cast<int64_t>(min(j, c1 - 1)) * f;
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) {
                const int i = i0 + wr * 64 + nn * 16 + (lane & 15);
                dst[nn][v] = (i < f && j < c1 && i >= j) ? col + i : junk;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) old[nn][v] = *dst[nn][v];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) *dst[nn][v] = old[nn][v] - acc[m][nn][v];
    }
}

// Tiles [lo, hi) of a launch's range of nt for this workgroup's XCD, which
// its workgroups stride over: workgroups are dispatched to the 8 XCDs round
// robin by linear index, so with a multiple of 8 per front, blockIdx.x % 8
// is the XCD. Each XCD then works through a contiguous run of tiles (one
// column block's operand panel stays in its L2) instead of every eighth.
struct XcdRange {
    int lo, hi, step;
};
__device__ __forceinline__ XcdRange xcd_range(int nt) {
    const int x = blockIdx.x & 7, per = gridDim.x >> 3;
    const int lo = static_cast<int>(static_cast<int64_t>(nt) * x / 8);
    const int hi = static_cast<int>(static_cast<int64_t>(nt) * (x + 1) / 8);
    return XcdRange{lo + static_cast<int>(blockIdx.x >> 3), hi, per};
}

__device__ __forceinline__ int big_tiles(int f, int c0, int c1) {
    if (c1 <= c0) return 0;
    const int nr = (f - c0 + kT - 1) / kT, nc = (c1 - c0 + kT - 1) / kT;
    return nc * nr - nc * (nc - 1) / 2;
}

// ---------------------------------------------------------------------------
// Speculative panel groups (first pass of the large-front path). When every
// pivot of a panel passes select_pivot's first test (the diagonal against u
// times the largest entry of its column, ldlt_kernel.cpp:145-178), the
// panel's pivots are 1x1, in order, without swaps, and each row of the front
// goes through a fixed sequence of row-local operations: eliminate_1x1's
// rank-1 updates within the panel (ldlt_kernel.cpp:350-432) and the deferred
// update of the later columns (705-761). A group of up to four panels of 16
// is evaluated under that assumption. The stats launch factors the group's
// diagonal region in one wave of every workgroup (the multipliers and the
// unscaled columns go to LDS), lets every thread carry one row below it
// through the same operations, and records per pivot the largest entry of
// its column. The panel kernel then applies the first test to each pivot in
// order and accepts the leading panels that pass; the apply launch redoes the
// accepted panels and writes their columns and W. A panel that fails is
// factored by the panel kernel's full search. Per row, the operations and
// their order are those of the one-panel path, so are the accepted values
// and the pivot decisions.
// Guarded loads are unconditional loads from clamped (in-front) addresses
// times a 0 / 1 mask, so the compiler keeps every load of a row in flight
// (inputs are finite: the factorization rejects non-finite values).
__device__ __forceinline__ double masked(double v, bool keep) { return v * (keep ? 1.0 : 0.0); }

__device__ __forceinline__ double rdlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Columns of the group the one-panel path would factor next if no panel
// stalled: panels of nb from k, each ending at min(live, start + nb), the
// group ending after the panel whose end triggers the outer block's flush.
// Every panel but the last has 16 columns.
__device__ __forceinline__ int group_cols(const BigLd &B, int nb) {
    if (B.pass != 0 || !B.spec || nb != 16 || B.k >= B.live) return 0;
    int pos = B.k;
    for (int m = 0; m < 4 && pos < B.live; ++m) {
        const int pl = min(B.live, pos + nb);
        if (pl > B.oe) break;
        pos = pl;
        if (pl + nb > B.oe || pl >= B.live) break;
    }
    return pos - B.k;
}

// Group buffer per front (doubles): the diagonal region's results, written
// by the region launch and read by the stats, panel and apply launches,
// then 64 column maxima per workgroup of the stats launch.
constexpr int64_t kGdMax = 0;                               // 64: largest |entry| below each pivot in the region
constexpr int64_t kGdPiv = 64;                              // 64: pivots
constexpr int64_t kGdInv = 128;                             // 64: 1 / pivot
constexpr int64_t kGdLm = 192;                              // 64 x 16: multipliers of each pivot for its panel
constexpr int64_t kGdWd = kGdLm + 64 * 16;                  // 64 x 64: [t][r] row r of the region at pivot t, unscaled
constexpr int64_t kGdFin = kGdWd + 64 * 64;                 // 64 x 64: [c][r] the region's final entries
constexpr int64_t kGMax = kGdFin + 64 * 64;
static_assert(kGMax == kGroupBuf, "group buffer layout");
__device__ __forceinline__ int64_t gpart_stride(int nch) { return kGMax + static_cast<int64_t>(nch) * kGroupCols; }

// Per-panel steps of a group (panel M: columns [16 M, 16 M + 16) of the
// group), templated so that every index into a row's registers is a
// constant; FULL: the group has all 64 columns (no column guards). A
// skipped update (multiplier 0, or a row above the column) is a select,
// the same value as the reference's skip. Diagonal region: lane r of
// wave 0 holds row ps + r.
template <int M, bool FULL>
__device__ __forceinline__ void gdiag_pivots(double (&a)[kGroupCols], int r, int ncols,
                                             double (*Wd)[kGroupCols + 1], double (*Lm)[16], double *invd,
                                             double *dd) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int kk = 16 * M + q;
        if (FULL || kk < ncols) {
            // stores without conditions: every lane writes the same value
            // to dd, invd and Lm; Wd[kk][r] for r <= kk is never read
            const double v = a[kk];
            const double dk = rdlane(v, kk);
            const double inv = 1.0 / dk;
            dd[kk] = dk;
            invd[kk] = inv;
            Wd[kk][r] = v;
#pragma unroll
            for (int jq = q + 1; jq < 16; ++jq) {
                const int j = 16 * M + jq;
                if (FULL || j < ncols) {
                    const double lj = __dmul_rn(rdlane(v, j), inv);
                    Lm[kk][jq] = lj;
                    const double t = __dsub_rn(a[j], __dmul_rn(v, lj));
                    a[j] = (lj != 0.0 && r >= j) ? t : a[j];
                }
            }
            const double sc = __dmul_rn(v, inv);
            a[kk] = r > kk ? sc : a[kk];
        }
    }
}

// Rows below the region (one per thread): the same pivots with the
// multipliers from LDS; STATS keeps |entry| per pivot in red[q][tid].
template <int M, bool STATS, bool FULL>
__device__ __forceinline__ void grow_pivots(double (&a)[kGroupCols], int ncols, const double (*Lm)[16],
                                            const double *invd, double (*red)[kLT + 1], double *W, int64_t wcol0,
                                            int f, int i, bool rowok) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int kk = 16 * M + q;
        if (FULL || kk < ncols) {
            const double v = a[kk];
            if (STATS) red[q][threadIdx.x] = fabs(v);
            if (rowok) W[(wcol0 + kk) * f + i] = v;
#pragma unroll
            for (int jq = q + 1; jq < 16; ++jq) {
                const int j = 16 * M + jq;
                if (FULL || j < ncols) {
                    const double lj = Lm[kk][jq];
                    const double t = __dsub_rn(a[j], __dmul_rn(v, lj));
                    a[j] = lj != 0.0 ? t : a[j];
                }
            }
            a[kk] = __dmul_rn(v, invd[kk]);
        }
    }
}

// Deferred update of the group's columns after panel M by its pivots
// (below = true for every row under the region, else rows at or below the
// column).
template <int M, bool FULL>
__device__ __forceinline__ void gdeferred(double (&a)[kGroupCols], int ncols, const double (*Wd)[kGroupCols + 1],
                                          int r, bool below) {
#pragma unroll
    for (int j = 16 * M + 16; j < kGroupCols; ++j) {
        if (FULL || j < ncols) {
            double sacc = 0.0;
#pragma unroll
            for (int t = 0; t < 16; ++t) sacc = __dadd_rn(sacc, __dmul_rn(a[16 * M + t], Wd[16 * M + t][j]));
            const double u = __dsub_rn(a[j], sacc);
            a[j] = (below || r >= j) ? u : a[j];
        }
    }
}

// Column maxima of panel M's pivots over the workgroup's rows (red holds
// |entry| per pivot and thread).
template <int M>
__device__ __forceinline__ void gpanel_max(double (*red)[kLT + 1], double (*red2)[16], double *wgmax) {
    const int tid = threadIdx.x;
    __syncthreads();
    {
        const int q = tid & 15, part = tid >> 4;
        double mx = 0.0;
#pragma unroll
        for (int rw = 0; rw < 16; ++rw) mx = fmax(mx, red[q][part * 16 + rw]);
        red2[part][q] = mx;
    }
    __syncthreads();
    if (tid < 16) {
        double mx = 0.0;
#pragma unroll
        for (int pp = 0; pp < 16; ++pp) mx = fmax(mx, red2[pp][tid]);
        wgmax[16 * M + tid] = mx;
    }
}

// The four panels of the diagonal region in turn (FULL: 64 columns).
template <bool FULL>
__device__ __forceinline__ void gdiag_run(double (&a)[kGroupCols], int r, int ncols, double (*Wd)[kGroupCols + 1],
                                          double (*Lm)[16], double *invd, double *dd) {
    gdiag_pivots<0, FULL>(a, r, ncols, Wd, Lm, invd, dd);
    __syncthreads();
    if (FULL || 16 < ncols) {
        gdeferred<0, FULL>(a, ncols, Wd, r, false);
        gdiag_pivots<1, FULL>(a, r, ncols, Wd, Lm, invd, dd);
    }
    __syncthreads();
    if (FULL || 32 < ncols) {
        gdeferred<1, FULL>(a, ncols, Wd, r, false);
        gdiag_pivots<2, FULL>(a, r, ncols, Wd, Lm, invd, dd);
    }
    __syncthreads();
    if (FULL || 48 < ncols) {
        gdeferred<2, FULL>(a, ncols, Wd, r, false);
        gdiag_pivots<3, FULL>(a, r, ncols, Wd, Lm, invd, dd);
    }
    __syncthreads();
}

// The rows below the region (one per thread) through the four panels.
template <bool STATS, bool FULL>
__device__ __forceinline__ void grow_run(double (&a)[kGroupCols], int ncols, const double (*Wd)[kGroupCols + 1],
                                         const double (*Lm)[16], const double *invd, double (*red)[kLT + 1],
                                         double (*red2)[16], double *wgmax, double *W, int64_t wcol0, int f, int i,
                                         bool rowok) {
    grow_pivots<0, STATS, FULL>(a, ncols, Lm, invd, red, W, wcol0, f, i, rowok);
    if (STATS) gpanel_max<0>(red, red2, wgmax);
    if (FULL || 16 < ncols) {
        gdeferred<0, FULL>(a, ncols, Wd, 0, true);
        grow_pivots<1, STATS, FULL>(a, ncols, Lm, invd, red, W, wcol0, f, i, rowok);
        if (STATS) gpanel_max<1>(red, red2, wgmax);
    }
    if (FULL || 32 < ncols) {
        gdeferred<1, FULL>(a, ncols, Wd, 0, true);
        grow_pivots<2, STATS, FULL>(a, ncols, Lm, invd, red, W, wcol0, f, i, rowok);
        if (STATS) gpanel_max<2>(red, red2, wgmax);
    }
    if (FULL || 48 < ncols) {
        gdeferred<2, FULL>(a, ncols, Wd, 0, true);
        grow_pivots<3, STATS, FULL>(a, ncols, Lm, invd, red, W, wcol0, f, i, rowok);
        if (STATS) gpanel_max<3>(red, red2, wgmax);
    }
}

// The group's diagonal region [ps, ps + ncols), one wave per front: lane r
// holds row ps + r. Results to the group buffer.
__global__ __launch_bounds__(64) void k_big_gdiag(LdltDev d, const BigLd *__restrict__ states,
                                                  double *__restrict__ gpart, int32_t nch) {
    __shared__ double Wd[kGroupCols][kGroupCols + 1];
    __shared__ double Lm[kGroupCols][16];
    __shared__ double invd[kGroupCols], dd[kGroupCols];
    const BigLd B = states[blockIdx.x];
    const int ps = B.k, ncols = group_cols(B, 16);
    if (ncols <= 0) return;
    const FrontLd F = d.fronts[B.s];
    const int f = F.f;
    const double *G = d.arena + F.fofs;
    double *gp = gpart + blockIdx.x * gpart_stride(nch);
    const int r = threadIdx.x;
    double a[kGroupCols];
    const bool mine = r < ncols;
    const int rr = ps + min(r, ncols - 1);
#pragma unroll
    for (int c = 0; c < kGroupCols; ++c)
        a[c] = masked(G[static_cast<int64_t>(ps + min(c, ncols - 1)) * f + rr], mine && c <= r);
    if (ncols == kGroupCols) gdiag_run<true>(a, r, ncols, Wd, Lm, invd, dd);
    else gdiag_run<false>(a, r, ncols, Wd, Lm, invd, dd);
#pragma unroll
    for (int c = 0; c < kGroupCols; ++c) gp[kGdFin + c * kGroupCols + r] = a[c];
    for (int e = r; e < kGroupCols * kGroupCols; e += 64) gp[kGdWd + e] = Wd[e >> 6][e & 63];
    for (int e = r; e < kGroupCols * 16; e += 64) gp[kGdLm + e] = Lm[e >> 4][e & 15];
    gp[kGdPiv + r] = dd[r];
    gp[kGdInv + r] = invd[r];
    double mx = 0.0;
    for (int rw = r + 1; rw < ncols; ++rw) mx = fmax(mx, fabs(Wd[r][rw]));
    gp[kGdMax + r] = mx;
}

// Rows of a group below its diagonal region, [ps + ncols, f), in chunks of
// kGroupRows per workgroup, one row per thread, with the region's tables:
// column maxima per pivot to the group buffer, each row's unscaled column
// entries to W (the columns the panel kernel's pivots will have: W column
// ps - pend0 + t, pend0 being the W base the panel kernel starts from) and
// its final group entries to W's speculation columns [kGroupSpec,
// kGroupSpec + 64). Nothing of the front square changes: a rejected panel
// finds its columns as they were. W columns of rejected pivots are
// rewritten (rows below the pivot) when those positions get their pivots.
__global__ __launch_bounds__(kLT) void k_big_group(LdltDev d, const BigLd *__restrict__ states,
                                                   double *__restrict__ gpart, int32_t nch) {
    __shared__ double Wd[kGroupCols][kGroupCols + 1];
    __shared__ double Lm[kGroupCols][16];
    __shared__ double invd[kGroupCols], wgmax[kGroupCols];
    __shared__ double red[16][kLT + 1];
    __shared__ double red2[16][16];
    const BigLd B = states[blockIdx.y];
    const int ps = B.k, ncols = group_cols(B, 16);
    if (ncols <= 0) return;
    const FrontLd F = d.fronts[B.s];
    const int f = F.f;
    const int gend = ps + ncols;
    const int bx = blockIdx.x;
    if (bx >= max(1, (f - gend + kGroupRows - 1) / kGroupRows)) return;
    const double *G = d.arena + F.fofs;
    double *W = d.work + F.wofs;
    const double *gp = gpart + blockIdx.y * gpart_stride(nch);
    const int64_t wcol0 = ps - B.pend0;  // W column of the group's first pivot
    const int tid = threadIdx.x;
    for (int e = tid; e < kGroupCols * kGroupCols; e += kLT) Wd[e >> 6][e & 63] = gp[kGdWd + e];
    for (int e = tid; e < kGroupCols * 16; e += kLT) Lm[e >> 4][e & 15] = gp[kGdLm + e];
    if (tid < kGroupCols) invd[tid] = gp[kGdInv + tid];
    __syncthreads();
    const int i = gend + bx * kGroupRows + tid;
    const bool rowok = i < f;
    const int il = min(i, f - 1);
    double a[kGroupCols];
#pragma unroll
    for (int c = 0; c < kGroupCols; ++c)
        a[c] = masked(G[static_cast<int64_t>(ps + min(c, ncols - 1)) * f + il], rowok && c < ncols);
    // (the unguarded form unrolls into more registers than a 256-thread
    // workgroup has)
    grow_run<true, false>(a, ncols, Wd, Lm, invd, red, red2, wgmax, W, wcol0, f, i, rowok);
    __syncthreads();
    if (tid < ncols) gpart[blockIdx.y * gpart_stride(nch) + kGMax + bx * kGroupCols + tid] = wgmax[tid];
    if (rowok) {
#pragma unroll
        for (int c = 0; c < kGroupCols; ++c)
            if (c < ncols) W[static_cast<int64_t>(kGroupSpec + c) * f + i] = a[c];
    }
}

// The accepted panels of a group, columns [ps, ps + nacc): the region's
// final entries and W rows (its rows [ps, ps + ncols), from the region
// launch's tables) and the rows below it from W's speculation columns.
// Per row, the region launch and k_big_group carried out the operations
// of the one-panel path in its order; a row of the region at or past nacc
// went through the same operations on the accepted columns as a row below
// it would have.
__global__ __launch_bounds__(kLT) void k_big_gapply(LdltDev d, const BigLd *__restrict__ states,
                                                    const double *__restrict__ gpart, int32_t nch) {
    const BigLd B = states[blockIdx.y];
    if (!B.g_apply) return;
    const int ps = B.g_ps, nacc = B.g_n, ncols = B.g_cols;
    const FrontLd F = d.fronts[B.s];
    const int f = F.f;
    const int gend = ps + ncols;
    const int bx = blockIdx.x;
    if (bx >= max(1, (f - gend + kGroupRows - 1) / kGroupRows)) return;
    double *G = d.arena + F.fofs;
    double *W = d.work + F.wofs;
    const double *gp = gpart + blockIdx.y * gpart_stride(nch);
    const int64_t wcol0 = ps - B.w_base;
    const int tid = threadIdx.x;
    if (bx == 0) {
        for (int e = tid; e < kGroupCols * kGroupCols; e += kLT) {
            const int c = e >> 6, rw = e & 63;
            if (c < nacc && rw < ncols && c <= rw) {
                G[static_cast<int64_t>(ps + c) * f + ps + rw] = gp[kGdFin + e];
                if (c < rw) W[(wcol0 + c) * f + ps + rw] = gp[kGdWd + e];
            }
        }
    }
    const int i = gend + bx * kGroupRows + tid;
    if (i < f) {
        // eight columns' loads in flight at a time
        const double *src = W + static_cast<int64_t>(kGroupSpec) * f + i;
        double *dst = G + static_cast<int64_t>(ps) * f + i;
        for (int c0 = 0; c0 < nacc; c0 += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = src[static_cast<int64_t>(min(c0 + u, nacc - 1)) * f];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (c0 + u < nacc) dst[static_cast<int64_t>(c0 + u) * f] = v[u];
        }
    }
}

__global__ __launch_bounds__(kLT) void k_big_panel(LdltDev d, BigLd *states, Prm prm, int32_t nb,
                                                  const double *__restrict__ gpart, int32_t nch) {
    __shared__ RedScratch rs;
    __shared__ int cnt[3];
    BigLd &B = states[blockIdx.x];
    const int pass = B.pass;
    if (pass == 2) {
        // done: nothing for this iteration's update kernel
        if (threadIdx.x == 0) {
            B.u_p1 = B.u_p0;
            B.x_p1 = B.x_p0;
            B.stalled = 0;
            B.g_apply = 0;
        }
        return;
    }
    FrontLd &Fd = d.fronts[B.s];
    const int f = Fd.f, ne = Fd.ne;
    int32_t *rows = d.rows + Fd.rofs;
    int8_t *tags = d.tags + Fd.rofs;
    double *W = d.work + Fd.wofs;
    FV A{d.arena + Fd.fofs, f, f};
    const int tid = threadIdx.x;
    int k = B.k, oe = B.oe, pend0 = B.pend0, full = B.full;
    const int live = B.live;
    if (tid < 3) cnt[tid] = B.cnt[tid];
    double flops = B.flops;
    __syncthreads();
    int u_c0 = 0, u_c1 = 0, u_p0 = 0, u_p1 = 0, x_c0 = ne, x_p0 = 0, x_p1 = 0, w_base = pend0;
    int stalled = 0, npass = pass, ps = k, pl = k;
    int spec = B.spec, g_apply = 0, g_ps = 0, g_n = 0, g_cols = 0, searched = 0;
    bool group_done = false;
    const int gcols = gpart ? group_cols(B, nb) : 0;
    if (gcols > 0) {
        // decide on the group the stats launch evaluated: column maxima over
        // its workgroups and the diagonal region, then the first test per
        // pivot in order; whole panels are accepted
        __shared__ double gmax[kGroupCols], gd[kGroupCols];
        __shared__ int acc_s;
        const int nact = max(1, (f - (k + gcols) + kGroupRows - 1) / kGroupRows);
        const double *gp = gpart + blockIdx.x * gpart_stride(nch);
        {
            const int c = tid & 63, part = tid >> 6;
            double mx = 0.0;
            if (c < gcols)
                for (int w = part; w < nact; w += 4) mx = fmax(mx, gp[kGMax + w * kGroupCols + c]);
            rs.v[tid] = mx;
        }
        __syncthreads();
        if (tid < kGroupCols) {
            const double pm = fmax(fmax(rs.v[tid], rs.v[tid + 64]), fmax(rs.v[tid + 128], rs.v[tid + 192]));
            gmax[tid] = tid < gcols ? fmax(pm, gp[kGdMax + tid]) : 0.0;
            gd[tid] = tid < gcols ? gp[kGdPiv + tid] : 0.0;
        }
        __syncthreads();
        if (tid == 0) {
            int nacc = 0;
            for (int c0 = 0; c0 < gcols; c0 += 16) {
                const int c1 = min(c0 + 16, gcols);
                bool ok = true;
                for (int kk = c0; kk < c1 && ok; ++kk) {
                    const double a0 = fabs(gd[kk]), cm = gmax[kk];
                    ok = a0 >= prm.eff_small && !(prm.u * cm > a0) && !(a0 < kWeakPivotRatio * cm);
                }
                if (!ok) break;
                nacc = c1;
            }
            acc_s = nacc;
            for (int kk = 0; kk < nacc; ++kk) {
                inertia_1x1(gd[kk], prm.eff_small, cnt);
                const double rem = static_cast<double>(f - (k + kk + 1));
                flops += rem + rem * (rem + 1.0);
            }
        }
        __syncthreads();
        const int nacc = acc_s;
        if (nacc > 0) {
            for (int c = tid; c < nacc; c += kLT) tags[k + c] = 1;
            const int kn = k + nacc;
            if (kn < oe) {
                u_c0 = kn;
                u_c1 = oe;
                u_p0 = k;
                u_p1 = kn;
            }
            if (kn + nb > oe || kn >= live) {
                if (kn > pend0 && oe < ne) {
                    x_c0 = oe;
                    x_p0 = pend0;
                    x_p1 = kn;
                }
                pend0 = kn;
                oe = min(ne, kn + B.outer);
            }
            g_apply = 1;
            g_ps = k;
            g_n = nacc;
            g_cols = gcols;
            spec = nacc == gcols ? 1 : 0;  // a panel failed inside: the next panel goes the one-panel way
            ps = k + ((nacc - 1) / 16) * 16;
            pl = kn;
            k = kn;
            group_done = true;
        }
    }
    if (group_done) {
    } else if (pass == 0) {
        spec = 1;
        if (k >= live) {
            // the first pass is over: what is held back goes out, then the
            // second pass (single-level updates)
            if (k > pend0 && oe < ne) {
                x_c0 = oe;
                x_p0 = pend0;
                x_p1 = k;
            }
            pend0 = k;
            oe = ne;
            npass = 1;
        } else {
            ps = k;
            pl = min(live, ps + nb);
            while (k < pl) {
                int kind, pa, pb;
                select_pivot_wg(A, k, ne, pl, prm, rs, kind, pa, pb);
                if (kind == 0) {
                    stalled = 1;
                    break;
                }
                apply_pivot_wg(A, kind, pa, pb, k, rows, tags, W + static_cast<int64_t>(ps - pend0) * f, f, k - ps, pl,
                               prm, cnt, flops, 0);
            }
            if (k > ps && pl < oe) {
                u_c0 = pl;
                u_c1 = oe;
                u_p0 = ps;
                u_p1 = k;
            }
            // flush before pulls / rotation (they move columns past the
            // outer block), at the end of the outer block and of the pass
            if (stalled || k + nb > oe || k >= live) {
                if (k > pend0 && oe < ne) {
                    x_c0 = oe;
                    x_p0 = pend0;
                    x_p1 = k;
                }
                pend0 = k;
                oe = min(ne, k + B.outer);
            }
        }
    } else {
        oe = ne;
        if (full) {
            // a stalled second-pass panel: one pivot from a search over all
            // the remaining columns (ldlt_kernel.cpp:1119-1131)
            searched = 1;
            const int ss = k;
            int kind, pa, pb;
            select_pivot_wg(A, k, ne, ne, prm, rs, kind, pa, pb);
            if (kind == 0) {
                npass = 2;
            } else {
                apply_pivot_wg(A, kind, pa, pb, k, rows, tags, W, f, 0, ss, prm, cnt, flops, 0);
                u_c0 = k;
                u_c1 = ne;
                u_p0 = ss;
                u_p1 = k;
                w_base = ss;
                pend0 = k;
            }
            full = 0;
        } else if (k >= ne) {
            npass = 2;
        } else {
            ps = k;
            pl = min(ne, ps + nb);
            while (k < pl) {
                int kind, pa, pb;
                select_pivot_wg(A, k, ne, pl, prm, rs, kind, pa, pb);
                if (kind == 0) {
                    stalled = 1;
                    break;
                }
                apply_pivot_wg(A, kind, pa, pb, k, rows, tags, W, f, k - ps, pl, prm, cnt, flops, 0);
            }
            w_base = ps;
            if (k > ps && pl < ne) {
                u_c0 = pl;
                u_c1 = ne;
                u_p0 = ps;
                u_p1 = k;
            }
            if (stalled) full = 1;
            pend0 = k;
        }
    }
    __syncthreads();
    if (tid == 0) {
        B.k = k;
        B.pass = npass;
        B.full = full;
        B.stalled = stalled;
        B.ps = ps;
        B.pl = pl;
        B.oe = oe;
        B.pend0 = pend0;
        B.u_c0 = u_c0;
        B.u_c1 = u_c1;
        B.u_p0 = u_p0;
        B.u_p1 = u_p1;
        B.x_c0 = x_c0;
        B.x_p0 = x_p0;
        B.x_p1 = x_p1;
        B.w_base = w_base;
        B.cnt[0] = cnt[0];
        B.cnt[1] = cnt[1];
        B.cnt[2] = cnt[2];
        B.flops = flops;
        B.searches += searched;
        B.spec = spec;
        B.g_apply = g_apply;
        B.g_ps = g_ps;
        B.g_n = g_n;
        B.g_cols = g_cols;
    }
}

// The requested update and flush of every front: grid (per_front, nbig),
// tiles by grid stride.
template <int MODE>
__global__ __launch_bounds__(kLT, 2) void k_big_update(LdltDev d, const BigLd *__restrict__ states) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[big_lds_bytes<MODE>()];
    const BigLd B = states[blockIdx.y];
    if (B.prec != MODE) return;  // another precision's launch
    const FrontLd F = d.fronts[B.s];
    const double *W = d.work + F.wofs;
    const int t1 = (B.u_p1 > B.u_p0) ? big_tiles(F.f, B.u_c0, B.u_c1) : 0;
    // BF16 modes on planes: the flush is k_big_wide's
    const bool wide = MODE >= 2 && d.xpl != nullptr;
    const int t2 = (!wide && B.x_p1 > B.x_p0) ? big_tiles(F.f, B.x_c0, F.ne) : 0;
    const XcdRange xr = xcd_range(t1 + t2);
    for (int tile = xr.lo; tile < xr.hi; tile += xr.step) {
        const bool u = tile < t1;
        big_tile<MODE>(d, F, W, B.w_base, u ? B.u_c0 : B.x_c0, u ? B.u_c1 : F.ne, u ? B.u_p0 : B.x_p0,
                    u ? B.u_p1 : B.x_p1, u ? tile : tile - t1, smem);
        __syncthreads();
    }
}

// Partner pulls or rotation of a stalled first-pass panel
// (ldlt_kernel.cpp:1052-1097); every column is current (flushed).
__global__ __launch_bounds__(kLT) void k_big_stall(LdltDev d, BigLd *states, int32_t nb) {
    __shared__ RedScratch rs;
    __shared__ int32_t order[128];
    BigLd &B = states[blockIdx.x];
    if (B.pass != 0 || !B.stalled) return;
    FrontLd &Fd = d.fronts[B.s];
    const int f = Fd.f, ne = Fd.ne;
    int32_t *rows = d.rows + Fd.rofs;
    FV A{d.arena + Fd.fofs, f, f};
    const int tid = threadIdx.x;
    const int k = B.k, pl = B.pl;
    int live = B.live;
    if (k != B.k_paired) {
        const int np = partner_pairs_wg(A, k, pl, ne, nb, rows, order, rs);
        if (np > 0) {
            place_rows_wg(A, k, ne, np, order, rows, rs);
            if (tid == 0) {
                B.live = max(live, k + np);
                B.k_paired = k;
                B.stalled = 0;
                B.stalls += 1;
                B.pulls += np / 2;
            }
            return;
        }
    }
    const int m = pl - k;
    const int mv = min(m, live - pl);
    for (int t = 0; t < mv; ++t) sym_swap_wg(A, k + t, live - mv + t, rows, nullptr, 0, 0, 0);
    if (tid == 0) {
        B.live = live - m;
        B.stalled = 0;
        B.stalls += 1;
        B.rotated += m;
    }
}

// L*D columns [c0, c1) of every done front, rows [ne, f), into W (the
// column order of schur_update's fill_ld_column, ldlt_kernel.cpp:764-816)
__global__ __launch_bounds__(kLT) void k_big_ld(LdltDev d, const BigLd *__restrict__ states, int32_t c0,
                                               int32_t width, int32_t wide) {
    const BigLd B = states[blockIdx.y];
    if ((B.prec >= 2 && d.xpl != nullptr) != (wide != 0)) return;  // the other chunk width's fronts
    const FrontLd F = d.fronts[B.s];
    const int f = F.f, ne = F.ne, p = B.k;
    const int c1 = min(p, c0 + width);
    if (c1 <= c0 || ne >= f) return;
    const double *G = d.arena + F.fofs;
    const int8_t *tags = d.tags + F.rofs;
    double *W = d.work + F.wofs;
    const int q = f - ne;
    const int64_t tot = static_cast<int64_t>(c1 - c0) * q;
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * kLT + threadIdx.x; idx < tot;
         idx += static_cast<int64_t>(gridDim.x) * kLT) {
        const int c = c0 + static_cast<int>(idx / q);
        const int j = ne + static_cast<int>(idx % q);
        const int8_t tg = tags[c];
        double v = 0.0;
        if (tg == 1) {
            v = G[static_cast<int64_t>(c) * f + j] * G[static_cast<int64_t>(c) * f + c];
        } else if (tg == 2) {
            v = __dadd_rn(G[static_cast<int64_t>(c) * f + j] * G[static_cast<int64_t>(c) * f + c],
                          G[static_cast<int64_t>(c + 1) * f + j] * G[static_cast<int64_t>(c) * f + c + 1]);
        } else if (tg == -1) {
            v = __dadd_rn(G[static_cast<int64_t>(c - 1) * f + j] * G[static_cast<int64_t>(c - 1) * f + c],
                          G[static_cast<int64_t>(c) * f + j] * G[static_cast<int64_t>(c) * f + c]);
        }
        W[static_cast<int64_t>(c - c0) * f + j] = v;
    }
}

template <int MODE>
__global__ __launch_bounds__(kLT, 2) void k_big_cbupdate(LdltDev d, const BigLd *__restrict__ states, int32_t c0,
                                                     int32_t width) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[big_lds_bytes<MODE>()];
    const BigLd B = states[blockIdx.y];
    if (B.prec != MODE) return;
    const FrontLd F = d.fronts[B.s];
    const int c1 = min(B.k, c0 + width);
    if (c1 <= c0 || F.ne >= F.f) return;
    const double *W = d.work + F.wofs;
    const XcdRange xr = xcd_range(big_tiles(F.f, F.ne, F.f));
    for (int tile = xr.lo; tile < xr.hi; tile += xr.step) {
        big_tile<MODE>(d, F, W, c0, F.ne, F.f, c0, c1, tile, smem);
        __syncthreads();
    }
}

// 256 x 256 tile of the rank-kBigWork updates on the bf16 planes (modes 2,
// 3): C[i][j] -= sum_c A[i][c] B[j][c] over plane columns [0, np), lower
// part of rows [c0, f) x columns [c0, c1) of front F, tile `tile` in
// super-block order.
template <int MODE>
__device__ __forceinline__ void big_tile_w(const LdltDev &d, const FrontLd &F, const uint16_t *pl, int c0, int c1,
                                           int np, int tile, unsigned char *smem) {
    const int f = F.f;
    int ib, jb;
    super_block_tile((f - c0 + kTW - 1) / kTW, (c1 - c0 + kTW - 1) / kTW, tile, ib, jb);
    const int i0 = c0 + ib * kTW, j0 = c0 + jb * kTW;
    double *G = d.arena + F.fofs;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = w & 3, wc = w >> 2;
    const bool active = (i0 + wr * 64 + 63) >= (j0 + wc * 128);
    v4f_t acc[8][4];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = v4f_t{0.f, 0.f, 0.f, 0.f};
    const int64_t ps = static_cast<int64_t>(kBigWork) * f;
    tile_loop_bf16p_w<MODE == 3, kBigKcBf>(pl, pl + ps, pl + 2 * ps, pl + 3 * ps, f, i0, j0, f, c1, 0, np, active, smem,
                                           acc);
    if (!active) return;
    const int ib0 = i0 + wr * 64, jb0 = j0 + wc * 128;
    if (ib0 + 63 < f && jb0 + 127 < c1 && ib0 >= jb0 + 127) {
        // the wave's 64 x 128 block lies inside the front's lower triangle:
        // no masks, and two column blocks (32 entries per lane) at a time,
        // their loads in flight together before their stores
        double *base = G + static_cast<int64_t>(jb0 + 4 * (lane >> 4)) * f + ib0 + (lane & 15);
#pragma unroll
        for (int m0 = 0; m0 < 8; m0 += 2) {
            double old[2][4][4];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const double *col = base + static_cast<int64_t>((m0 + h) * 16 + v) * f;
#pragma unroll
                    for (int nn = 0; nn < 4; ++nn) old[h][nn][v] = col[nn * 16];
                }
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double *col = base + static_cast<int64_t>((m0 + h) * 16 + v) * f;
#pragma unroll
                    for (int nn = 0; nn < 4; ++nn) col[nn * 16] = old[h][nn][v] - acc[m0 + h][nn][v];
                }
        }
        return;
    }
    double *junk = d.junk + threadIdx.x;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        double *dst[4][4];
        double old[4][4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = jb0 + m * 16 + 4 * (lane >> 4) + v;
            double *col = G + static_cast<int64_t>(min(j, c1 - 1)) * f;
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) {
                const int i = ib0 + nn * 16 + (lane & 15);
                dst[nn][v] = (i < f && j < c1 && i >= j) ? col + i : junk;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) old[nn][v] = *dst[nn][v];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) *dst[nn][v] = old[nn][v] - acc[m][nn][v];
    }
}

__device__ __forceinline__ int big_tiles_w(int f, int c0, int c1) {
    if (c1 <= c0) return 0;
    const int nr = (f - c0 + kTW - 1) / kTW, nc = (c1 - c0 + kTW - 1) / kTW;
    return nc * nr - nc * (nc - 1) / 2;
}

// The rank-kBigWork updates of BF16 / split-BF16 fronts on their planes, in
// 256 x 256 tiles: CB, the contribution block's chunk at c0 (after
// k_big_ld and k_big_pack<true>); else the flush request (after
// k_big_pack<false>).
template <int MODE, bool CB>
__global__ __launch_bounds__(kWideThreads) void k_big_wide(LdltDev d, const BigLd *__restrict__ states, int32_t c0,
                                                           int32_t width) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[bfw_lds_bytes<MODE == 3, kBigKcBf>()];
    const BigLd B = states[blockIdx.y];
    if (B.prec != MODE) return;
    const FrontLd F = d.fronts[B.s];
    int r0, rc, np;
    if (CB) {
        const int c1 = min(B.k, c0 + width);
        if (c1 <= c0 || F.ne >= F.f) return;
        r0 = F.ne;
        rc = F.f;
        np = c1 - c0;
    } else {
        if (B.x_p1 <= B.x_p0) return;
        r0 = B.x_c0;
        rc = F.ne;
        np = B.x_p1 - B.x_p0;
    }
    const uint16_t *pl = d.xpl + 4 * F.wofs;
    const XcdRange xr = xcd_range(big_tiles_w(F.f, r0, rc));
    for (int tile = xr.lo; tile < xr.hi; tile += xr.step) {
        big_tile_w<MODE>(d, F, pl, r0, rc, np, tile, smem);
        __syncthreads();
    }
}

// bf16 planes of a rank-kBigWork update's operands in the packed layout of
// tile_loop_bf16p (row i, plane column c at (c / 8) f 8 + 8 i + c % 8), for
// the fronts of BF16 (high part) or split-BF16 (both parts) precision:
// A = front columns [p0, p0 + np), rows [r0, f); B = W columns from p0 - w0,
// rows [r0, rb); plane columns [np, np rounded up to 8) zero. CB: the
// contribution-block chunk at c0 (W from k_big_ld), else the flush request.
template <bool CB>
__global__ __launch_bounds__(kLT) void k_big_pack(LdltDev d, const BigLd *__restrict__ states, int32_t c0,
                                                  int32_t width) {
    const BigLd B = states[blockIdx.y];
    if (B.prec < 2) return;
    const FrontLd F = d.fronts[B.s];
    const int f = F.f;
    int p0, np, w0, r0, rb;
    if (CB) {
        const int c1 = min(B.k, c0 + width);
        if (c1 <= c0 || F.ne >= f) return;
        p0 = c0;
        np = c1 - c0;
        w0 = c0;
        r0 = F.ne;
        rb = f;
    } else {
        if (B.x_p1 <= B.x_p0) return;
        p0 = B.x_p0;
        np = B.x_p1 - B.x_p0;
        w0 = B.w_base;
        r0 = B.x_c0;
        rb = F.ne;
    }
    const bool split = B.prec == 3;
    const double *G = d.arena + F.fofs;
    const double *W = d.work + F.wofs;
    uint16_t *pl = d.xpl + 4 * F.wofs;
    const int64_t ps = static_cast<int64_t>(kBigWork) * f;
    const int ng = (np + 7) / 8;
    const int nr = f - r0;
    const int64_t tot = 2 * static_cast<int64_t>(ng) * nr;
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * kLT + threadIdx.x; idx < tot;
         idx += static_cast<int64_t>(gridDim.x) * kLT) {
        const int i = r0 + static_cast<int>(idx % nr);
        const int64_t q = idx / nr;
        const int g = static_cast<int>(q % ng);
        const int op = static_cast<int>(q / ng);  // 0: A, 1: B
        if (op == 1 && i >= rb) continue;
        const double *src = op == 0 ? G + static_cast<int64_t>(p0) * f : W + static_cast<int64_t>(p0 - w0) * f;
        v8bf_t hi, lo;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int c = 8 * g + u;
            const float x = c < np ? static_cast<float>(src[static_cast<int64_t>(c) * f + i]) : 0.f;
            hi[u] = static_cast<__bf16>(x);
            lo[u] = static_cast<__bf16>(x - static_cast<float>(hi[u]));
        }
        const int64_t o = static_cast<int64_t>(op) * 2 * ps + static_cast<int64_t>(g) * f * 8 + static_cast<int64_t>(i) * 8;
        *reinterpret_cast<v8bf_t *>(pl + o) = hi;
        if (split) *reinterpret_cast<v8bf_t *>(pl + o + ps) = lo;
    }
}

__global__ void k_big_finish(LdltDev d, const BigLd *__restrict__ states, int32_t nbig) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbig) return;
    const BigLd B = states[b];
    FrontLd &F = d.fronts[B.s];
    F.p = B.k;
    F.npos = B.cnt[0];
    F.nneg = B.cnt[1];
    F.nzero = B.cnt[2];
    F.flops = B.flops;
    F.stalls = B.stalls;
    F.pulls = B.pulls;
    F.rotated = B.rotated;
    F.searches = B.searches;
}

void launch_big_panel(hipStream_t st, const LdltDev &d, BigLd *states, int32_t nbig, double u, double eff_small,
                      double small_pivot, int32_t nb, const double *gpart, int32_t nch) {
    Prm prm{u, eff_small, small_pivot};
    k_big_panel<<<dim3(nbig), kLT, 0, st>>>(d, states, prm, nb, gpart, nch);
}

void launch_big_group_stats(hipStream_t st, const LdltDev &d, const BigLd *states, int32_t nbig, double *gpart,
                            int32_t nch) {
    k_big_gdiag<<<dim3(nbig), 64, 0, st>>>(d, states, gpart, nch);
    k_big_group<<<dim3(nch, nbig), kLT, 0, st>>>(d, states, gpart, nch);
}

void launch_big_group_apply(hipStream_t st, const LdltDev &d, const BigLd *states, int32_t nbig, double *gpart,
                            int32_t nch) {
    k_big_gapply<<<dim3(nch, nbig), kLT, 0, st>>>(d, states, gpart, nch);
}

// workgroups per front of the 256 x 256 tiles (one per CU; a multiple of 8)
static int32_t wide_per_front(int32_t per_front) { return std::max(8, per_front / 16 * 8); }

void launch_big_update(hipStream_t st, const LdltDev &d, const BigLd *states, int32_t nbig, int32_t per_front,
                       const int32_t *nmode) {
    if (nbig <= 0) return;
    const dim3 g(per_front, nbig);
    if (d.xpl != nullptr && nmode[2] + nmode[3] > 0) k_big_pack<false><<<g, kLT, 0, st>>>(d, states, 0, 0);
    if (nmode[0] > 0) k_big_update<0><<<g, kLT, 0, st>>>(d, states);
    if (nmode[1] > 0) k_big_update<1><<<g, kLT, 0, st>>>(d, states);
    if (nmode[2] > 0) k_big_update<2><<<g, kLT, 0, st>>>(d, states);
    if (nmode[3] > 0) k_big_update<3><<<g, kLT, 0, st>>>(d, states);
    if (d.xpl != nullptr) {
        const dim3 gw(wide_per_front(per_front), nbig);
        if (nmode[2] > 0) k_big_wide<2, false><<<gw, kWideThreads, 0, st>>>(d, states, 0, 0);
        if (nmode[3] > 0) k_big_wide<3, false><<<gw, kWideThreads, 0, st>>>(d, states, 0, 0);
    }
}

void launch_big_stall(hipStream_t st, const LdltDev &d, BigLd *states, int32_t nbig, int32_t nb) {
    if (nbig <= 0) return;
    k_big_stall<<<dim3(nbig), kLT, 0, st>>>(d, states, nb);
}

void launch_big_cb(hipStream_t st, const LdltDev &d, const BigLd *states, int32_t nbig, int32_t c0, int32_t width,
                   int32_t per_front, const int32_t *nmode, bool wide) {
    if (nbig <= 0) return;
    const dim3 g(per_front, nbig);
    k_big_ld<<<g, kLT, 0, st>>>(d, states, c0, width, wide ? 1 : 0);
    if (wide) {
        // the BF16 modes' fronts on planes
        const dim3 gw(wide_per_front(per_front), nbig);
        k_big_pack<true><<<g, kLT, 0, st>>>(d, states, c0, width);
        if (nmode[2] > 0) k_big_wide<2, true><<<gw, kWideThreads, 0, st>>>(d, states, c0, width);
        if (nmode[3] > 0) k_big_wide<3, true><<<gw, kWideThreads, 0, st>>>(d, states, c0, width);
        return;
    }
    if (nmode[0] > 0) k_big_cbupdate<0><<<g, kLT, 0, st>>>(d, states, c0, width);
    if (nmode[1] > 0) k_big_cbupdate<1><<<g, kLT, 0, st>>>(d, states, c0, width);
    if (d.xpl == nullptr) {
        if (nmode[2] > 0) k_big_cbupdate<2><<<g, kLT, 0, st>>>(d, states, c0, width);
        if (nmode[3] > 0) k_big_cbupdate<3><<<g, kLT, 0, st>>>(d, states, c0, width);
    }
}

void launch_big_finish(hipStream_t st, const LdltDev &d, const BigLd *states, int32_t nbig) {
    if (nbig <= 0) return;
    k_big_finish<<<dim3((nbig + 63) / 64), 64, 0, st>>>(d, states, nbig);
}

// ---------------------------------------------------------------------------
// Assembly of the level's fronts (f x f squares in the arena, lower
// triangle). One workgroup per front column, rows from the diagonal down:
// zero, original entries of the front's own pivot columns, then each
// child's CB column in child order; the CB of a child is the trailing
// (f_c - p_c) square of its own front.
constexpr int kAsmChunkL = 2048;

__device__ __forceinline__ int lbound(const int32_t *a, int lo, int hi, int32_t key) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kLT) void k_ldlt_assemble(LdltDev d, const AsmTask *__restrict__ tasks,
                                                      const ChildCol *__restrict__ cc) {
    __shared__ double buf[kAsmChunkL];
    const AsmTask t = tasks[blockIdx.x];
    const FrontLd F = d.fronts[t.s];
    const int c = t.c;
    const int f = F.f;
    double *colp = d.arena + F.fofs + static_cast<int64_t>(c) * f;
    const int tid = threadIdx.x;
    // the strict upper triangle is never read: left as it is
    const int oc = c - F.nd;  // own pivot column index, when in [0, psym)
    for (int r0 = c; r0 < f; r0 += kAsmChunkL) {
        const int r1 = min(f, r0 + kAsmChunkL);
        for (int r = r0 + tid; r < r1; r += kLT) buf[r - r0] = 0.0;
        __syncthreads();
        if (oc >= 0 && oc < F.psym) {
            const int j = F.col0 + oc;
            const int k0 = d.pcol_ptr[j], k1 = d.pcol_ptr[j + 1];
            for (int k = k0 + tid; k < k1; k += kLT) {
                const int lr = F.nd + d.amap[k];
                if (lr >= r0 && lr < r1) buf[lr - r0] += d.av[k];
            }
            __syncthreads();
        }
        for (int e = t.cc0; e < t.cc1; ++e) {
            const ChildCol ch = cc[e];
            const FrontLd C = d.fronts[ch.ch];
            const int q = C.f - C.p;
            const int32_t *cm = d.cmap + C.cmofs;
            const double *cbcol = d.arena + C.fofs + static_cast<int64_t>(C.p + ch.jj) * C.f + C.p;
            const int lo = lbound(cm, ch.jj, q, r0);
            const int hi = lbound(cm, lo, q, r1);
            for (int ii = lo + tid; ii < hi; ii += kLT) buf[cm[ii] - r0] += cbcol[ii];
            __syncthreads();
        }
        for (int r = r0 + tid; r < r1; r += kLT) colp[r] = buf[r - r0];
        __syncthreads();
    }
}

void launch_ldlt_assemble(hipStream_t st, const LdltDev &d, const AsmTask *tasks, int64_t ntasks, const ChildCol *cc) {
    if (ntasks <= 0) return;
    for (int64_t o = 0; o < ntasks; o += kMaxLaunchGroups)
        k_ldlt_assemble<<<dim3(static_cast<unsigned>(std::min(kMaxLaunchGroups, ntasks - o))), kLT, 0, st>>>(
            d, tasks + o, cc);
}

// ---------------------------------------------------------------------------
// Front epilogue (multifrontal_serial.cpp:892-945): a root's unaccepted
// eligible columns become zero pivots (tag 0); D^-1 per pivot as
// build_inline_dinv (solve_diag.cpp:49-123): kinds 1 / 2 / 0. The stored
// factor columns keep D on the diagonal, a 2x2 block's off-diagonal moved
// to the strict upper slot (c, c+1), zeros elsewhere above the diagonal
// and in a zero pivot's column; the square itself is left as it is (its
// trailing part is the contribution block).
__device__ __forceinline__ double stored_entry(const double *G, int f, int c, int i, int8_t tg) {
    if (i < c) return (tg == -1 && i == c - 1) ? G[static_cast<int64_t>(c - 1) * f + c] : 0.0;
    if (tg == 0 || (tg == 2 && i == c + 1)) return 0.0;
    return G[static_cast<int64_t>(c) * f + i];
}

// The epilogue of one front whose square (leading dimension f) is at G, in
// global memory or LDS; store: also write its stored factor columns.
template <int NT = kLT>
__device__ void finish_front_wg(const LdltDev &d, FrontLd &F, const double *G, bool store) {
    const int f = F.f;
    int p = F.p;
    int8_t *tags = d.tags + F.rofs;
    const int tid = threadIdx.x;
    const bool root = F.root != 0;
    int ps = p;
    if (root && p < F.ne) {
        for (int c = p + tid; c < F.ne; c += NT) tags[c] = 0;
        ps = F.ne;
    }
    __syncthreads();
    if (tid == 0) {
        F.pstore = ps;
        if (root && p < F.ne) F.nzero += F.ne - p;
        int n2 = 0;
        for (int c = 0; c < ps; ++c) n2 += tags[c] == 2;
        F.n2x2 = n2;
        F.delayed = root ? 0 : F.ne - p;
    }
    for (int c = tid; c < ps; c += NT) {
        const double *Lc = G + static_cast<int64_t>(c) * f;
        const int8_t tg = tags[c];
        const double d00 = Lc[c];
        double *dv = d.dinv + 3 * (F.dofs + c);
        int8_t *dk = d.dkind + F.dofs + c;
        if (tg == 2) {
            const double d10 = Lc[c + 1];
            const double d11 = G[static_cast<int64_t>(c + 1) * f + c + 1];
            double sc = fabs(d00);
            if (sc < fabs(d10)) sc = fabs(d10);
            if (sc < fabs(d11)) sc = fabs(d11);
            double det_s = 0.0;
            if (isfinite(sc) && sc != 0.0) det_s = __dsub_rn(__dmul_rn(d00 / sc, d11 / sc), __dmul_rn(d10 / sc, d10 / sc));
            if (sc == 0.0 || !isfinite(det_s) || det_s == 0.0) {
                dv[0] = dv[1] = dv[2] = 0.0;
            } else {
                dv[0] = (d11 / sc) / det_s / sc;
                dv[1] = -(d10 / sc) / det_s / sc;
                dv[2] = (d00 / sc) / det_s / sc;
            }
            *dk = 2;
        } else if (tg == -1) {
            dv[0] = dv[1] = dv[2] = 0.0;
            *dk = 0;
        } else if (tg == 1) {
            dv[0] = (d00 == 0.0 || !isfinite(d00)) ? 0.0 : 1.0 / d00;
            dv[1] = dv[2] = 0.0;
            *dk = 1;
        } else {
            dv[0] = dv[1] = dv[2] = 0.0;
            *dk = 1;
        }
    }
    if (!store) return;
    double *O = d.fac + F.lofs;
    const int64_t tot = static_cast<int64_t>(ps) * f;
    for (int64_t idx = tid; idx < tot; idx += NT) {
        const int c = static_cast<int>(idx / f), i = static_cast<int>(idx - static_cast<int64_t>(c) * f);
        O[idx] = stored_entry(G, f, c, i, tags[c]);
    }
}

__global__ __launch_bounds__(kLT) void k_ldlt_epilogue(LdltDev d, const int32_t *__restrict__ flist,
                                                       int32_t n_inline) {
    FrontLd &F = d.fronts[flist[blockIdx.x]];
    finish_front_wg(d, F, d.arena + F.fofs, static_cast<int>(blockIdx.x) < n_inline);
}

// ---------------------------------------------------------------------------
// Levels of small fronts laid out on the device.
//
// Layout: one thread per front. Its size follows from the children's
// results (nd = their delayed columns, f = nd + symbolic rows, ne = nd +
// symbolic pivots, as the host layout); its four offsets are a workgroup
// scan plus one bump of each counter. A front larger than L.fmax, or a
// bump past the room, refuses the level (atomicMin of the level index).
__global__ __launch_bounds__(kLT) void k_dev_layout(LdltDev d, DevLevel L, const int32_t *__restrict__ flist,
                                                    int32_t nf, int32_t lv) {
    __shared__ long long sc[4][kLT];
    __shared__ unsigned long long base[4];
    __shared__ int refuse;
    // after a refused level nothing more is laid out
    if (__atomic_load_n(&L.ctr[kDcLevel], __ATOMIC_RELAXED) < static_cast<unsigned long long>(lv)) return;
    const int tid = threadIdx.x;
    const int i = blockIdx.x * kLT + tid;
    FrontLd F{};
    long long sz[4] = {0, 0, 0, 0};
    bool big = false;
    if (tid == 0) refuse = 0;
    if (i < nf) {
        const int s = flist[i];
        const FrontDev Sy = L.sym[s];
        int nd = 0;
        for (int c = L.child_ptr[s]; c < L.child_ptr[s + 1]; ++c) nd += d.fronts[L.child_idx[c]].delayed;
        F.nd = nd;
        F.psym = Sy.p;
        F.col0 = Sy.col0;
        F.f = nd + Sy.f;
        F.ne = nd + Sy.p;
        F.root = L.parent[s] < 0 ? 1 : 0;
        F.dev = 1;
        big = F.f > L.fmax;
        const long long f = F.f;
        sz[0] = (f * f + 63) / 64 * 64;
        sz[1] = f * F.ne;
        sz[2] = f;
        sz[3] = F.ne;
    }
    for (int k = 0; k < 4; ++k) sc[k][tid] = sz[k];
    __syncthreads();
    // inclusive scan (Hillis-Steele)
    for (int w = 1; w < kLT; w <<= 1) {
        long long v[4];
        for (int k = 0; k < 4; ++k) v[k] = tid >= w ? sc[k][tid - w] : 0;
        __syncthreads();
        for (int k = 0; k < 4; ++k) sc[k][tid] += v[k];
        __syncthreads();
    }
    if (tid == 0) {
        for (int k = 0; k < 4; ++k) {
            const unsigned long long tot = static_cast<unsigned long long>(sc[k][kLT - 1]);
            base[k] = atomicAdd(&L.ctr[k], tot);
            if (static_cast<long long>(base[k] + tot) > L.cap[k]) refuse = 1;
        }
    }
    __syncthreads();
    if (refuse && tid == 0) {
        atomicMin(&L.ctr[kDcLevel], static_cast<unsigned long long>(lv));
        atomicOr(&L.ctr[kDcWhy], kDevWhyRoom);
    }
    if (i >= nf) return;
    if (big) {
        atomicMin(&L.ctr[kDcLevel], static_cast<unsigned long long>(lv));
        atomicOr(&L.ctr[kDcWhy], kDevWhySize);
    }
    F.fofs = static_cast<int64_t>(base[0] + (sc[0][tid] - sz[0]));
    F.lofs = static_cast<int64_t>(base[1] + (sc[1][tid] - sz[1]));
    F.rofs = static_cast<int64_t>(base[2] + (sc[2][tid] - sz[2]));
    F.dofs = static_cast<int64_t>(base[3] + (sc[3][tid] - sz[3]));
    d.fronts[flist[i]] = F;
}

void launch_dev_layout(hipStream_t st, const LdltDev &d, const DevLevel &L, const int32_t *flist, int32_t nf,
                       int32_t lv) {
    if (nf <= 0) return;
    k_dev_layout<<<dim3((nf + kLT - 1) / kLT), kLT, 0, st>>>(d, L, flist, nf, lv);
}

// One front per workgroup, the square in LDS: row list (children's delayed
// columns in child order, then the symbolic rows), assembly (zero, the
// front's original entries, each child's contribution block in child
// order: per entry the host layout's order of sums), the unblocked
// factorization, the epilogue, and the contribution block written back to
// the front's square for the parent.
template <int NT, int FMIN, int FMAX>
__global__ __launch_bounds__(NT) void k_dev_fronts(LdltDev d, DevLevel L, const int32_t *__restrict__ flist, int32_t lv,
                                                   Prm prm) {
    __shared__ double Fs[FMAX * FMAX];
    __shared__ RedScratchT<NT> rs;
    __shared__ int cnt[3];
    __shared__ int32_t cpos[FMAX];
    __shared__ int32_t cptr[FMAX + 1];
    // a refused level is left whole to the host
    if (__atomic_load_n(&L.ctr[kDcLevel], __ATOMIC_RELAXED) <= static_cast<unsigned long long>(lv)) return;
    const int s = flist[blockIdx.x];
    FrontLd &Fd = d.fronts[s];
    const FrontLd F = Fd;
    const FrontDev Sy = L.sym[s];
    const int f = F.f, ne = F.ne, nd = F.nd;
    if (f <= FMIN || f > FMAX) return;  // the other size class's launch
    int32_t *rows = d.rows + F.rofs;
    int8_t *tags = d.tags + F.rofs;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < f * f; idx += NT) Fs[idx] = 0.0;
    for (int t = tid; t < f; t += NT) tags[t] = 0;
    for (int t = tid; t < Sy.f; t += NT) rows[nd + t] = L.sym_rows[Sy.rofs + t];
    if (tid < 3) cnt[tid] = 0;
    __syncthreads();
    // original entries of the own pivot columns (one per row and column):
    // the columns' entries are contiguous; their pointers go to LDS once and
    // each lane finds the column of its entries there
    for (int t = tid; t <= F.psym; t += NT) cptr[t] = d.pcol_ptr[F.col0 + t];
    __syncthreads();
    {
        const int e0 = cptr[0], e1 = cptr[F.psym];
        for (int e = e0 + tid; e < e1; e += NT) {
            int lo = 0, hi = F.psym;  // the column oc with cptr[oc] <= e < cptr[oc + 1]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (cptr[mid] <= e) lo = mid;
                else hi = mid;
            }
            Fs[(nd + lo) * f + nd + d.amap[e]] += d.av[e];
        }
    }
    __syncthreads();
    int dp = 0;
    for (int c = L.child_ptr[s]; c < L.child_ptr[s + 1]; ++c) {
        const int ch = L.child_idx[c];
        const FrontLd C = d.fronts[ch];
        const int q = C.f - C.p, del = C.delayed;
        const int32_t *crow = d.rows + C.rofs + C.p;
        const int32_t *sm = L.sym_cmap + L.sym_cmofs[ch];
        for (int ii = tid; ii < q; ii += NT) cpos[ii] = ii < del ? dp + ii : nd + sm[ii - del];
        for (int t = tid; t < del; t += NT) rows[dp + t] = crow[t];
        __syncthreads();
        const double *cb = d.arena + C.fofs + static_cast<int64_t>(C.p) * C.f + C.p;
        for (int idx = tid; idx < q * q; idx += NT) {
            const int jj = idx / q, ii = idx - jj * q;
            if (ii >= jj) Fs[cpos[jj] * f + cpos[ii]] += cb[static_cast<int64_t>(jj) * C.f + ii];
        }
        __syncthreads();
        dp += del;
    }
    FV A{Fs, f, f};
    int k = 0;
    double flops = 0.0;
    Prm fp;
    const double sp = front_prm(prm, F, fp);
    int nstatic = 0;
    double growth = 0.0;
    while (k < ne) {
        int kind, pa, pb;
        select_pivot_wg(A, k, ne, ne, fp, rs, kind, pa, pb);
        bool az = false;
        if (kind == 0) {
            if (!(sp > 0.0)) break;
            int raised;
            pa = force_static_wg(A, k, ne, sp, fp, rs, raised, nstatic, growth);
            kind = 1;
            az = static_as_zero(raised, fp, F);
        }
        apply_pivot_wg<NT>(A, kind, pa, pb, k, rows, tags, nullptr, 0, 0, f, fp, cnt, flops, 0, az);
    }
    __syncthreads();
    if (tid == 0) {
        Fd.p = k;
        Fd.npos = cnt[0];
        Fd.nneg = cnt[1];
        Fd.nzero = cnt[2];
        Fd.flops = flops;
        Fd.nstatic = nstatic;
        Fd.growth = growth;
        Fd.stalls = Fd.pulls = Fd.rotated = Fd.searches = 0;  // unblocked: no panels
    }
    __syncthreads();
    finish_front_wg<NT>(d, Fd, Fs, true);
    if (F.root) return;
    // contribution block (lower part of columns [k, f)) for the parent
    double *G = d.arena + F.fofs;
    const int q = f - k;
    for (int idx = tid; idx < q * q; idx += NT) {
        const int jj = idx / q, ii = idx - jj * q;
        if (ii >= jj) G[static_cast<int64_t>(k + jj) * f + k + ii] = Fs[(k + jj) * f + k + ii];
    }
}

void launch_dev_fronts(hipStream_t st, const LdltDev &d, const DevLevel &L, const int32_t *flist, int32_t nf,
                       int32_t lv, double u, double eff_small, double small_pivot, const StaticPrm &S, bool wave) {
    if (nf <= 0) return;
    Prm prm{u, eff_small, small_pivot, 0.0, S};
    if (!wave) {
        // every front on kLT threads (shorter per-front latency where a
        // level has few fronts)
        for (int64_t o = 0; o < nf; o += kMaxLaunchGroups)
            k_dev_fronts<kLT, 0, 95><<<dim3(static_cast<unsigned>(std::min<int64_t>(kMaxLaunchGroups, nf - o))), kLT, 0,
                                        st>>>(d, L, flist + o, lv, prm);
        return;
    }
    // two size classes: one wavefront and a 48 x 48 square (many fronts per
    // CU), a workgroup of kLT and the largest square
    static_assert(kDevFrontMax == 95, "size classes");
    for (int64_t o = 0; o < nf; o += kMaxLaunchGroups) {
        const dim3 grid(static_cast<unsigned>(std::min<int64_t>(kMaxLaunchGroups, nf - o)));
        k_dev_fronts<64, 0, 48><<<grid, 64, 0, st>>>(d, L, flist + o, lv, prm);
        k_dev_fronts<kLT, 48, 95><<<grid, kLT, 0, st>>>(d, L, flist + o, lv, prm);
    }
}

// Stored columns of large fronts: grid (column stride, fronts).
__global__ __launch_bounds__(kLT) void k_ldlt_store(LdltDev d, const int32_t *__restrict__ flist) {
    const FrontLd F = d.fronts[flist[blockIdx.y]];
    const int f = F.f;
    const int8_t *tags = d.tags + F.rofs;
    const double *G = d.arena + F.fofs;
    double *O = d.fac + F.lofs;
    for (int c = blockIdx.x; c < F.pstore; c += gridDim.x) {
        const int8_t tg = tags[c];
        for (int i = threadIdx.x; i < f; i += kLT) O[static_cast<int64_t>(c) * f + i] = stored_entry(G, f, c, i, tg);
    }
}

__global__ void k_front_patch(FrontLd *fronts, const FrontLd *patch, const int32_t *ids, int32_t np) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) fronts[ids[i]] = patch[i];
}

__global__ void k_front_gather(const FrontLd *fronts, FrontLd *out, const int32_t *ids, int32_t np) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) out[i] = fronts[ids[i]];
}

void launch_front_gather(hipStream_t st, const FrontLd *fronts, FrontLd *out, const int32_t *ids, int32_t np) {
    if (np <= 0) return;
    k_front_gather<<<dim3((np + 127) / 128), 128, 0, st>>>(fronts, out, ids, np);
}

void launch_front_patch(hipStream_t st, FrontLd *fronts, const FrontLd *patch, const int32_t *ids, int32_t np) {
    if (np <= 0) return;
    k_front_patch<<<dim3((np + 127) / 128), 128, 0, st>>>(fronts, patch, ids, np);
}

void launch_ldlt_epilogue(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, int32_t n_inline) {
    if (nf <= 0) return;
    for (int64_t o = 0; o < nf; o += kMaxLaunchGroups) {
        const int64_t m = std::min<int64_t>(kMaxLaunchGroups, nf - o);
        const int32_t inl = static_cast<int32_t>(std::max<int64_t>(0, std::min<int64_t>(m, n_inline - o)));
        k_ldlt_epilogue<<<dim3(static_cast<unsigned>(m)), kLT, 0, st>>>(d, flist + o, inl);
    }
}

void launch_ldlt_store(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, int32_t max_cols) {
    if (nf <= 0 || max_cols <= 0) return;
    k_ldlt_store<<<dim3(max_cols < 2048 ? max_cols : 2048, nf), kLT, 0, st>>>(d, flist);
}

// ---------------------------------------------------------------------------
// D^-1 step of the LDL^T solve (apply_dinv, solve_backward.cpp:156-173): per pivot,
// kind 1 scales, kind 2 applies the 2x2 inverse, kind 0 (second column of a
// 2x2) is skipped. The L sweeps are in kernels_solve.hip.
__global__ void k_ldlt_diag(LdltDev d, const int32_t *__restrict__ flist, int32_t nf, double *__restrict__ y,
                            int32_t n, int32_t nrhs) {
    const int fi = blockIdx.x;
    if (fi >= nf) return;
    const FrontLd F = d.fronts[flist[fi]];
    const int32_t *rows = d.rows + F.rofs;
    for (int c = threadIdx.x; c < F.pstore; c += blockDim.x) {
        const int8_t kd = d.dkind[F.dofs + c];
        const double *dv = d.dinv + 3 * (F.dofs + c);
        for (int r = 0; r < nrhs; ++r) {
            double *yr = y + static_cast<int64_t>(r) * n;
            if (kd == 1) {
                yr[rows[c]] *= dv[0];
            } else if (kd == 2) {
                const double r0 = yr[rows[c]], r1 = yr[rows[c + 1]];
                yr[rows[c]] = dv[0] * r0 + dv[1] * r1;
                yr[rows[c + 1]] = dv[1] * r0 + dv[2] * r1;
            }
        }
    }
}

// The diagonal-only solve job (solve_diag, solve_diag.cpp:236-261, over
// the entries process_front pushes, multifrontal_serial.cpp:921-942): D
// read from the stored columns (d on the diagonal, a 2x2 block's
// off-diagonal in slot (c, c+1)); a 1x1 pivot divides (a zero or
// non-finite d, and a root's zero pivot, give 0), a 2x2 block takes
// solve_2x2_scaled (pivoting.cpp:30-62).
__global__ void k_ldlt_diag_div(LdltDev d, const int32_t *__restrict__ flist, int32_t nf, double *__restrict__ y,
                                int32_t n, int32_t nrhs) {
    const int fi = blockIdx.x;
    if (fi >= nf) return;
    const FrontLd F = d.fronts[flist[fi]];
    const int32_t *rows = d.rows + F.rofs;
    const int8_t *tags = d.tags + F.rofs;
    const double *L = d.fac + F.lofs;
    const int64_t f = F.f;
    for (int c = threadIdx.x; c < F.pstore; c += blockDim.x) {
        const int8_t tg = tags[c];
        if (tg == -1) continue;
        const double d00 = L[c * f + c];
        for (int r = 0; r < nrhs; ++r) {
            double *yr = y + static_cast<int64_t>(r) * n;
            if (tg != 2) {
                yr[rows[c]] = (d00 == 0.0 || !isfinite(d00)) ? 0.0 : yr[rows[c]] / d00;
                continue;
            }
            const double d10 = L[(c + 1) * f + c], d11 = L[(c + 1) * f + c + 1];
            const double x0 = yr[rows[c]], x1 = yr[rows[c + 1]];
            double s = fabs(d00);
            if (s < fabs(d10)) s = fabs(d10);
            if (s < fabs(d11)) s = fabs(d11);
            double y0 = 0.0, y1 = 0.0;
            if (isfinite(s) && s != 0.0) {
                const double a = d00 / s, b = d10 / s, cc = d11 / s;
                const double det_s = a * cc - b * b;
                if (isfinite(det_s) && det_s != 0.0) {
                    if (!isfinite(x0) || !isfinite(x1)) {
                        y0 = x0;
                        y1 = x1;
                    } else {
                        const double rs = fmax(fabs(x0), fabs(x1));
                        if (rs != 0.0) {
                            const double r0 = x0 / rs, r1 = x1 / rs;
                            const double u0 = (cc * r0 - b * r1) / det_s;
                            const double u1 = (-b * r0 + a * r1) / det_s;
                            const double ratio = rs / s;
                            y0 = u0 * ratio;
                            y1 = u1 * ratio;
                        }
                    }
                }
            }
            yr[rows[c]] = y0;
            yr[rows[c + 1]] = y1;
        }
    }
}

void launch_ldlt_diag_div(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double *y, int32_t n,
                          int32_t nrhs) {
    if (nf <= 0) return;
    for (int64_t o = 0; o < nf; o += kMaxLaunchGroups) {
        const int32_t m = static_cast<int32_t>(std::min<int64_t>(kMaxLaunchGroups, nf - o));
        k_ldlt_diag_div<<<dim3(static_cast<unsigned>(m)), 64, 0, st>>>(d, flist + o, m, y, n, nrhs);
    }
}

void launch_ldlt_diag(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double *y, int32_t n,
                      int32_t nrhs) {
    if (nf <= 0) return;
    for (int64_t o = 0; o < nf; o += kMaxLaunchGroups) {
        const int32_t m = static_cast<int32_t>(std::min<int64_t>(kMaxLaunchGroups, nf - o));
        k_ldlt_diag<<<dim3(static_cast<unsigned>(m)), 64, 0, st>>>(d, flist + o, m, y, n, nrhs);
    }
}

}  // namespace smfgpu
