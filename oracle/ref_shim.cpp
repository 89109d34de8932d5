// Plain C entry points over the reference smf library (built from the
// reference sources by oracle/Makefile.ref). Test infrastructure only: the
// parity tests and bench.py's CPU-baseline leg call these through ctypes to
// obtain the reference's own analysis, factor, solve and dense-kernel
// results. Nothing in the product links or loads this file.
//
// Reference interfaces used (danlil240/opt_solver, smf 0.2.3):
//   Solver::analyse / factor(span) / solve   solver/include/smf/solver.hpp:60-114
//   AnalysisKeep fields                      solver/include/smf/analysis.hpp:25-144
//   FactorKeep layout                        solver/include/smf/factor_posdef.hpp:16-220
//   ldlt_factor_front                        solver/include/smf/ldlt_kernel.hpp:66-70
//   factor_indef (work_limit)                solver/include/smf/factor_indef.hpp:28-39
//   posdef_factor_front_lapack               solver/include/smf/ldlt_kernel.hpp:95
#include "smf/analysis.hpp"
#include "smf/control.hpp"
#include "smf/factor_indef.hpp"
#include "smf/factor_posdef.hpp"
#include "smf/info.hpp"
#include "smf/ldlt_kernel.hpp"
#include "smf/solver.hpp"
#include "smf/workspace.hpp"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

using namespace smf;

namespace {

struct RefAnalysis {
    std::unique_ptr<AnalysisKeep> ak;
    std::vector<Int> user_perm;
    int posdef = 0;
};

struct RefFactor {
    RefAnalysis *an = nullptr;
    FactorKeep fk;
    NumericWorkspace ws;
    Info info;
    Control ctrl;
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

Control base_control(int posdef, int num_threads) {
    Control c;
    c.matrix_type = posdef ? MatrixType::RealSymmetricPositiveDefinite
                           : MatrixType::RealSymmetricIndefinite;
    c.num_threads = num_threads > 0 ? num_threads : 1;
    return c;
}

}  // namespace

extern "C" {

// ordering: OrderingMethod value (2 AMD, 3 METIS, 4 UserSupplied with
// user_perm[new] = old). Returns an opaque handle or null.
void *ref_analyse(int n, const int *col_ptr, const int *row_idx, const double *values,
                  int ordering, const int *user_perm, int posdef, double *seconds) {
    auto *h = new RefAnalysis;
    h->posdef = posdef;
    CscLower A;
    A.n = n;
    A.col_ptr.assign(col_ptr, col_ptr + n + 1);
    const std::size_t nnz = static_cast<std::size_t>(col_ptr[n]);
    A.row_idx.assign(row_idx, row_idx + nnz);
    if (values != nullptr) A.values.assign(values, values + nnz);
    else A.values.assign(nnz, 1.0);
    Control c = base_control(posdef, 1);
    c.ordering = static_cast<OrderingMethod>(ordering);
    if (c.ordering == OrderingMethod::UserSupplied) {
        h->user_perm.assign(user_perm, user_perm + n);
        c.user_perm = h->user_perm.data();
    }
    Info info;
    Solver s;
    const double t0 = now_s();
    h->ak = s.analyse(A, c, info);
    if (seconds != nullptr) *seconds = now_s() - t0;
    if (!h->ak) {
        delete h;
        return nullptr;
    }
    return h;
}

// out[0..7] = n, nsup, child_idx size, front_row_idx size, permuted nnz,
// original nnz, identity ordering flag, slack plan flag.
void ref_analysis_sizes(void *handle, int64_t *out) {
    const AnalysisKeep &k = *static_cast<RefAnalysis *>(handle)->ak;
    out[0] = k.n;
    out[1] = static_cast<int64_t>(k.supernodes.size());
    out[2] = static_cast<int64_t>(k.child_idx.size());
    out[3] = static_cast<int64_t>(k.front_row_idx.size());
    const bool ident = k.perm_col_ptr.empty();
    out[4] = ident ? static_cast<int64_t>(k.cleaned.row_idx.size())
                   : static_cast<int64_t>(k.perm_row_idx.size());
    out[5] = static_cast<int64_t>(k.orig_row_idx.size());
    out[6] = ident ? 1 : 0;
    out[7] = k.slack ? 1 : 0;
}

// Any pointer may be null. orig_to_slot: for each entry of the analysed
// matrix, its slot in the permuted lower pattern (-1: dropped).
void ref_analysis_export(void *handle, int *sn_start, int *sn_end, int *sn_parent,
                         int *child_ptr, int *child_idx, int *front_row_ptr,
                         int *front_row_idx, int *postorder, int *perm, int *pcol_ptr,
                         int *prow_idx, int *orig_to_slot) {
    const AnalysisKeep &k = *static_cast<RefAnalysis *>(handle)->ak;
    const std::size_t ns = k.supernodes.size();
    for (std::size_t s = 0; s < ns; ++s) {
        if (sn_start) sn_start[s] = k.supernodes[s].col_start;
        if (sn_end) sn_end[s] = k.supernodes[s].col_end;
        if (sn_parent) sn_parent[s] = k.supernodes[s].parent;
        if (postorder) postorder[s] = k.postorder[s];
    }
    if (child_ptr) std::memcpy(child_ptr, k.child_ptr.data(), k.child_ptr.size() * sizeof(Int));
    if (child_idx) std::memcpy(child_idx, k.child_idx.data(), k.child_idx.size() * sizeof(Int));
    if (front_row_ptr)
        std::memcpy(front_row_ptr, k.front_row_ptr.data(), k.front_row_ptr.size() * sizeof(Int));
    if (front_row_idx)
        std::memcpy(front_row_idx, k.front_row_idx.data(), k.front_row_idx.size() * sizeof(Int));
    if (perm) {
        if (k.perm.empty())
            for (Int i = 0; i < k.n; ++i) perm[i] = i;
        else
            std::memcpy(perm, k.perm.data(), k.perm.size() * sizeof(Int));
    }
    const bool ident = k.perm_col_ptr.empty();
    const std::vector<Int> &cp = ident ? k.cleaned.col_ptr : k.perm_col_ptr;
    const std::vector<Int> &ri = ident ? k.cleaned.row_idx : k.perm_row_idx;
    if (pcol_ptr) std::memcpy(pcol_ptr, cp.data(), cp.size() * sizeof(Int));
    if (prow_idx) std::memcpy(prow_idx, ri.data(), ri.size() * sizeof(Int));
    if (orig_to_slot) {
        const std::size_t no = k.orig_row_idx.size();
        if (!k.orig_to_perm_direct.empty()) {
            std::memcpy(orig_to_slot, k.orig_to_perm_direct.data(), no * sizeof(Int));
        } else if (!k.orig_to_perm_idx.empty()) {
            std::memcpy(orig_to_slot, k.orig_to_perm_idx.data(), no * sizeof(Int));
        } else {
            for (std::size_t e = 0; e < no; ++e) orig_to_slot[e] = static_cast<Int>(e);
        }
    }
}

void ref_analysis_free(void *handle) { delete static_cast<RefAnalysis *>(handle); }

// Factor with values in the analysed matrix's entry order. Threshold
// pivoting with u; scale-on-delay off and no scaling, so the factor is the
// plain multifrontal one (the parity configuration). stats out:
// [0] status (FactorStatus) [1] ErrorCode [2] pos [3] neg [4] zero
// [5] delayed pivots [6] 2x2 pivots [7] actual flops [8] front work
// [9] seconds [10] static pivots [11] max front
// static pivoting (control.hpp:337-364): tolerances and flags bit 0
// static_pivots_as_zero, bit 1 static_zero_stays_zero, bit 2
// static_pivot_in_order; stats[12] = Info::pivot_growth.
void *ref_factor_ex(void *handle, const double *values, int64_t nvalues, double u, int num_threads,
                    double static_tol, double fallback_tol, int static_flags, double *stats) {
    auto *an = static_cast<RefAnalysis *>(handle);
    auto *h = new RefFactor;
    h->an = an;
    h->ctrl = base_control(an->posdef, num_threads);
    h->ctrl.pivot_u = u;
    h->ctrl.static_pivot_tol = static_tol;
    h->ctrl.fallback_static_pivot_tol = fallback_tol;
    h->ctrl.static_pivots_as_zero = (static_flags & 1) != 0;
    h->ctrl.static_zero_stays_zero = (static_flags & 2) != 0;
    h->ctrl.static_pivot_in_order = (static_flags & 4) != 0;
    h->ctrl.scaling = ScalingMethod::None;
    h->ctrl.scale_on_delay_ratio = -1.0;
    h->ctrl.matching_trigger = MatchingTrigger::Predicted;
    Solver s;
    const double t0 = now_s();
    const FactorStatus st = s.factor(*an->ak, std::span<const double>(values, static_cast<std::size_t>(nvalues)),
                                     h->ctrl, h->info, h->fk, h->ws);
    const double t1 = now_s();
    if (stats) {
        stats[0] = static_cast<double>(st);
        stats[1] = static_cast<double>(h->info.status);
        stats[2] = h->info.num_positive;
        stats[3] = h->info.num_negative;
        stats[4] = h->info.num_zero;
        stats[5] = h->info.delayed_pivots;
        stats[6] = h->info.two_by_two_pivots;
        stats[7] = h->info.actual_flops;
        stats[8] = h->info.front_work;
        stats[9] = t1 - t0;
        stats[10] = h->info.static_pivots;
        stats[11] = h->info.max_front_size;
        stats[12] = h->info.pivot_growth;
    }
    return h;
}

// The early stops on the values the analysis was built from:
// Control::max_negative_pivots (-1 off) and factor_indef's work_limit
// argument (factor_indef.hpp:28-39, -1 off), one thread, no scaling. stats
// as ref_factor_ex ([1] ErrorCode, [0] FactorStatus: 3 MaxPivotDelays,
// 4 InertiaExceeded; the partial counts of a stopped run).
void *ref_factor_limits(void *handle, double u, int max_negative_pivots, double work_limit, double *stats) {
    auto *an = static_cast<RefAnalysis *>(handle);
    auto *h = new RefFactor;
    h->an = an;
    h->ctrl = base_control(an->posdef, 1);
    h->ctrl.pivot_u = u;
    h->ctrl.max_negative_pivots = max_negative_pivots;
    h->ctrl.scaling = ScalingMethod::None;
    h->ctrl.scale_on_delay_ratio = -1.0;
    h->ctrl.matching_trigger = MatchingTrigger::Predicted;
    const double t0 = now_s();
    const FactorStatus st = factor_indef(*an->ak, h->ctrl, h->info, h->fk, work_limit);
    const double t1 = now_s();
    if (stats) {
        stats[0] = static_cast<double>(st);
        stats[1] = static_cast<double>(h->info.status);
        stats[2] = h->info.num_positive;
        stats[3] = h->info.num_negative;
        stats[4] = h->info.num_zero;
        stats[5] = h->info.delayed_pivots;
        stats[6] = h->info.two_by_two_pivots;
        stats[7] = h->info.actual_flops;
        stats[8] = h->info.front_work;
        stats[9] = t1 - t0;
        stats[10] = h->info.static_pivots;
        stats[11] = h->info.max_front_size;
        stats[12] = h->info.pivot_growth;
    }
    return h;
}

// ref_factor_ex's plain threshold factorization with a symmetric scaling:
// method 1 ScalingMethod::UserSupplied (user_scale, n doubles), 2
// ScalingMethod::Equilibration (scaling.cpp:57-101). scale_out (n doubles,
// may be null) receives the scale the factor holds (FactorKeep::scale);
// stats as ref_factor_ex.
void *ref_factor_scaled(void *handle, const double *values, int64_t nvalues, double u, int method,
                        const double *user_scale, double *scale_out, double *stats) {
    auto *an = static_cast<RefAnalysis *>(handle);
    auto *h = new RefFactor;
    h->an = an;
    h->ctrl = base_control(an->posdef, 1);
    h->ctrl.pivot_u = u;
    h->ctrl.scaling = method == 1 ? ScalingMethod::UserSupplied
                                  : (method == 2 ? ScalingMethod::Equilibration : ScalingMethod::None);
    h->ctrl.user_scale = user_scale;
    h->ctrl.scale_on_delay_ratio = -1.0;
    h->ctrl.matching_trigger = MatchingTrigger::Predicted;
    Solver s;
    const double t0 = now_s();
    const FactorStatus st = s.factor(*an->ak, std::span<const double>(values, static_cast<std::size_t>(nvalues)),
                                     h->ctrl, h->info, h->fk, h->ws);
    const double t1 = now_s();
    h->ctrl.user_scale = nullptr;  // copied into the factor; the caller's array may go
    if (scale_out)
        for (std::size_t i = 0; i < h->fk.scale.size(); ++i) scale_out[i] = h->fk.scale[i];
    if (stats) {
        stats[0] = static_cast<double>(st);
        stats[1] = static_cast<double>(h->info.status);
        stats[2] = h->info.num_positive;
        stats[3] = h->info.num_negative;
        stats[4] = h->info.num_zero;
        stats[5] = h->info.delayed_pivots;
        stats[6] = h->info.two_by_two_pivots;
        stats[7] = h->info.actual_flops;
        stats[8] = h->info.front_work;
        stats[9] = t1 - t0;
        stats[10] = h->info.static_pivots;
        stats[11] = h->info.max_front_size;
        stats[12] = h->info.pivot_growth;
        stats[13] = static_cast<double>(h->fk.scale.size());
    }
    return h;
}

void *ref_factor(void *handle, const double *values, int64_t nvalues, double u, int num_threads, double *stats) {
    return ref_factor_ex(handle, values, nvalues, u, num_threads, 0.0, 0.0, 0, stats);
}

// out: [0] steps [1] factor extent (doubles) [2] row index extent
// [3] diag entries
void ref_factor_sizes(void *handle, int64_t *out) {
    const FactorKeep &fk = static_cast<RefFactor *>(handle)->fk;
    out[0] = static_cast<int64_t>(fk.solve_steps.size());
    out[1] = static_cast<int64_t>(fk.factor_values.size());
    out[2] = static_cast<int64_t>(fk.solve_row_indices.size());
    out[3] = static_cast<int64_t>(fk.diag().size());
}

// steps: 4 int64 per step (row_offset, factor_offset, p, f); pivot_types:
// per step in step order, p entries each (SPD: empty). Any pointer may be
// null.
void ref_factor_export(void *handle, int64_t *steps, double *factor_values, int *row_indices,
                       int8_t *pivot_types, int *step_supernode) {
    const FactorKeep &fk = static_cast<RefFactor *>(handle)->fk;
    const std::size_t ns = fk.solve_steps.size();
    for (std::size_t t = 0; t < ns; ++t) {
        const SolveStep &st = fk.solve_steps[t];
        if (steps) {
            steps[4 * t + 0] = st.row_offset;
            steps[4 * t + 1] = st.factor_offset;
            steps[4 * t + 2] = st.p;
            steps[4 * t + 3] = st.f;
        }
    }
    if (factor_values) std::memcpy(factor_values, fk.factor_values.data(), fk.factor_values.size() * sizeof(double));
    if (row_indices) std::memcpy(row_indices, fk.solve_row_indices.data(), fk.solve_row_indices.size() * sizeof(Int));
    // Steps are in supernode postorder; the supernode of each step is found
    // from its factor offset.
    if (step_supernode) {
        const std::size_t nsup = fk.factor_col_ptr.size();
        std::size_t s = 0;
        for (std::size_t t = 0; t < ns; ++t) {
            const LongInt fo = fk.solve_steps[t].factor_offset;
            while (s < nsup && fk.factor_col_ptr[s] != fo) ++s;
            step_supernode[t] = s < nsup ? static_cast<int>(s) : -1;
            if (s < nsup) ++s;
        }
    }
    if (pivot_types) {
        std::size_t at = 0;
        for (std::size_t s = 0; s < fk.pivot_types.size(); ++s) {
            const auto &pt = fk.pivot_types[s];
            for (int8_t v : pt) pivot_types[at++] = v;
        }
    }
}

int64_t ref_pivot_types_size(void *handle) {
    const FactorKeep &fk = static_cast<RefFactor *>(handle)->fk;
    int64_t total = 0;
    for (const auto &pt : fk.pivot_types) total += static_cast<int64_t>(pt.size());
    return total;
}

// In-place solve of b (n x nrhs, column-major).
int ref_solve(void *handle, double *b, int n, int nrhs, double *seconds) {
    auto *h = static_cast<RefFactor *>(handle);
    Solver s;
    const double t0 = now_s();
    const int rc = s.solve(h->fk, h->ctrl, h->info, b, n, nrhs);
    if (seconds) *seconds = now_s() - t0;
    return rc;
}

// Solver::solve with a SolveJob (solver.hpp:101-114, types.hpp:250-256).
int ref_solve_job(void *handle, double *b, int n, int nrhs, int job) {
    auto *h = static_cast<RefFactor *>(handle);
    Solver s;
    return s.solve(h->fk, h->ctrl, h->info, b, n, nrhs, static_cast<SolveJob>(job));
}

void ref_factor_free(void *handle) { delete static_cast<RefFactor *>(handle); }

// The reference's dense TPP / in-order LDL^T front kernel on F (f x f,
// column-major, lower triangle). stats: [0] num_accepted [1] pos [2] neg
// [3] zero [4] flops [5] static pivots.
int ref_ldlt_front(double *F, int f, int num_elig, int *rows, double u, double eff_small,
                   double small_pivot, int nb, int posdef, double static_pivot, int8_t *tags,
                   double *stats) {
    LdltKernelParams prm;
    prm.u = u;
    prm.eff_small = eff_small;
    prm.small_pivot = small_pivot;
    prm.nb = nb;
    prm.posdef = posdef != 0;
    prm.static_pivot = static_pivot;
    const std::size_t wsize = ldlt_schur_work_size(f, num_elig, prm);
    std::vector<double> work(wsize + 1);
    const LdltKernelResult r = ldlt_factor_front(F, f, num_elig, rows, prm, tags, work.data(), 0, wsize);
    if (stats) {
        stats[0] = r.num_accepted;
        stats[1] = r.inertia.positive;
        stats[2] = r.inertia.negative;
        stats[3] = r.inertia.zero;
        stats[4] = r.flops;
        stats[5] = r.static_pivots;
    }
    return r.num_accepted;
}

// The reference's SPD front kernel (dpotrf + dtrsm + dsyrk).
int ref_posdef_front(double *F, int f, int num_elig, int8_t *tags, double *stats) {
    const LdltKernelResult r = posdef_factor_front_lapack(F, f, num_elig, tags);
    if (stats) {
        stats[0] = r.num_accepted;
        stats[4] = r.flops;
    }
    return r.num_accepted;
}

}  // extern "C"
