// The adapter smf would add (see gpu_numeric.hpp).
#include "gpu_numeric.hpp"

#include <numeric>

namespace smf {

int make_gpu_plan(const AnalysisKeep &k, bool posdef, GpuPlan &g) {
    const std::size_t ns = k.supernodes.size();
    g.sn_start.resize(ns);
    g.sn_end.resize(ns);
    g.sn_parent.resize(ns);
    for (std::size_t s = 0; s < ns; ++s) {
        g.sn_start[s] = k.supernodes[s].col_start;
        g.sn_end[s] = k.supernodes[s].col_end;
        g.sn_parent[s] = k.supernodes[s].parent;
    }
    g.perm = k.perm;  // perm[new] = old; empty: identity
    if (g.perm.empty()) {
        g.perm.resize(static_cast<std::size_t>(k.n));
        std::iota(g.perm.begin(), g.perm.end(), Int{0});
    }
    const bool ident = k.perm_col_ptr.empty();
    const std::vector<Int> &cp = ident ? k.cleaned.col_ptr : k.perm_col_ptr;
    const std::vector<Int> &ri = ident ? k.cleaned.row_idx : k.perm_row_idx;
    // each input entry's slot in the permuted lower pattern: the direct map
    // when the analysis kept one, else input -> cleaned -> permuted
    const std::size_t no = k.orig_row_idx.size();
    if (!k.orig_to_perm_direct.empty()) {
        g.slot = k.orig_to_perm_direct;
    } else {
        g.slot.resize(no);
        for (std::size_t e = 0; e < no; ++e) {
            const Int c = k.orig_to_clean_idx.empty() ? static_cast<Int>(e) : k.orig_to_clean_idx[e];
            g.slot[e] = c < 0 ? Int{-1} : (k.orig_to_perm_idx.empty() ? c : k.orig_to_perm_idx[static_cast<std::size_t>(c)]);
        }
    }
    smfgpu_analysis_desc d{};
    d.n = k.n;
    d.nsup = static_cast<int32_t>(ns);
    d.sn_start = g.sn_start.data();
    d.sn_end = g.sn_end.data();
    d.sn_parent = g.sn_parent.data();
    d.child_ptr = k.child_ptr.data();
    d.child_idx = k.child_idx.data();
    d.front_row_ptr = k.front_row_ptr.data();
    d.front_row_idx = k.front_row_idx.data();
    d.pcol_ptr = cp.data();
    d.prow_idx = ri.data();
    d.n_orig = static_cast<int64_t>(no);
    d.orig_to_slot = g.slot.data();
    d.perm = g.perm.data();
    return smfgpu_plan_create(&d, posdef ? 1 : 0, &g.plan);
}

FactorStatus factor_gpu(GpuPlan &g, std::span<const double> values, const Control &ctrl, double work_limit,
                        Info &info, smfgpu_factor_t &out) {
    smfgpu_factor_options o;
    smfgpu_default_options(&o);
    o.posdef = ctrl.matrix_type == MatrixType::RealSymmetricPositiveDefinite ? 1 : 0;
    o.pivot_u = ctrl.pivot_u;
    o.small_pivot = ctrl.small_pivot;
    o.max_negative_pivots = ctrl.max_negative_pivots;
    if (work_limit >= 0.0) {
        o.work_limit_on = 1;
        o.work_limit = work_limit;
    }
    // static pivoting (control.hpp:337-364): tolerances relative to max|A|
    o.static_pivot_tol = ctrl.static_pivot_tol;
    o.fallback_static_pivot_tol = ctrl.fallback_static_pivot_tol;
    o.static_flags = (ctrl.static_pivots_as_zero ? SMFGPU_STATIC_AS_ZERO : 0) |
                     (ctrl.static_zero_stays_zero ? SMFGPU_STATIC_ZERO_STAYS_ZERO : 0) |
                     (ctrl.static_pivot_in_order ? SMFGPU_STATIC_IN_ORDER : 0);
    smfgpu_factor_info fi{};
    const int rc =
        smfgpu_factor(g.plan, values.data(), static_cast<int64_t>(values.size()), &o, &out, &fi);
    if (rc == SMFGPU_ERR_VALUE) {
        info.status = ErrorCode::IllegalValue;
        return FactorStatus::Singular;
    }
    if (rc != SMFGPU_OK) {
        info.status = ErrorCode::AllocationError;
        return FactorStatus::Singular;
    }
    info.num_positive = fi.num_positive;
    info.num_negative = fi.num_negative;
    info.num_zero = fi.num_zero;
    info.delayed_pivots = fi.delayed_pivots;
    info.two_by_two_pivots = fi.two_by_two_pivots;
    info.actual_flops = fi.actual_flops;
    info.front_work = fi.front_work;
    info.max_front_size = fi.max_front;
    info.factor_seconds = fi.factor_seconds;
    info.static_pivots = fi.static_pivots;
    info.pivot_growth = fi.pivot_growth;
    info.factor_status = static_cast<FactorStatus>(fi.status);  // the same numbering
    return info.factor_status;
}

}  // namespace smf
