// smf's simple C API (smf_c.h) routed through the GPU adapter: the bodies
// smf_analyse / smf_factor / smf_solve / smf_inertia
// (solver/src/smf_c.cpp:94-287) would have with the GPU path on. The
// signatures and argument meanings are smf_c.h's; the handles are this
// file's own, as smf_c.cpp's structs are private to it.
#include "smf_gpu_c.h"

#include "gpu_numeric.hpp"
#include "smf/solver.hpp"

#include <memory>

struct smf_gpu_analysis_s {
    std::unique_ptr<smf::AnalysisKeep> keep;
    int n = 0;
    std::size_t nnz = 0;
    smf::Pivoting pivoting = smf::Pivoting::Threshold;
    smf::GpuPlan plan[2];  // per matrix type (0 indefinite, 1 positive definite), made on first use
};

struct smf_gpu_factor_s {
    smfgpu_factor_t f = nullptr;
    int pos = 0, neg = 0, zero = 0;
    ~smf_gpu_factor_s() { smfgpu_factor_free(&f); }
};

extern "C" {

int smf_gpu_analyse(const smf_csc_t *csc, smf_gpu_analysis_t *out) {
    if (!csc || !out || csc->n < 0 || !csc->col_ptr || !csc->row_idx || !csc->values) return SMF_ERR_INVALID_ARG;
    try {
        smf::CscLower A;
        A.n = static_cast<smf::Int>(csc->n);
        const int nnz = csc->col_ptr[csc->n];
        if (csc->col_ptr[0] != 0 || nnz < 0) return SMF_ERR_INVALID_ARG;
        A.col_ptr.assign(csc->col_ptr, csc->col_ptr + csc->n + 1);
        A.row_idx.assign(csc->row_idx, csc->row_idx + nnz);
        A.values.assign(csc->values, csc->values + nnz);
        smf::Control ctrl;
        ctrl.matrix_type = smf::MatrixType::RealSymmetricIndefinite;
        smf::Info info;
        smf::Solver solver;
        auto keep = solver.analyse(A, ctrl, info);
        if (!keep) return SMF_ERR_INTERNAL;
        auto *h = new smf_gpu_analysis_s;
        h->keep = std::move(keep);
        h->n = csc->n;
        h->nnz = static_cast<std::size_t>(nnz);
        *out = h;
        return SMF_OK;
    } catch (...) {
        return SMF_ERR_INTERNAL;
    }
}

int smf_gpu_factor(smf_gpu_analysis_t a, const double *values, int mtype, smf_gpu_factor_t *out) {
    if (!a || !values || !out) return SMF_ERR_INVALID_ARG;
    if (mtype != SMF_POSDEF && mtype != SMF_INDEF) return SMF_ERR_INVALID_ARG;
    try {
        smf::Control ctrl;
        const bool posdef = mtype == SMF_POSDEF;
        ctrl.matrix_type = posdef ? smf::MatrixType::RealSymmetricPositiveDefinite
                                  : smf::MatrixType::RealSymmetricIndefinite;
        if (!posdef) ctrl.pivoting = a->pivoting;
        smf::GpuPlan &g = a->plan[posdef ? 1 : 0];
        if (!g.plan && smf::make_gpu_plan(*a->keep, posdef, g) != SMFGPU_OK) return SMF_ERR_INTERNAL;
        smf::Info info;
        auto h = std::make_unique<smf_gpu_factor_s>();
        const smf::FactorStatus fs =
            smf::factor_gpu(g, std::span<const double>(values, a->nnz), ctrl, -1.0, info, h->f);
        if (fs != smf::FactorStatus::Success) {
            if (info.status == smf::ErrorCode::IllegalValue) return SMF_ERR_INVALID_ARG;
            if (fs == smf::FactorStatus::Singular || fs == smf::FactorStatus::NotPositiveDefinite)
                return SMF_ERR_SINGULAR;
            return SMF_ERR_INTERNAL;
        }
        h->pos = info.num_positive;
        h->neg = info.num_negative;
        h->zero = info.num_zero;
        *out = h.release();
        return SMF_OK;
    } catch (...) {
        return SMF_ERR_INTERNAL;
    }
}

int smf_gpu_solve(smf_gpu_analysis_t a, smf_gpu_factor_t f, int nrhs, double *rhs, int job) {
    if (!a || !f || !rhs || nrhs <= 0) return SMF_ERR_INVALID_ARG;
    int gjob;
    switch (job) {  // SolveJob numbering (Full, Forward, DiagOnly, Backward)
    case SMF_SOLVE_FULL: gjob = 0; break;
    case SMF_SOLVE_FORWARD: gjob = 1; break;
    case SMF_SOLVE_DIAGONAL: gjob = 2; break;
    case SMF_SOLVE_BACKWARD: gjob = 3; break;
    default: return SMF_ERR_INVALID_ARG;
    }
    const int rc = smfgpu_solve(f->f, rhs, nrhs, a->n, gjob, SMFGPU_MEM_HOST);
    return rc == SMFGPU_OK ? SMF_OK : (rc == SMFGPU_ERR_INVALID_ARG ? SMF_ERR_INVALID_ARG : SMF_ERR_INTERNAL);
}

int smf_gpu_inertia(smf_gpu_factor_t f, smf_inertia_t *inertia) {
    if (!f || !inertia) return SMF_ERR_INVALID_ARG;
    inertia->pos = f->pos;
    inertia->neg = f->neg;
    inertia->zero = f->zero;
    return SMF_OK;
}

void smf_gpu_free_analysis(smf_gpu_analysis_t *h) {
    if (!h || !*h) return;
    delete *h;
    *h = nullptr;
}

void smf_gpu_free_factor(smf_gpu_factor_t *h) {
    if (!h || !*h) return;
    delete *h;
    *h = nullptr;
}

}  // extern "C"
