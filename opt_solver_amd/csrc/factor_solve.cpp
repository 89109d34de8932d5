// smfgpu_factor_solve: factor + solve + refinement with the fallback of
// reduced-precision factors to plain FP64 when refinement stalls.
#include "smfgpu_plan.hpp"

#include <cstring>

using namespace smfgpu;

namespace {

// x = b, both n x nrhs with leading dimension ld
int copy_rhs(hipStream_t st, const double *b, double *x, int32_t n, int32_t nrhs, int32_t ld, int32_t mem) {
    if (b == x) return SMFGPU_OK;
    if (mem == SMFGPU_MEM_DEVICE) {
        HIP_OK(hipMemcpy2DAsync(x, ld * sizeof(double), b, ld * sizeof(double), n * sizeof(double), nrhs,
                                hipMemcpyDeviceToDevice, st));
        HIP_OK(hipStreamSynchronize(st));
    } else {
        for (int32_t r = 0; r < nrhs; ++r)
            std::memcpy(x + static_cast<int64_t>(r) * ld, b + static_cast<int64_t>(r) * ld, n * sizeof(double));
    }
    return SMFGPU_OK;
}

}  // namespace

extern "C" {

SMFGPU_API int smfgpu_factor_solve(smfgpu_plan_t P, const double *values, int64_t nvalues,
                                   const smfgpu_factor_options *opt_in, const double *b, double *x, int32_t nrhs,
                                   int32_t ld, const smfgpu_refine_options *ropt_in, int32_t mem,
                                   smfgpu_factor_t *out, smfgpu_factor_info *finfo, smfgpu_refine_info *rinfo) {
    if (!P || !b || !x || !out) return set_error(SMFGPU_ERR_INVALID_ARG, "factor_solve: null argument");
    *out = nullptr;
    if (nrhs <= 0 || ld < P->n) return set_error(SMFGPU_ERR_INVALID_ARG, "factor_solve: bad nrhs/ld");
    smfgpu_factor_options opt;
    if (opt_in) opt = *opt_in;
    else smfgpu_default_options(&opt);
    smfgpu_refine_options ropt;
    if (ropt_in) ropt = *ropt_in;
    else smfgpu_refine_default_options(&ropt);
    smfgpu_refine_info ri{};
    smfgpu_factor_info fi{};
    auto attempt = [&](const smfgpu_factor_options &o, smfgpu_factor_t *f) -> int {
        int rc = smfgpu_factor(P, values, nvalues, &o, f, &fi);
        if (rc != SMFGPU_OK) return rc;
        if (fi.status != SMFGPU_STATUS_SUCCESS) return SMFGPU_OK;
        if ((rc = copy_rhs(P->stream, b, x, P->n, nrhs, ld, mem))) return rc;
        if ((rc = smfgpu_solve(*f, x, nrhs, ld, 0, mem))) return rc;
        return smfgpu_refine_ex(*f, nullptr, b, ld, x, ld, nrhs, &ropt, mem, &ri);
    };
    smfgpu_factor_t f = nullptr;
    int rc = attempt(opt, &f);
    const bool reduced_run = opt.precision != SMFGPU_PREC_FP64;
    if (rc == SMFGPU_OK && fi.status == SMFGPU_STATUS_SUCCESS && reduced_run && !ri.reached) {
        // the reduced-precision factor did not get there: plain FP64
        smfgpu_factor_free(&f);
        smfgpu_factor_options o64 = opt;
        o64.precision = SMFGPU_PREC_FP64;
        o64.collect = 0;
        ri = smfgpu_refine_info{};
        rc = attempt(o64, &f);
        ri.fallback = 1;
        ++P->sel_fallbacks;
    }
    if (finfo) *finfo = fi;
    if (rinfo) *rinfo = ri;
    *out = f;
    return rc;
}

}  // extern "C"
