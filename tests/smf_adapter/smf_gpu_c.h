/* smf's simple C API (solver/include/smf/smf_c.h:165-178) on the GPU
 * adapter: the same arguments and return codes, smf_gpu_ handles. */
#pragma once

#include "smf/smf_c.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct smf_gpu_analysis_s *smf_gpu_analysis_t;
typedef struct smf_gpu_factor_s *smf_gpu_factor_t;

int smf_gpu_analyse(const smf_csc_t *csc, smf_gpu_analysis_t *out);
int smf_gpu_factor(smf_gpu_analysis_t analysis, const double *values, int mtype, smf_gpu_factor_t *out);
int smf_gpu_solve(smf_gpu_analysis_t analysis, smf_gpu_factor_t factor, int nrhs, double *rhs, int job);
int smf_gpu_inertia(smf_gpu_factor_t factor, smf_inertia_t *inertia);
void smf_gpu_free_analysis(smf_gpu_analysis_t *handle);
void smf_gpu_free_factor(smf_gpu_factor_t *handle);

#ifdef __cplusplus
}
#endif
