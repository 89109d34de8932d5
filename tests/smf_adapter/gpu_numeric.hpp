// The adapter smf would add to route its numeric phase through smfgpu
// (the code of INTEGRATION.md, compiled here against the reference's own
// headers by oracle/Makefile.ref's `adapter` target). Test infrastructure:
// tests/test_gpu_smf_adapter.py runs the resulting binary.
#pragma once

#include "smfgpu.h"
#include "smf/analysis.hpp"
#include "smf/control.hpp"
#include "smf/info.hpp"

#include <span>
#include <vector>

namespace smf {

// A device plan of one AnalysisKeep, created once and reused for every
// refactorization with the same pattern (what an IPM session does).
struct GpuPlan {
    smfgpu_plan_t plan = nullptr;
    std::vector<Int> sn_start, sn_end, sn_parent, perm, slot;
    GpuPlan() = default;
    GpuPlan(const GpuPlan &) = delete;
    GpuPlan &operator=(const GpuPlan &) = delete;
    ~GpuPlan() { smfgpu_plan_free(&plan); }
};

// smfgpu_analysis_desc from the AnalysisKeep fields the numeric phase reads
// (analysis.hpp:25-144), then smfgpu_plan_create.
int make_gpu_plan(const AnalysisKeep &k, bool posdef, GpuPlan &g);

// Solver::factor's numeric work on the GPU (solver.hpp:88): Control's
// pivoting parameters to smfgpu_factor_options, smfgpu_factor_info back into
// Info. work_limit is factor_indef's argument (< 0: none).
FactorStatus factor_gpu(GpuPlan &g, std::span<const double> values, const Control &ctrl, double work_limit,
                        Info &info, smfgpu_factor_t &out);

}  // namespace smf
