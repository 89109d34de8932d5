// smfgpu_solve_noalloc: the device-memory solve with the right-hand-side
// work block in a fixed device array instead of a plan buffer allocated on
// first use, for callers that cannot allocate (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).
#include "smfgpu_plan.hpp"

#include <mutex>

namespace smfgpu {
namespace {

// Work block of the no-allocation solve: the permuted right-hand sides of
// one pass through the sweeps (n rows x up to 4 (Cholesky) or 16 (LDL^T)
// columns, as many as fit). 64 MiB per device, held until the process ends.
constexpr int64_t kScratchDoubles = int64_t{1} << 23;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};  // scratch's address on each device (null: not resolved yet)

}  // namespace

int scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return set_error(SMFGPU_ERR_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return SMFGPU_OK;
    int cur = -1;
    HIP_OK(hipGetDevice(&cur));
    if (cur != device) return set_error(SMFGPU_ERR_STATE, "scratch_init: device is not current");
    void *p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(SMFGPU_ERR_DEVICE, "scratch_init: no address for the work block");
    g_ptr[device] = static_cast<double *>(p);
    return SMFGPU_OK;
}

static double *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    return (device >= 0 && device < kMaxDevices) ? g_ptr[device] : nullptr;
}

}  // namespace smfgpu

extern "C" {

SMFGPU_API int smfgpu_solve_noalloc(smfgpu_factor_t f, double *b, int32_t nrhs, int32_t ldb, int32_t job) {
    if (!f || !b) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    if (!f->plan) return set_error(SMFGPU_ERR_STATE, "solve: the factor's plan was freed");
    smfgpu_plan_s *P = f->plan;
    if (P->live != f || !f->valid) return set_error(SMFGPU_ERR_STATE, "solve: factor not usable");
    if (job < 0 || job > 4) return set_error(SMFGPU_ERR_INVALID_ARG, "solve: job must be 0..4");
    if (nrhs <= 0 || ldb < P->n) return set_error(SMFGPU_ERR_INVALID_ARG, "solve: bad nrhs/ldb");
    if (static_cast<int64_t>(P->n) > kScratchDoubles)
        return set_error(SMFGPU_ERR_INVALID_ARG, "solve_noalloc: n = " + std::to_string(P->n) +
                                                     " exceeds the work block's " + std::to_string(kScratchDoubles) +
                                                     " rows");
    double *y = scratch_ptr(P->device);
    if (!y) return set_error(SMFGPU_ERR_STATE, "solve_noalloc: work block not resolved on the plan's device");
    int cur = -1;
    HIP_OK(hipGetDevice(&cur));
    if (cur != P->device) HIP_OK(hipSetDevice(P->device));
    int rc = solve_device(P, b, ldb, b, ldb, nrhs, job, y, kScratchDoubles);
    const hipError_t e = hipStreamSynchronize(P->stream);
    if (cur != P->device) (void)hipSetDevice(cur);
    if (rc == SMFGPU_OK && e != hipSuccess)
        rc = set_error(SMFGPU_ERR_DEVICE, std::string("solve_noalloc: ") + hipGetErrorString(e));
    if (rc == SMFGPU_OK) rc = sweep_check(P);
    return rc;
}

}  // extern "C"
