// Learned per-supernode precision selector: the C ABI (weights, training
// pairs, training, inference) around the kernels of selector.hip.
#include "smfgpu_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace smfgpu;

namespace {

// device buffers of one training / inference call, freed on return
struct Tmp {
    std::vector<void *> ptrs;
    ~Tmp() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **p, size_t n) {
        *p = nullptr;
        if (hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
            *p = nullptr;
            return set_error(SMFGPU_ERR_ALLOC, "selector: device allocation of " + std::to_string(n * sizeof(T)) +
                                                   " bytes failed");
        }
        ptrs.push_back(*p);
        return SMFGPU_OK;
    }
};

constexpr int kCbTile = 128;  // contribution-block tile of the label kernels

}  // namespace

extern "C" {

SMFGPU_API int smfgpu_selector_set(smfgpu_plan_t P, const float *w) {
    if (!P || !w) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < kSelWeights; ++i)
        if (!std::isfinite(w[i])) return set_error(SMFGPU_ERR_VALUE, "selector: non-finite weight");
    if (P->d_selw.p == nullptr) {
        const int rc = P->d_selw.alloc(kSelWeights);
        if (rc) return rc;
    }
    HIP_OK(hipMemcpyAsync(P->d_selw.p, w, kSelWeights * sizeof(float), hipMemcpyHostToDevice, P->stream));
    HIP_OK(hipStreamSynchronize(P->stream));
    P->sel_loaded = true;
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_factor_precisions(smfgpu_factor_t f, int8_t *out) {
    if (!f || !out) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    if (!f->plan) return set_error(SMFGPU_ERR_STATE, "precisions: the factor's plan was freed");
    smfgpu_plan_s *P = f->plan;
    if (P->live != f) return set_error(SMFGPU_ERR_STATE, "precisions: factor was superseded by a newer factorization");
    static const int8_t kAbi[4] = {SMFGPU_PREC_FP64, SMFGPU_PREC_FP32, SMFGPU_PREC_BF16, SMFGPU_PREC_BF16X3};
    for (int32_t s = 0; s < P->nsup; ++s) {
        const int8_t c = s < static_cast<int32_t>(f->prec.size()) ? f->prec[s] : 0;
        out[s] = (c >= 0 && c < 4) ? kAbi[c] : SMFGPU_PREC_FP64;
    }
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_selector_pairs(smfgpu_factor_t f, float *X, float *Y, int32_t *fronts, int32_t cap) {
    if (!f) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    if (!f->plan) return set_error(SMFGPU_ERR_STATE, "pairs: the factor's plan was freed");
    smfgpu_plan_s *P = f->plan;
    if (P->live != f || !f->valid) return set_error(SMFGPU_ERR_STATE, "pairs: factor not usable");
    if (!f->collected) return set_error(SMFGPU_ERR_STATE, "pairs: the factorization did not collect features");
    hipStream_t st = P->stream;
    auto emit = [&](const std::vector<int32_t> &cand, const std::vector<unsigned int> &hl,
                    const std::vector<float> &hf) {
        int32_t k = 0;
        for (int32_t s : cand) {
            if (k >= cap) break;
            float v[4];
            std::memcpy(v, hl.data() + 4 * static_cast<size_t>(s), sizeof(v));
            if (X)
                std::copy(hf.begin() + static_cast<size_t>(s) * kSelNF, hf.begin() + static_cast<size_t>(s + 1) * kSelNF,
                          X + static_cast<size_t>(k) * kSelNF);
            if (Y)
                for (int o = 0; o < kSelNO; ++o) {
                    const double rel = v[3] > 0.f ? static_cast<double>(v[o]) / v[3] : 1.0;
                    Y[static_cast<size_t>(k) * kSelNO + o] = static_cast<float>(std::log10(std::max(rel, 1e-30)));
                }
            if (fronts) fronts[k] = s;
            ++k;
        }
        return k;
    };
    if (!P->posdef) {
        // LDL^T: the labels were measured during the factorization
        const std::vector<int32_t> &cand = P->sel_cand;
        if (!X && !Y && !fronts) return static_cast<int32_t>(cand.size());
        std::vector<unsigned int> hl(static_cast<size_t>(P->nsup) * 4);
        std::vector<float> hf(static_cast<size_t>(P->nsup) * kSelNF);
        HIP_OK(hipMemcpyAsync(hl.data(), P->d_lab.p, hl.size() * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(hf.data(), P->d_feat.p, hf.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        return emit(cand, hl, hf);
    }
    // candidate fronts with a contribution block, and their CB tiles
    std::vector<int32_t> cand;
    for (int32_t s = 0; s < P->nsup; ++s)
        if (P->fronts[s].f >= f->mixed_min_front && P->fronts[s].q > 0) cand.push_back(s);
    const int32_t ncand = static_cast<int32_t>(cand.size());
    if (!X && !Y && !fronts) return ncand;
    Tmp tmp;
    unsigned int *lab = nullptr;
    int rc;
    if ((rc = tmp.alloc(&lab, static_cast<size_t>(P->nsup) * 4))) return rc;
    HIP_OK(hipMemsetAsync(lab, 0, static_cast<size_t>(P->nsup) * 4 * sizeof(unsigned int), st));
    // tiles in batches: the FP64 tile of each item is kept in scratch
    constexpr int64_t kBatch = 2048;
    double *zs = nullptr;
    SyrkItem *d_items = nullptr;
    if ((rc = tmp.alloc(&zs, static_cast<size_t>(kBatch) * kCbTile * kCbTile)) || (rc = tmp.alloc(&d_items, kBatch)))
        return rc;
    std::vector<SyrkItem> items;
    DevPtrs d{};
    d.fronts = P->d_fronts.p;
    d.L = P->d_L.p;
    d.CB = P->d_CB.p;
    auto flush = [&]() -> int {
        if (items.empty()) return SMFGPU_OK;
        HIP_OK(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(SyrkItem), hipMemcpyHostToDevice, st));
        launch_label_spd(st, d, d_items, static_cast<int64_t>(items.size()), zs, lab);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(st));  // items and zs are reused by the next batch
        items.clear();
        return SMFGPU_OK;
    };
    for (int32_t s : cand) {
        const FrontDev &F = P->fronts[s];
        for (int32_t j0 = F.p; j0 < F.f; j0 += kCbTile)
            for (int32_t i0 = j0; i0 < F.f; i0 += kCbTile) {
                items.push_back(SyrkItem{s, i0, j0, F.f, 0, F.p, 0, 0});
                if (static_cast<int64_t>(items.size()) == kBatch && (rc = flush())) return rc;
            }
    }
    if ((rc = flush())) return rc;
    std::vector<unsigned int> hl(static_cast<size_t>(P->nsup) * 4);
    std::vector<float> hf(static_cast<size_t>(P->nsup) * kSelNF);
    HIP_OK(hipMemcpyAsync(hl.data(), lab, hl.size() * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(hf.data(), P->d_feat.p, hf.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return emit(cand, hl, hf);
}

SMFGPU_API int smfgpu_selector_train(const float *X, const float *Y, const float *M, int32_t n, int32_t epochs,
                                     float lr, float *w, float *loss) {
    if (!X || !Y || !w || n <= 0 || epochs < 0 || !(lr > 0.f))
        return set_error(SMFGPU_ERR_INVALID_ARG, "selector_train: bad arguments");
    // input normalization from the training features
    float *mu = w + kSelParams, *isd = mu + kSelNF;
    for (int i = 0; i < kSelNF; ++i) {
        double a = 0.0, b = 0.0;
        for (int32_t r = 0; r < n; ++r) {
            const double x = X[static_cast<size_t>(r) * kSelNF + i];
            a += x;
            b += x * x;
        }
        const double m = a / n, var = std::max(b / n - m * m, 0.0);
        mu[i] = static_cast<float>(m);
        isd[i] = static_cast<float>(var > 1e-12 ? 1.0 / std::sqrt(var) : 1.0);
    }
    std::vector<float> ones;
    if (!M) {
        ones.assign(static_cast<size_t>(n) * kSelNO, 1.f);
        M = ones.data();
    }
    double cnt = 0.0;
    for (size_t i = 0; i < static_cast<size_t>(n) * kSelNO; ++i) cnt += M[i];
    if (cnt <= 0.0) return set_error(SMFGPU_ERR_INVALID_ARG, "selector_train: no labels");
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    Tmp tmp;
    float *dX, *dY, *dM, *dw, *dm, *dv, *dpart, *dloss;
    const int32_t maxblk = 64;
    int rc;
    if ((rc = tmp.alloc(&dX, static_cast<size_t>(n) * kSelNF)) || (rc = tmp.alloc(&dY, static_cast<size_t>(n) * kSelNO)) ||
        (rc = tmp.alloc(&dM, static_cast<size_t>(n) * kSelNO)) || (rc = tmp.alloc(&dw, kSelWeights)) ||
        (rc = tmp.alloc(&dm, kSelParams)) || (rc = tmp.alloc(&dv, kSelParams)) ||
        (rc = tmp.alloc(&dpart, static_cast<size_t>(maxblk) * (kSelParams + 1))) ||
        (rc = tmp.alloc(&dloss, std::max(epochs, 1))))
        return rc;
    HIP_OK(hipMemcpyAsync(dX, X, static_cast<size_t>(n) * kSelNF * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dY, Y, static_cast<size_t>(n) * kSelNO * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dM, M, static_cast<size_t>(n) * kSelNO * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dw, w, kSelWeights * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemsetAsync(dm, 0, kSelParams * sizeof(float), st));
    HIP_OK(hipMemsetAsync(dv, 0, kSelParams * sizeof(float), st));
    for (int32_t e = 0; e < epochs; ++e) {
        const int32_t nblk = launch_mlp_grad(st, dw, dX, dY, dM, n, dpart);
        launch_mlp_adam(st, dw, dm, dv, dpart, nblk, static_cast<float>(cnt), e + 1, lr, dloss + e);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(w, dw, kSelWeights * sizeof(float), hipMemcpyDeviceToHost, st));
    if (loss && epochs > 0)
        HIP_OK(hipMemcpyAsync(loss, dloss, static_cast<size_t>(epochs) * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_selector_predict(const float *w, const float *X, int32_t n, float *Y) {
    if (!w || !X || !Y || n < 0) return set_error(SMFGPU_ERR_INVALID_ARG, "selector_predict: bad arguments");
    if (n == 0) return SMFGPU_OK;
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    Tmp tmp;
    float *dw, *dX, *dY;
    int rc;
    if ((rc = tmp.alloc(&dw, kSelWeights)) || (rc = tmp.alloc(&dX, static_cast<size_t>(n) * kSelNF)) ||
        (rc = tmp.alloc(&dY, static_cast<size_t>(n) * kSelNO)))
        return rc;
    HIP_OK(hipMemcpyAsync(dw, w, kSelWeights * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(dX, X, static_cast<size_t>(n) * kSelNF * sizeof(float), hipMemcpyHostToDevice, st));
    launch_mlp_predict(st, dw, dX, n, dY);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(Y, dY, static_cast<size_t>(n) * kSelNO * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return SMFGPU_OK;
}

}  // extern "C"
