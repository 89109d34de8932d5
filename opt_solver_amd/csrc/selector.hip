// Learned per-supernode precision selector (gfx950): front features, MLP
// inference inside the factorization, the MLP's training kernels, and the
// labels the training reads, measured by running a front's Schur update
// both in FP64 and in each reduced precision on the MFMA tiles.
//
// Features of a front (computed from its assembled pivot columns before it
// is factored; log10 of ratios unless said otherwise):
//   0 log2 f (front order)           1 log2 p (pivot / eligible columns)
//   2 q / f (contribution-block share)
//   3 max / min pivot-column 2-norm  4 min over the columns of |a_jj| / ||a_j||
//   5 mean over the columns of log10(|a_jj| / ||a_j||)
//   6 max / min |a_jj|               7 max column norm / max |A|
//   8 delayed columns received / p   9 scaling state (1: scaled values)
// Outputs: predicted log10 relative error of the Schur update (max |dZ| /
// max |S| over the contribution block) in FP32, split BF16 and BF16.
#include "smfgpu_internal.hpp"
#include "mfma_tile.hpp"

#include <cfloat>

namespace smfgpu {

namespace {

constexpr int kThr = 256;

struct MlpView {
    const float *W1, *b1, *W2, *b2, *W3, *b3, *mu, *isd;
};
__device__ __forceinline__ MlpView mlp_view(const float *w) {
    MlpView v;
    v.W1 = w;
    v.b1 = v.W1 + kSelH * kSelNF;
    v.W2 = v.b1 + kSelH;
    v.b2 = v.W2 + kSelH * kSelH;
    v.W3 = v.b2 + kSelH;
    v.b3 = v.W3 + kSelNO * kSelH;
    v.mu = v.b3 + kSelNO;
    v.isd = v.mu + kSelNF;
    return v;
}

// forward pass of one sample: normalized input x, hidden layers h1, h2,
// outputs y
__device__ __forceinline__ void mlp_forward(const MlpView &m, const float *xin, float *x, float *h1, float *h2,
                                            float *y) {
#pragma unroll
    for (int i = 0; i < kSelNF; ++i) x[i] = (xin[i] - m.mu[i]) * m.isd[i];
    for (int j = 0; j < kSelH; ++j) {
        float a = m.b1[j];
#pragma unroll
        for (int i = 0; i < kSelNF; ++i) a += m.W1[j * kSelNF + i] * x[i];
        h1[j] = tanhf(a);
    }
    for (int j = 0; j < kSelH; ++j) {
        float a = m.b2[j];
        for (int i = 0; i < kSelH; ++i) a += m.W2[j * kSelH + i] * h1[i];
        h2[j] = tanhf(a);
    }
#pragma unroll
    for (int o = 0; o < kSelNO; ++o) {
        float a = m.b3[o];
        for (int i = 0; i < kSelH; ++i) a += m.W3[o * kSelH + i] * h2[i];
        y[o] = a;
    }
}

// cheapest allowed precision whose predicted error passes: BF16, then
// split BF16, then FP32; otherwise FP64
__device__ __forceinline__ int8_t choose(const float *y, float log_tol, int32_t mask) {
    if ((mask & 4) && y[kSelOutOfCode[2]] <= log_tol) return 2;
    if ((mask & 8) && y[kSelOutOfCode[3]] <= log_tol) return 3;
    if ((mask & 2) && y[kSelOutOfCode[1]] <= log_tol) return 1;
    return 0;
}

// One workgroup per front: reduce the pivot columns' statistics, form the
// features, run the MLP and pick the precision.
__global__ __launch_bounds__(kThr) void k_select(const SelFront *__restrict__ fr, const float4 *__restrict__ colstat,
                                                 SelPrm prm, float *__restrict__ feat, int8_t *__restrict__ prec) {
    const SelFront F = fr[blockIdx.x];
    const int tid = threadIdx.x;
    double nmax = 0.0, nmin = DBL_MAX, dmax = 0.0, dmin = DBL_MAX, rmin = DBL_MAX, lsum = 0.0;
    for (int j = tid; j < F.p; j += kThr) {
        const float4 c = colstat[F.cs0 + j];
        const double nrm = fmax(static_cast<double>(c.x), 1e-300);
        const double dg = fmax(static_cast<double>(c.y), 1e-300);
        nmax = fmax(nmax, nrm);
        nmin = fmin(nmin, nrm);
        dmax = fmax(dmax, dg);
        dmin = fmin(dmin, dg);
        rmin = fmin(rmin, dg / nrm);
        lsum += log10(dg / nrm);
    }
    __shared__ double red[6][kThr / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        nmax = fmax(nmax, __shfl_xor(nmax, o));
        nmin = fmin(nmin, __shfl_xor(nmin, o));
        dmax = fmax(dmax, __shfl_xor(dmax, o));
        dmin = fmin(dmin, __shfl_xor(dmin, o));
        rmin = fmin(rmin, __shfl_xor(rmin, o));
        lsum += __shfl_xor(lsum, o);
    }
    if ((tid & 63) == 0) {
        const int w = tid >> 6;
        red[0][w] = nmax;
        red[1][w] = nmin;
        red[2][w] = dmax;
        red[3][w] = dmin;
        red[4][w] = rmin;
        red[5][w] = lsum;
    }
    __syncthreads();
    if (tid != 0) return;
    for (int w = 1; w < kThr / 64; ++w) {
        red[0][0] = fmax(red[0][0], red[0][w]);
        red[1][0] = fmin(red[1][0], red[1][w]);
        red[2][0] = fmax(red[2][0], red[2][w]);
        red[3][0] = fmin(red[3][0], red[3][w]);
        red[4][0] = fmin(red[4][0], red[4][w]);
        red[5][0] += red[5][w];
    }
    double amax;
    {
        const unsigned long long b = prm.maxabs_bits ? *prm.maxabs_bits : 0ull;
        amax = fmax(__longlong_as_double(static_cast<long long>(b)), 1e-300);
    }
    const double p = fmax(static_cast<double>(F.p), 1.0);
    float xin[kSelNF];
    xin[0] = static_cast<float>(log2(fmax(static_cast<double>(F.f), 1.0)));
    xin[1] = static_cast<float>(log2(p));
    xin[2] = static_cast<float>(static_cast<double>(F.f - F.p) / fmax(static_cast<double>(F.f), 1.0));
    xin[3] = static_cast<float>(log10(red[0][0] / red[1][0]));
    xin[4] = static_cast<float>(log10(red[4][0]));
    xin[5] = static_cast<float>(red[5][0] / p);
    xin[6] = static_cast<float>(log10(red[2][0] / red[3][0]));
    xin[7] = static_cast<float>(log10(red[0][0] / amax));
    xin[8] = static_cast<float>(static_cast<double>(F.nd) / p);
    xin[9] = prm.scaled ? 1.f : 0.f;
    if (feat != nullptr) {
#pragma unroll
        for (int i = 0; i < kSelNF; ++i) feat[static_cast<int64_t>(F.s) * kSelNF + i] = xin[i];
    }
    int8_t code = 0;
    if (prm.w != nullptr) {
        float x[kSelNF], h1[kSelH], h2[kSelH], y[kSelNO];
        mlp_forward(mlp_view(prm.w), xin, x, h1, h2, y);
        code = choose(y, prm.log_tol, prm.mask);
    }
    if (prec != nullptr) prec[F.s] = code;
}

__global__ __launch_bounds__(kThr) void k_mlp_predict(const float *__restrict__ w, const float *__restrict__ X,
                                                      int32_t n, float *__restrict__ Y) {
    const int i = blockIdx.x * kThr + threadIdx.x;
    if (i >= n) return;
    float xin[kSelNF], x[kSelNF], h1[kSelH], h2[kSelH], y[kSelNO];
#pragma unroll
    for (int k = 0; k < kSelNF; ++k) xin[k] = X[static_cast<int64_t>(i) * kSelNF + k];
    mlp_forward(mlp_view(w), xin, x, h1, h2, y);
#pragma unroll
    for (int o = 0; o < kSelNO; ++o) Y[static_cast<int64_t>(i) * kSelNO + o] = y[o];
}

// Training gradient of the masked squared error sum_o M_o (y_o - t_o)^2.
// Samples go through in tiles of kTile: threads [0, kTile) each take one
// sample's forward and backward pass and leave its activations and
// back-propagated errors in LDS; then every thread sums, over the tile,
// the gradient of the parameters it owns (parameter p goes to thread
// p mod kThr), in registers across tiles.
constexpr int kTile = 128;
constexpr int kOwn = (kSelParams + kThr - 1) / kThr;

__global__ __launch_bounds__(kThr) void k_mlp_grad(const float *__restrict__ w, const float *__restrict__ X,
                                                   const float *__restrict__ Yt, const float *__restrict__ Mk,
                                                   int32_t n, float *__restrict__ part) {
    __shared__ float sx[kTile][kSelNF];
    __shared__ float sh1[kTile][kSelH];
    __shared__ float sh2[kTile][kSelH];
    __shared__ float sg1[kTile][kSelH];
    __shared__ float sg2[kTile][kSelH];
    __shared__ float sg3[kTile][kSelNO];
    __shared__ float sloss[kThr];
    const MlpView m = mlp_view(w);
    const int tid = threadIdx.x;
    float acc[kOwn];
#pragma unroll
    for (int k = 0; k < kOwn; ++k) acc[k] = 0.f;
    float loss = 0.f;
    for (int t0 = blockIdx.x * kTile; t0 < n; t0 += gridDim.x * kTile) {
        const int cnt = min(kTile, n - t0);
        if (tid < kTile) {
            float xin[kSelNF], x[kSelNF], h1[kSelH], h2[kSelH], y[kSelNO], g3[kSelNO], g2[kSelH];
            const bool live = tid < cnt;
            const int64_t i = t0 + (live ? tid : 0);
#pragma unroll
            for (int k = 0; k < kSelNF; ++k) xin[k] = X[i * kSelNF + k];
            mlp_forward(m, xin, x, h1, h2, y);
#pragma unroll
            for (int o = 0; o < kSelNO; ++o) {
                const float mk = live ? Mk[i * kSelNO + o] : 0.f;
                const float e = y[o] - Yt[i * kSelNO + o];
                g3[o] = 2.f * mk * e;
                loss += mk * e * e;
                sg3[tid][o] = g3[o];
            }
            for (int j = 0; j < kSelH; ++j) {
                float a = 0.f;
#pragma unroll
                for (int o = 0; o < kSelNO; ++o) a += m.W3[o * kSelH + j] * g3[o];
                g2[j] = a * (1.f - h2[j] * h2[j]);
                sg2[tid][j] = g2[j];
                sh2[tid][j] = h2[j];
                sh1[tid][j] = h1[j];
            }
            for (int j = 0; j < kSelH; ++j) {
                float a = 0.f;
                for (int k = 0; k < kSelH; ++k) a += m.W2[k * kSelH + j] * g2[k];
                sg1[tid][j] = a * (1.f - h1[j] * h1[j]);
            }
#pragma unroll
            for (int k = 0; k < kSelNF; ++k) sx[tid][k] = x[k];
        }
        __syncthreads();
        // parameter p: its gradient summed over the tile's samples
#pragma unroll
        for (int k = 0; k < kOwn; ++k) {
            const int p = tid + k * kThr;
            if (p >= kSelParams) break;
            float g = 0.f;
            int q = p;
            if (q < kSelH * kSelNF) {  // W1[j][i]
                const int j = q / kSelNF, ii = q % kSelNF;
                for (int s = 0; s < cnt; ++s) g += sg1[s][j] * sx[s][ii];
            } else if ((q -= kSelH * kSelNF) < kSelH) {  // b1
                for (int s = 0; s < cnt; ++s) g += sg1[s][q];
            } else if ((q -= kSelH) < kSelH * kSelH) {  // W2[j][i]
                const int j = q / kSelH, ii = q % kSelH;
                for (int s = 0; s < cnt; ++s) g += sg2[s][j] * sh1[s][ii];
            } else if ((q -= kSelH * kSelH) < kSelH) {  // b2
                for (int s = 0; s < cnt; ++s) g += sg2[s][q];
            } else if ((q -= kSelH) < kSelNO * kSelH) {  // W3[o][i]
                const int o = q / kSelH, ii = q % kSelH;
                for (int s = 0; s < cnt; ++s) g += sg3[s][o] * sh2[s][ii];
            } else {  // b3
                q -= kSelNO * kSelH;
                for (int s = 0; s < cnt; ++s) g += sg3[s][q];
            }
            acc[k] += g;
        }
        __syncthreads();
    }
    float *out = part + static_cast<int64_t>(blockIdx.x) * (kSelParams + 1);
#pragma unroll
    for (int k = 0; k < kOwn; ++k) {
        const int p = tid + k * kThr;
        if (p < kSelParams) out[p] = acc[k];
    }
    sloss[tid] = loss;
    __syncthreads();
    for (int s = kThr / 2; s > 0; s >>= 1) {
        if (tid < s) sloss[tid] += sloss[tid + s];
        __syncthreads();
    }
    if (tid == 0) out[kSelParams] = sloss[0];
}

// Adam (Kingma and Ba, 2015) with beta1 0.9, beta2 0.999, eps 1e-8 on the
// mean gradient over cnt labels.
__global__ __launch_bounds__(kThr) void k_mlp_adam(float *__restrict__ w, float *__restrict__ mm, float *__restrict__ vv,
                                                   const float *__restrict__ part, int32_t nblk, float cnt, int32_t step,
                                                   float lr, float *__restrict__ loss) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, static_cast<float>(step));
    const float c2 = 1.f - powf(b2, static_cast<float>(step));
    for (int p = threadIdx.x; p < kSelParams; p += kThr) {
        float g = 0.f;
        for (int b = 0; b < nblk; ++b) g += part[static_cast<int64_t>(b) * (kSelParams + 1) + p];
        g /= cnt;
        const float m1 = b1 * mm[p] + (1.f - b1) * g;
        const float v1 = b2 * vv[p] + (1.f - b2) * g * g;
        mm[p] = m1;
        vv[p] = v1;
        w[p] -= lr * (m1 / c1) / (sqrtf(v1 / c2) + eps);
    }
    if (threadIdx.x == 0) {
        float l = 0.f;
        for (int b = 0; b < nblk; ++b) l += part[static_cast<int64_t>(b) * (kSelParams + 1) + kSelParams];
        loss[0] = l / cnt;
    }
}

// ---------------------------------------------------------------------------
// Labels (SPD): a contribution-block tile Z = L21 L21^T over all p pivot
// columns of the factored front, from the FP64 factor: MODE 0 in FP64 MFMA
// (stored to zs, one 128 x 128 tile per item, and max |S| of the tile's
// updated contribution-block entries into lab[4 s + 3]); MODE 1, 2, 3 in
// FP32, BF16 and split BF16 (operands rounded from FP64 as they enter LDS,
// the production tiles' rounding), max |Z_m - Z_64| into lab[4 s + out].
// Values go into lab as float bits (non-negative floats order like their
// bits).
template <int MODE>
struct LabelLds {
    static constexpr int bytes = MODE == 0   ? tile_lds_bytes<double, 16>()
                                 : MODE == 1 ? tile_lds_bytes<float, 32>()
                                             : bf_lds_bytes<MODE == 3, 32>();
};

__device__ __forceinline__ void atomic_max_float(unsigned int *a, float v) {
    atomicMax(a, __float_as_uint(fmaxf(v, 0.f)));
}

template <int MODE>
__global__ __launch_bounds__(kThr) void k_label_spd(DevPtrs d, const SyrkItem *__restrict__ items, double *zs,
                                                    unsigned int *lab) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[LabelLds<MODE>::bytes];
    const SyrkItem it = items[blockIdx.x];
    const FrontDev F = d.fronts[it.s];
    const int f = F.f;
    const int i0 = it.i0, j0 = it.j0, j1 = min(it.j1, f);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
    const bool active = (i0 + wr * 64 + 63) >= (j0 + wc * 64);
    const double *Lb = d.L + F.lofs;
    double *z = zs + static_cast<int64_t>(blockIdx.x) * kT * kT;
    float worst = 0.f;
    if constexpr (MODE == 0) {
        typedef SyrkMfma<double> M;
        M::acc_t acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = M::zero();
        if (active) tile_loop<double, true, 16, 1>(Lb, Lb, f, 0, i0, j0, f, j1, it.t0, it.t1, smem, acc);
        else tile_loop<double, false, 16, 1>(Lb, Lb, f, 0, i0, j0, f, j1, it.t0, it.t1, smem, acc);
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int jl = wc * 64 + m * 16 + M::row(v, lane);
                const int j = j0 + jl;
#pragma unroll
                for (int nn = 0; nn < 4; ++nn) {
                    const int il = wr * 64 + nn * 16 + (lane & 15);
                    const int i = i0 + il;
                    z[jl * kT + il] = acc[m][nn][v];
                    if (active && i < f && j < j1 && i >= j) {
                        const double sv = d.CB[F.cbofs + static_cast<int64_t>(j - F.p) * F.q + (i - F.p)];
                        worst = fmaxf(worst, static_cast<float>(fabs(sv)));
                    }
                }
            }
    } else {
        typedef SyrkMfma<float> M;
        M::acc_t acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = M::zero();
        if constexpr (MODE == 1) {
            if (active) tile_loop<float, true, 32, 1, double>(Lb, Lb, f, 0, i0, j0, f, j1, it.t0, it.t1, smem, acc);
            else tile_loop<float, false, 32, 1, double>(Lb, Lb, f, 0, i0, j0, f, j1, it.t0, it.t1, smem, acc);
        } else {
            constexpr bool kSplit = MODE == 3;
            if (active) tile_loop_bf16<kSplit, true, 32, double>(Lb, Lb, f, 0, i0, j0, f, j1, it.t0, it.t1, smem, acc);
            else tile_loop_bf16<kSplit, false, 32, double>(Lb, Lb, f, 0, i0, j0, f, j1, it.t0, it.t1, smem, acc);
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int jl = wc * 64 + m * 16 + M::row(v, lane);
                const int j = j0 + jl;
#pragma unroll
                for (int nn = 0; nn < 4; ++nn) {
                    const int il = wr * 64 + nn * 16 + (lane & 15);
                    const int i = i0 + il;
                    if (active && i < f && j < j1 && i >= j)
                        worst = fmaxf(worst, static_cast<float>(fabs(static_cast<double>(acc[m][nn][v]) - z[jl * kT + il])));
                }
            }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) worst = fmaxf(worst, __shfl_xor(worst, o));
    const int slot = MODE == 0 ? 3 : kSelOutOfCode[MODE];
    if (lane == 0) atomic_max_float(lab + 4 * static_cast<int64_t>(it.s) + slot, worst);
}

// LDL^T square statistics: one workgroup per (front, eligible column)
__global__ __launch_bounds__(kThr) void k_sq_colstat(const double *__restrict__ arena, const SelFront *__restrict__ fr,
                                                     float4 *__restrict__ colstat) {
    const SelFront F = fr[blockIdx.y];
    const int j = blockIdx.x;
    if (j >= F.p) return;
    const double *col = arena + F.fofs + static_cast<int64_t>(j) * F.f;
    double ssq = 0.0, dg = 0.0, off = 0.0;
    for (int r = j + threadIdx.x; r < F.f; r += kThr) {
        const double v = col[r];
        ssq += v * v;
        if (r == j) dg = fabs(v);
        else off = fmax(off, fabs(v));
    }
    __shared__ double red[3][kThr / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ssq += __shfl_xor(ssq, o);
        dg = fmax(dg, __shfl_xor(dg, o));
        off = fmax(off, __shfl_xor(off, o));
    }
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = ssq;
        red[1][tid >> 6] = dg;
        red[2][tid >> 6] = off;
    }
    __syncthreads();
    if (tid == 0) {
        double a = 0.0, b = 0.0, m = 0.0;
        for (int w = 0; w < kThr / 64; ++w) {
            a += red[0][w];
            b = fmax(b, red[1][w]);
            m = fmax(m, red[2][w]);
        }
        colstat[F.cs0 + j] = make_float4(static_cast<float>(sqrt(a)), static_cast<float>(b), static_cast<float>(m), 0.f);
    }
}

// LDL^T labels: one workgroup per (front, contribution-block column)
__global__ __launch_bounds__(kThr) void k_label_sq(const double *__restrict__ arena, const double *__restrict__ snap,
                                                   int64_t stride, const SelFront *__restrict__ fr,
                                                   unsigned int *lab) {
    const SelFront F = fr[blockIdx.y];
    const int j = F.p + blockIdx.x;
    if (j >= F.f) return;
    const int64_t co = static_cast<int64_t>(j) * F.f;
    const double *sq = arena + F.fofs + co;
    float d0 = 0.f, d1 = 0.f, d2 = 0.f, sm = 0.f;
    for (int i = j + threadIdx.x; i < F.f; i += kThr) {
        const double v = sq[i];
        sm = fmaxf(sm, static_cast<float>(fabs(v)));
        d0 = fmaxf(d0, static_cast<float>(fabs(snap[F.sofs + co + i] - v)));
        d1 = fmaxf(d1, static_cast<float>(fabs(snap[stride + F.sofs + co + i] - v)));
        d2 = fmaxf(d2, static_cast<float>(fabs(snap[2 * stride + F.sofs + co + i] - v)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sm = fmaxf(sm, __shfl_xor(sm, o));
        d0 = fmaxf(d0, __shfl_xor(d0, o));
        d1 = fmaxf(d1, __shfl_xor(d1, o));
        d2 = fmaxf(d2, __shfl_xor(d2, o));
    }
    if ((threadIdx.x & 63) == 0) {
        unsigned int *l = lab + 4 * static_cast<int64_t>(F.s);
        atomic_max_float(l + 0, d0);
        atomic_max_float(l + 1, d1);
        atomic_max_float(l + 2, d2);
        atomic_max_float(l + 3, sm);
    }
}

}  // namespace

void launch_sq_colstat(hipStream_t st, const double *arena, const SelFront *fr, int32_t nf, int32_t maxcols,
                       float4 *colstat) {
    if (nf <= 0 || maxcols <= 0) return;
    k_sq_colstat<<<dim3(static_cast<unsigned>(maxcols), static_cast<unsigned>(nf)), kThr, 0, st>>>(arena, fr, colstat);
}

void launch_label_sq(hipStream_t st, const double *arena, const double *snap, int64_t stride, const SelFront *fr,
                     int32_t nf, int32_t maxq, unsigned int *lab) {
    if (nf <= 0 || maxq <= 0) return;
    k_label_sq<<<dim3(static_cast<unsigned>(maxq), static_cast<unsigned>(nf)), kThr, 0, st>>>(arena, snap, stride, fr,
                                                                                            lab);
}

void launch_select(hipStream_t st, const SelFront *fr, int32_t nf, const float4 *colstat, const SelPrm &prm,
                   float *feat, int8_t *prec) {
    if (nf <= 0) return;
    k_select<<<dim3(static_cast<unsigned>(nf)), kThr, 0, st>>>(fr, colstat, prm, feat, prec);
}

void launch_mlp_predict(hipStream_t st, const float *w, const float *X, int32_t n, float *Y) {
    if (n <= 0) return;
    k_mlp_predict<<<dim3(static_cast<unsigned>((n + kThr - 1) / kThr)), kThr, 0, st>>>(w, X, n, Y);
}

int32_t launch_mlp_grad(hipStream_t st, const float *w, const float *X, const float *Y, const float *M, int32_t n,
                        float *part) {
    const int32_t nblk = std::max(1, std::min<int32_t>(64, (n + kTile - 1) / kTile));
    k_mlp_grad<<<dim3(static_cast<unsigned>(nblk)), kThr, 0, st>>>(w, X, Y, M, n, part);
    return nblk;
}

void launch_mlp_adam(hipStream_t st, float *w, float *m, float *v, const float *part, int32_t nblk, float cnt,
                     int32_t step, float lr, float *loss) {
    k_mlp_adam<<<1, kThr, 0, st>>>(w, m, v, part, nblk, cnt, step, lr, loss);
}

void launch_label_spd(hipStream_t st, const DevPtrs &d, const SyrkItem *items, int64_t n, double *zs,
                      unsigned int *lab) {
    if (n <= 0) return;
    const dim3 g(static_cast<unsigned>(n));
    k_label_spd<0><<<g, kThr, 0, st>>>(d, items, zs, lab);
    k_label_spd<1><<<g, kThr, 0, st>>>(d, items, zs, lab);
    k_label_spd<2><<<g, kThr, 0, st>>>(d, items, zs, lab);
    k_label_spd<3><<<g, kThr, 0, st>>>(d, items, zs, lab);
}

}  // namespace smfgpu
