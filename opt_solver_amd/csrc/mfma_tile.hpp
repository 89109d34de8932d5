// MFMA tile of a symmetric rank-K update, shared by the SPD Schur update and
// the indefinite large-front updates (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace smfgpu {

typedef double v4d_t __attribute__((ext_vector_type(4)));
typedef float v4f_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// A10/A11: Schur-complement update on MFMA. Tile 128 x 128 of
// C[i, j] -= sum_{t in [t0, t1)} L[i, t] L[j, t], i >= j, four waves of
// 64 x 64, each 4 x 4 MFMA 16x16x4 blocks. The MFMA computes the
// transposed block (A operand = the j rows, B operand = the i rows) so that
// the 16 lanes of a row group hold 16 consecutive rows of one column on the
// store. prec 0: FP64 MFMA; 1: operands rounded to FP32, FP32 MFMA with FP32
// accumulation over the panel, subtracted from the FP64 Schur complement.
constexpr int kT = 128;
constexpr int kTileThreads = 256;

template <typename E>
struct SyrkMfma;
template <>
struct SyrkMfma<double> {
    typedef v4d_t acc_t;
    static __device__ __forceinline__ acc_t zero() { return v4d_t{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row (within the 16 x 16 block) of accumulator value v in this lane
    static __device__ __forceinline__ int row(int v, int lane) { return 4 * v + (lane >> 4); }
};
template <>
struct SyrkMfma<float> {
    typedef v4f_t acc_t;
    static __device__ __forceinline__ acc_t zero() { return v4f_t{0.f, 0.f, 0.f, 0.f}; }
    static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int v, int lane) { return 4 * (lane >> 4) + v; }
};

// Word 3 of a raw buffer resource on gfx9-family targets (gfx950): the
// data format field that makes a raw buffer read dwords.
constexpr int kBufferWord3 = 0x00020000;

template <typename GE>
__device__ __forceinline__ GE buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ float buffer_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, static_cast<int>(off), 0, 0));
}
template <>
__device__ __forceinline__ double buffer_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, static_cast<int>(off), 0, 0));
}

// LDS row length of a staged chunk: FP32 rows are padded by 16 floats so
// that the four panel columns a wave reads at once (lanes 16 k .. 16 k + 15
// take column k) fall on different banks; FP64 reads do not conflict.
template <typename E>
__host__ __device__ constexpr int tile_lds_ld() { return kT + (sizeof(E) == 4 ? 16 : 0); }
// LDS bytes of the two stages of both operands
template <typename E, int kKc>
__host__ __device__ constexpr int tile_lds_bytes() { return 4 * kKc * tile_lds_ld<E>() * static_cast<int>(sizeof(E)); }

// One 128 x 128 tile, operands of type E staged in a double-buffered LDS
// ring. kAhead = 1: the global loads of chunk k+1 are in flight during the
// MFMAs of chunk k; kAhead = 2: those of chunk k+2 (two register sets, the
// loads get two MFMA phases to arrive). Loads are branch-free buffer loads
// (out-of-range rows and columns read 0); waves
// wholly above the diagonal run the same loop without MFMAs (ACTIVE =
// false).
// A operand element (row i, panel column t) = Ab[t * ld + i], B operand
// element (row j, t) = Bb[(t - boff) * ld + j]; rows i < f, j < j1 are
// valid, panel columns [t0, t1). GE is the operands' type in global memory:
// GE = double with E = float rounds FP64 operands to FP32 as they are
// written to LDS (FP32 MFMA on FP64 data).
template <typename E, bool ACTIVE, int kKc, int kAhead, typename GE = E>
__device__ __forceinline__ void tile_loop(const GE *Ab, const GE *Bb, int ld, int boff, int i0, int j0, int f, int j1,
                                          int t0, int t1, unsigned char *smem,
                                          typename SyrkMfma<E>::acc_t (&acc)[4][4]) {
    typedef SyrkMfma<E> M;
    constexpr int kThreads = kTileThreads;
    constexpr int kLoadPer = (kKc * kT) / kThreads;  // elements of A (and of B) per thread per chunk
    constexpr int kLd = tile_lds_ld<E>();
    E (*As)[kKc][kLd] = reinterpret_cast<E (*)[kKc][kLd]>(smem);                             // [2][kKc][kLd]
    E (*Bs)[kKc][kLd] = reinterpret_cast<E (*)[kKc][kLd]>(smem + 2 * kKc * kLd * sizeof(E));  // [2][kKc][kLd]
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    typedef GE Regs[kLoadPer];
    // Buffer loads: per chunk one descriptor over its panel columns
    // [kc, min(kc + kKc, t1)) and per operand one lane offset. Columns past
    // t1, and rows past the front (A) or past j1 (B), whose offsets are
    // pushed past the descriptor's end, read 0 without a branch, a clamp or
    // a mask (the range check covers the lane offset).
    constexpr int kPass = kThreads / kT;  // panel columns of one load instruction
    constexpr uint32_t kOut = 0x80000000u;
    const int rl = tid % kT, th = tid / kT;
    const uint32_t colb = static_cast<uint32_t>(ld) * static_cast<uint32_t>(sizeof(GE));
    const uint32_t voa = (i0 + rl < f) ? th * colb + static_cast<uint32_t>(i0 + rl) * sizeof(GE) : kOut;
    const uint32_t vob = (j0 + rl < j1) ? th * colb + static_cast<uint32_t>(j0 + rl) * sizeof(GE) : kOut;
    auto gload = [&](Regs &ra, Regs &rb, int kc) {
        const int nrec = min(kKc, t1 - kc) * static_cast<int>(colb);
        const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<GE *>(Ab + static_cast<int64_t>(kc) * ld), 0, nrec, kBufferWord3);
        const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<GE *>(Bb + static_cast<int64_t>(kc - boff) * ld), 0, nrec, kBufferWord3);
#pragma unroll
        for (int e = 0; e < kLoadPer; ++e) {
            const uint32_t co = static_cast<uint32_t>(e * kPass) * colb;
            ra[e] = buffer_load<GE>(sa, voa + co);
            rb[e] = buffer_load<GE>(sb, vob + co);
        }
    };
    auto sstore = [&](const Regs &ra, const Regs &rb, int stg, int kc) {
        (void)kc;
#pragma unroll
        for (int e = 0; e < kLoadPer; ++e) {
            const int t = kPass * e + th;
            As[stg][t][rl] = static_cast<E>(ra[e]);
            Bs[stg][t][rl] = static_cast<E>(rb[e]);
        }
    };
    auto mfma_stage = [&](int stg) {
        if (!ACTIVE) return;
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 4) {
            E a[4], b[4];
            const int kr = kk + (lane >> 4);
#pragma unroll
            for (int m = 0; m < 4; ++m) a[m] = Bs[stg][kr][wc * 64 + m * 16 + (lane & 15)];
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) b[nn] = As[stg][kr][wr * 64 + nn * 16 + (lane & 15)];
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int nn = 0; nn < 4; ++nn) acc[m][nn] = M::mfma(a[m], b[nn], acc[m][nn]);
        }
    };
    int stg = 0;
    if (kAhead == 1) {
        Regs ra, rb;
        gload(ra, rb, t0);
        sstore(ra, rb, 0, t0);
        __syncthreads();
        for (int kc = t0; kc < t1; kc += kKc) {
            const bool more = kc + kKc < t1;
            if (more) gload(ra, rb, kc + kKc);
            mfma_stage(stg);
            if (more) sstore(ra, rb, stg ^ 1, kc + kKc);
            __syncthreads();
            stg ^= 1;
        }
    } else {
        // chunk c in LDS stage stg, chunk c+1 in one register set; the
        // step loads chunk c+2 into the other set (the one just stored)
        Regs ra0, rb0, ra1, rb1;
        gload(ra0, rb0, t0);
        if (t0 + kKc < t1) gload(ra1, rb1, t0 + kKc);
        sstore(ra0, rb0, 0, t0);
        __syncthreads();
        auto step = [&](Regs &la, Regs &lb, const Regs &sa, const Regs &sb, int kc) {
            if (kc + 2 * kKc < t1) gload(la, lb, kc + 2 * kKc);
            mfma_stage(stg);
            if (kc + kKc < t1) sstore(sa, sb, stg ^ 1, kc + kKc);
            __syncthreads();
            stg ^= 1;
        };
        for (int kc = t0; kc < t1; kc += 2 * kKc) {
            step(ra0, rb0, ra1, rb1, kc);
            if (kc + kKc < t1) step(ra1, rb1, ra0, rb0, kc + kKc);
        }
    }
}

// ---------------------------------------------------------------------------
// BF16 MFMA tile (v_mfma_f32_16x16x32_bf16, FP32 accumulation). Same tile
// geometry and output layout as the FP32 tile (accumulator value v of a lane
// is row 4 (lane / 16) + v of the 16 x 16 block), but the operands sit in
// LDS row-major ([row][k], kKc k-values plus 8 of padding per row) so that
// each lane reads the 8 consecutive k-values of one row that the
// instruction takes per lane as one 16-byte read. The A and B operands use
// the same lane -> k assignment, so the order of k inside the 32-wide step
// does not matter for the dot products.
// SPLIT = false: each operand x is rounded to bf16 (8 significant bits).
// SPLIT = true: x = hi + lo with hi = bf16(x), lo = bf16(x - hi), and the
// product is lo*lo + lo*hi + hi*lo + hi*hi in four MFMAs (smallest terms
// first): about 16 significant bits per operand at a quarter of the BF16
// rate, the split-bfloat16 scheme of Henry, Tang and Heinecke, "Leveraging
// the bfloat16 artificial intelligence datatype for higher-precision
// computations" (ARITH 2019), here with the lo*lo term kept: on the 100^3
// 27-point operator it halves the first solve's backward error, so one
// refinement step reaches 1e-12 where three products needed two (the
// tile's loads, not its MFMAs, bound its time).
typedef __bf16 v8bf_t __attribute__((ext_vector_type(8)));

template <int kKc>
__host__ __device__ constexpr int bf_lds_ld() { return kKc + 8; }
// LDS bytes: 2 stages x 2 operands (x 2 for hi / lo) of kT rows
template <bool SPLIT, int kKc>
__host__ __device__ constexpr int bf_lds_bytes() { return (SPLIT ? 2 : 1) * 2 * 2 * kT * bf_lds_ld<kKc>() * 2; }

__device__ __forceinline__ v4f_t mfma_bf16(v8bf_t a, v8bf_t b, v4f_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// GE: operand type in global memory (float: an FP32 panel copy; double: FP64
// data rounded on the way into LDS). Element (row i, panel column t) of A is
// Ab[t * ld + i], of B Bb[(t - boff) * ld + j]; rows i < f, j < j1 and panel
// columns [t0, t1) are valid, the rest read as 0 (buffer loads past the
// descriptor). Each thread loads one row's 8 consecutive k-values per group
// (the 256 threads cover 128 rows x kKc columns of each operand), converts
// them and writes one 16-byte LDS vector per operand (two with SPLIT).
template <bool SPLIT, bool ACTIVE, int kKc, typename GE>
__device__ __forceinline__ void tile_loop_bf16(const GE *Ab, const GE *Bb, int ld, int boff, int i0, int j0, int f,
                                               int j1, int t0, int t1, unsigned char *smem, v4f_t (&acc)[4][4]) {
    constexpr int kLd = bf_lds_ld<kKc>();
    constexpr int kGroups = kKc / 16;  // 8-column groups per thread and operand
    constexpr int kLoads = 8 * kGroups;
    constexpr int kPlane = 2 * kT * kLd;  // one operand's two stages, in bf16 elements
    __bf16 *lds = reinterpret_cast<__bf16 *>(smem);
    // planes: A (hi), B (hi), then with SPLIT A (lo), B (lo); stage s of a
    // plane starts at s * kT * kLd
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    constexpr uint32_t kOut = 0x80000000u;
    const int rl = tid & (kT - 1), th = tid >> 7;
    const uint32_t colb = static_cast<uint32_t>(ld) * static_cast<uint32_t>(sizeof(GE));
    const uint32_t voa = (i0 + rl < f) ? static_cast<uint32_t>(i0 + rl) * sizeof(GE) : kOut;
    const uint32_t vob = (j0 + rl < j1) ? static_cast<uint32_t>(j0 + rl) * sizeof(GE) : kOut;
    typedef GE Regs[kLoads];
    auto gload = [&](Regs &ra, Regs &rb, int kc) {
        const int nrec = min(kKc, t1 - kc) * static_cast<int>(colb);
        const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<GE *>(Ab + static_cast<int64_t>(kc) * ld), 0, nrec, kBufferWord3);
        const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<GE *>(Bb + static_cast<int64_t>(kc - boff) * ld), 0, nrec, kBufferWord3);
#pragma unroll
        for (int e = 0; e < kGroups; ++e)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t co = static_cast<uint32_t>(8 * (th + 2 * e) + u) * colb;
                ra[8 * e + u] = buffer_load<GE>(sa, voa + co);
                rb[8 * e + u] = buffer_load<GE>(sb, vob + co);
            }
    };
    auto sstore = [&](const Regs &ra, const Regs &rb, int stg) {
#pragma unroll
        for (int e = 0; e < kGroups; ++e) {
            const int o = stg * kT * kLd + rl * kLd + 8 * (th + 2 * e);
            v8bf_t ha, hb, la, lb;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const float xa = static_cast<float>(ra[8 * e + u]);
                const float xb = static_cast<float>(rb[8 * e + u]);
                ha[u] = static_cast<__bf16>(xa);
                hb[u] = static_cast<__bf16>(xb);
                if (SPLIT) {
                    la[u] = static_cast<__bf16>(xa - static_cast<float>(ha[u]));
                    lb[u] = static_cast<__bf16>(xb - static_cast<float>(hb[u]));
                }
            }
            *reinterpret_cast<v8bf_t *>(lds + o) = ha;
            *reinterpret_cast<v8bf_t *>(lds + kPlane + o) = hb;
            if (SPLIT) {
                *reinterpret_cast<v8bf_t *>(lds + 2 * kPlane + o) = la;
                *reinterpret_cast<v8bf_t *>(lds + 3 * kPlane + o) = lb;
            }
        }
    };
    auto mfma_stage = [&](int stg) {
        if (!ACTIVE) return;
        const __bf16 *As = lds + stg * kT * kLd;
        const __bf16 *Bs = lds + kPlane + stg * kT * kLd;
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 32) {
            const int ko = kk + 8 * (lane >> 4);
            v8bf_t a[4], b[4];
#pragma unroll
            for (int m = 0; m < 4; ++m)
                a[m] = *reinterpret_cast<const v8bf_t *>(Bs + (wc * 64 + m * 16 + (lane & 15)) * kLd + ko);
#pragma unroll
            for (int nn = 0; nn < 4; ++nn)
                b[nn] = *reinterpret_cast<const v8bf_t *>(As + (wr * 64 + nn * 16 + (lane & 15)) * kLd + ko);
            if (SPLIT) {
                v8bf_t al[4], bl[4];
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    al[m] = *reinterpret_cast<const v8bf_t *>(Bs + 2 * kPlane + (wc * 64 + m * 16 + (lane & 15)) * kLd + ko);
#pragma unroll
                for (int nn = 0; nn < 4; ++nn)
                    bl[nn] = *reinterpret_cast<const v8bf_t *>(As + 2 * kPlane + (wr * 64 + nn * 16 + (lane & 15)) * kLd + ko);
                // small terms first
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int nn = 0; nn < 4; ++nn) {
                        acc[m][nn] = mfma_bf16(al[m], bl[nn], acc[m][nn]);
                        acc[m][nn] = mfma_bf16(al[m], b[nn], acc[m][nn]);
                        acc[m][nn] = mfma_bf16(a[m], bl[nn], acc[m][nn]);
                    }
            }
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int nn = 0; nn < 4; ++nn) acc[m][nn] = mfma_bf16(a[m], b[nn], acc[m][nn]);
        }
    };
    // one register set: the loads of chunk c+1 are in flight during the
    // MFMAs of chunk c
    int stg = 0;
    Regs ra, rb;
    gload(ra, rb, t0);
    sstore(ra, rb, 0);
    __syncthreads();
    for (int kc = t0; kc < t1; kc += kKc) {
        const bool more = kc + kKc < t1;
        if (more) gload(ra, rb, kc + kKc);
        mfma_stage(stg);
        if (more) sstore(ra, rb, stg ^ 1);
        __syncthreads();
        stg ^= 1;
    }
}

// ---------------------------------------------------------------------------
// BF16 tile on operands already stored as bf16 in the packed panel layout:
// element (row i, panel column t) of a front with ld rows sits at
// P[(t / 8) * ld * 8 + i * 8 + t % 8], so one row's 8 consecutive k-values
// are one 16-byte vector, loaded by one raw buffer load and written to LDS
// as is (the LDS layout is tile_loop_bf16's). SPLIT: a second plane holds
// the low parts. Panel columns [t0, t1) with t0 a multiple of 8; the plane
// holds zeros in a front's columns past its last pivot, so the last group of
// a chunk ending at the front's last pivot reads zeros there.
typedef unsigned int v4u_t __attribute__((ext_vector_type(4)));
template <bool SPLIT, bool ACTIVE, int kKc>
__device__ __forceinline__ void tile_loop_bf16p(const uint16_t *Ah, const uint16_t *Al, const uint16_t *Bh,
                                                const uint16_t *Bl, int ld, int i0, int j0, int f, int j1, int t0,
                                                int t1, unsigned char *smem, v4f_t (&acc)[4][4]) {
    constexpr int kLd = bf_lds_ld<kKc>();
    constexpr int kVec = kKc / 16;        // 16-byte vectors per thread, operand and plane
    constexpr int kPlane = 2 * kT * kLd;  // one operand's two stages, in bf16 elements
    __bf16 *lds = reinterpret_cast<__bf16 *>(smem);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    constexpr uint32_t kOut = 0x80000000u;
    const int rl = tid & (kT - 1), th = tid >> 7;
    const uint32_t grp = static_cast<uint32_t>(ld) * 16u;  // bytes of one 8-column group
    const uint32_t voa = (i0 + rl < f) ? static_cast<uint32_t>(i0 + rl) * 16u : kOut;
    const uint32_t vob = (j0 + rl < j1) ? static_cast<uint32_t>(j0 + rl) * 16u : kOut;
    typedef v4u_t Regs[SPLIT ? 2 * kVec : kVec];
    auto gload = [&](Regs &ra, Regs &rb, int kc) {
        const int ng = (min(kKc, t1 - kc) + 7) / 8;  // groups of the chunk inside [t0, t1)
        const int nrec = ng * static_cast<int>(grp);
        const int64_t go = static_cast<int64_t>(kc / 8) * ld * 8;
        const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(Ah + go), 0, nrec, kBufferWord3);
        const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(Bh + go), 0, nrec, kBufferWord3);
#pragma unroll
        for (int e = 0; e < kVec; ++e) {
            const uint32_t co = static_cast<uint32_t>(th + 2 * e) * grp;
            ra[e] = __builtin_amdgcn_raw_buffer_load_b128(sa, static_cast<int>(voa + co), 0, 0);
            rb[e] = __builtin_amdgcn_raw_buffer_load_b128(sb, static_cast<int>(vob + co), 0, 0);
        }
        if (SPLIT) {
            const __amdgpu_buffer_rsrc_t la = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(Al + go), 0, nrec, kBufferWord3);
            const __amdgpu_buffer_rsrc_t lb = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(Bl + go), 0, nrec, kBufferWord3);
#pragma unroll
            for (int e = 0; e < kVec; ++e) {
                const uint32_t co = static_cast<uint32_t>(th + 2 * e) * grp;
                ra[kVec + e] = __builtin_amdgcn_raw_buffer_load_b128(la, static_cast<int>(voa + co), 0, 0);
                rb[kVec + e] = __builtin_amdgcn_raw_buffer_load_b128(lb, static_cast<int>(vob + co), 0, 0);
            }
        }
    };
    auto sstore = [&](const Regs &ra, const Regs &rb, int stg) {
#pragma unroll
        for (int e = 0; e < kVec; ++e) {
            const int o = stg * kT * kLd + rl * kLd + 8 * (th + 2 * e);
            *reinterpret_cast<v4u_t *>(lds + o) = ra[e];
            *reinterpret_cast<v4u_t *>(lds + kPlane + o) = rb[e];
            if (SPLIT) {
                *reinterpret_cast<v4u_t *>(lds + 2 * kPlane + o) = ra[kVec + e];
                *reinterpret_cast<v4u_t *>(lds + 3 * kPlane + o) = rb[kVec + e];
            }
        }
    };
    auto mfma_stage = [&](int stg) {
        if (!ACTIVE) return;
        const __bf16 *As = lds + stg * kT * kLd;
        const __bf16 *Bs = lds + kPlane + stg * kT * kLd;
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 32) {
            const int ko = kk + 8 * (lane >> 4);
            v8bf_t a[4], b[4];
#pragma unroll
            for (int m = 0; m < 4; ++m)
                a[m] = *reinterpret_cast<const v8bf_t *>(Bs + (wc * 64 + m * 16 + (lane & 15)) * kLd + ko);
#pragma unroll
            for (int nn = 0; nn < 4; ++nn)
                b[nn] = *reinterpret_cast<const v8bf_t *>(As + (wr * 64 + nn * 16 + (lane & 15)) * kLd + ko);
            if (SPLIT) {
                v8bf_t al[4], bl[4];
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    al[m] = *reinterpret_cast<const v8bf_t *>(Bs + 2 * kPlane + (wc * 64 + m * 16 + (lane & 15)) * kLd + ko);
#pragma unroll
                for (int nn = 0; nn < 4; ++nn)
                    bl[nn] = *reinterpret_cast<const v8bf_t *>(As + 2 * kPlane + (wr * 64 + nn * 16 + (lane & 15)) * kLd + ko);
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int nn = 0; nn < 4; ++nn) {
                        acc[m][nn] = mfma_bf16(al[m], bl[nn], acc[m][nn]);
                        acc[m][nn] = mfma_bf16(al[m], b[nn], acc[m][nn]);
                        acc[m][nn] = mfma_bf16(a[m], bl[nn], acc[m][nn]);
                    }
            }
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int nn = 0; nn < 4; ++nn) acc[m][nn] = mfma_bf16(a[m], b[nn], acc[m][nn]);
        }
    };
    // global loads of chunk c+1 in flight during the MFMAs of chunk c
    int stg = 0;
    Regs ra, rb;
    gload(ra, rb, t0);
    sstore(ra, rb, 0);
    __syncthreads();
    for (int kc = t0; kc < t1; kc += kKc) {
        const bool more = kc + kKc < t1;
        if (more) gload(ra, rb, kc + kKc);
        mfma_stage(stg);
        if (more) sstore(ra, rb, stg ^ 1);
        __syncthreads();
        stg ^= 1;
    }
}

// ---------------------------------------------------------------------------
// 256 x 256 tile of eight waves on operands in the packed bf16 layout (for
// the rank-512 updates of large fronts). Wave w covers A rows
// [i0 + 64 (w % 4), +64) x B rows [j0 + 128 (w / 4), +128): 8 x 4 MFMA
// blocks of 16 x 16 in acc[m][nn] (m: B-row block, nn: A-row block, the
// transposed blocks of tile_loop_bf16p). Per 32-column chunk the workgroup
// stages 64 KB (split) for four times the products of the 128 x 128 tile's
// 32 KB. Each accumulator receives the same MFMAs on the same fragments in
// the same order as in tile_loop_bf16p, so the sums agree bit for bit.
// LDS: two stages of the four planes, 160 KB (one workgroup per CU).
constexpr int kTW = 256;
constexpr int kWideThreads = 512;
template <bool SPLIT, int kKc>
__host__ __device__ constexpr int bfw_lds_bytes() { return (SPLIT ? 2 : 1) * 2 * 2 * kTW * bf_lds_ld<kKc>() * 2; }
template <bool SPLIT, int kKc>
__device__ __forceinline__ void tile_loop_bf16p_w(const uint16_t *Ah, const uint16_t *Al, const uint16_t *Bh,
                                                  const uint16_t *Bl, int ld, int i0, int j0, int f, int j1, int t0,
                                                  int t1, bool active, unsigned char *smem, v4f_t (&acc)[8][4]) {
    constexpr int kLd = bf_lds_ld<kKc>();
    constexpr int kVec = kKc / 16;         // 16-byte vectors per thread, operand and plane
    constexpr int kPlane = 2 * kTW * kLd;  // one operand's two stages, in bf16 elements
    __bf16 *lds = reinterpret_cast<__bf16 *>(smem);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int wr = w & 3, wc = w >> 2;
    constexpr uint32_t kOut = 0x80000000u;
    const int rl = tid & (kTW - 1), th = tid >> 8;
    const uint32_t grp = static_cast<uint32_t>(ld) * 16u;  // bytes of one 8-column group
    const uint32_t voa = (i0 + rl < f) ? static_cast<uint32_t>(i0 + rl) * 16u : kOut;
    const uint32_t vob = (j0 + rl < j1) ? static_cast<uint32_t>(j0 + rl) * 16u : kOut;
    typedef v4u_t Regs[SPLIT ? 2 * kVec : kVec];
    auto gload = [&](Regs &ra, Regs &rb, int kc) {
        const int ng = (min(kKc, t1 - kc) + 7) / 8;
        const int nrec = ng * static_cast<int>(grp);
        const int64_t go = static_cast<int64_t>(kc / 8) * ld * 8;
        const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(Ah + go), 0, nrec, kBufferWord3);
        const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(Bh + go), 0, nrec, kBufferWord3);
#pragma unroll
        for (int e = 0; e < kVec; ++e) {
            const uint32_t co = static_cast<uint32_t>(th + 2 * e) * grp;
            ra[e] = __builtin_amdgcn_raw_buffer_load_b128(sa, static_cast<int>(voa + co), 0, 0);
            rb[e] = __builtin_amdgcn_raw_buffer_load_b128(sb, static_cast<int>(vob + co), 0, 0);
        }
        if (SPLIT) {
            const __amdgpu_buffer_rsrc_t la = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(Al + go), 0, nrec, kBufferWord3);
            const __amdgpu_buffer_rsrc_t lb = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(Bl + go), 0, nrec, kBufferWord3);
#pragma unroll
            for (int e = 0; e < kVec; ++e) {
                const uint32_t co = static_cast<uint32_t>(th + 2 * e) * grp;
                ra[kVec + e] = __builtin_amdgcn_raw_buffer_load_b128(la, static_cast<int>(voa + co), 0, 0);
                rb[kVec + e] = __builtin_amdgcn_raw_buffer_load_b128(lb, static_cast<int>(vob + co), 0, 0);
            }
        }
    };
    auto sstore = [&](const Regs &ra, const Regs &rb, int stg) {
#pragma unroll
        for (int e = 0; e < kVec; ++e) {
            const int o = stg * kTW * kLd + rl * kLd + 8 * (th + 2 * e);
            *reinterpret_cast<v4u_t *>(lds + o) = ra[e];
            *reinterpret_cast<v4u_t *>(lds + kPlane + o) = rb[e];
            if (SPLIT) {
                *reinterpret_cast<v4u_t *>(lds + 2 * kPlane + o) = ra[kVec + e];
                *reinterpret_cast<v4u_t *>(lds + 3 * kPlane + o) = rb[kVec + e];
            }
        }
    };
    auto mfma_stage = [&](int stg) {
        const __bf16 *As = lds + stg * kTW * kLd;
        const __bf16 *Bs = lds + kPlane + stg * kTW * kLd;
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 32) {
            const int ko = kk + 8 * (lane >> 4);
            v8bf_t b[4], bl[4];
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) {
                b[nn] = *reinterpret_cast<const v8bf_t *>(As + (wr * 64 + nn * 16 + (lane & 15)) * kLd + ko);
                if (SPLIT)
                    bl[nn] = *reinterpret_cast<const v8bf_t *>(As + 2 * kPlane + (wr * 64 + nn * 16 + (lane & 15)) * kLd + ko);
            }
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int ro = (wc * 128 + m * 16 + (lane & 15)) * kLd + ko;
                const v8bf_t a = *reinterpret_cast<const v8bf_t *>(Bs + ro);
                if (SPLIT) {
                    const v8bf_t al = *reinterpret_cast<const v8bf_t *>(Bs + 2 * kPlane + ro);
#pragma unroll
                    for (int nn = 0; nn < 4; ++nn) {
                        acc[m][nn] = mfma_bf16(al, bl[nn], acc[m][nn]);
                        acc[m][nn] = mfma_bf16(al, b[nn], acc[m][nn]);
                        acc[m][nn] = mfma_bf16(a, bl[nn], acc[m][nn]);
                    }
                }
#pragma unroll
                for (int nn = 0; nn < 4; ++nn) acc[m][nn] = mfma_bf16(a, b[nn], acc[m][nn]);
            }
        }
    };
    int stg = 0;
    Regs ra, rb;
    gload(ra, rb, t0);
    sstore(ra, rb, 0);
    __syncthreads();
    for (int kc = t0; kc < t1; kc += kKc) {
        const bool more = kc + kKc < t1;
        if (more) gload(ra, rb, kc + kKc);
        if (active) mfma_stage(stg);
        if (more) sstore(ra, rb, stg ^ 1);
        __syncthreads();
        stg ^= 1;
    }
}

}  // namespace smfgpu
