// Internal declarations shared by the host orchestration and the HIP
// kernels of the smfgpu numeric phase.
#pragma once

#include "../../include/smfgpu.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace smfgpu {

// A dispatch packet counts a grid's work-items in 32 bits: a launch over a
// list of items (one workgroup each, up to 1024 threads) covers at most
// this many, and longer lists are launched in pieces.
constexpr int64_t kMaxLaunchGroups = (int64_t{1} << 22) - 1;

int set_error(int code, const std::string &msg);
const char *last_error();  // message of the last set_error (this thread)

// ---------------------------------------------------------------------
// Host analysis result (the AnalysisKeep fields of the numeric phase).
struct Analysis {
    int32_t n = 0, nsup = 0;
    std::vector<int32_t> sn_start, sn_end, sn_parent;
    std::vector<int32_t> child_ptr, child_idx;
    std::vector<int32_t> front_row_ptr, front_row_idx;
    std::vector<int32_t> pcol_ptr, prow_idx, orig_to_slot, perm;
    void describe(smfgpu_analysis_desc &d) const;
};

int analyse(int32_t n, const int32_t *col_ptr, const int32_t *row_idx, const int32_t *perm,
            int32_t nemin, Analysis &out);

// ---------------------------------------------------------------------
// Device data layout.
//
// Front s (supernode index = postorder position) has f rows, p pivot
// columns, q = f - p. Its columns [0, p) live in the factor array at lofs
// (f x p, column-major, leading dimension f: the FactorKeep layout of
// SolveStep::factor_offset), its trailing q x q contribution block in the
// CB array at cbofs (column-major, leading dimension q, lower triangle).
// Element (r, c), r >= c, of the front is therefore
//   c <  p : L[lofs + c * f + r]
//   c >= p : CB[cbofs + (c - p) * q + (r - p)].
struct FrontDev {
    int32_t f, p, q, col0;  // col0: first pivot column (permuted numbering)
    int64_t lofs;           // factor offset
    int64_t cbofs;          // contribution-block offset
    int64_t rofs;           // row-index offset (rows[rofs .. rofs + f))
    int64_t l16ofs;         // offset of the pivot columns in the packed bf16 planes (f x p rounded up to 8 columns)
};

// One column of one front to assemble (a workgroup per task).
struct AsmTask {
    int32_t s, c;       // front, local column
    int32_t cc0, cc1;   // range in the child-contribution list
};
// A child's CB column that lands in the task's column: child ch's column jj.
struct ChildCol {
    int32_t ch, jj;
};

// Dense-kernel work items (expanded on the host at plan time).
struct PotrfItem {
    int32_t s, k, kb, pad;
};
struct TrsmItem {
    int32_t s, k, kb, r0;  // rows [r0, min(r0 + 64, f)) of columns [k, k + kb)
};
struct SyrkItem {
    int32_t s, i0, j0, j1;  // tile rows [i0, i0 + T), cols [j0, min(j0 + T, j1)), i >= j
    int32_t t0, t1, prec, pad;  // panel columns [t0, t1); prec: 0 fp64, 1 fp32, 2 bf16
};

// One workgroup of a solve step on front s.
// Forward: update rows [a0, a1) by pivot block [k, k + kb) (kb = 0: none),
//   then solve the triangle of block [tk, tk + tkb) (tkb = 0: none; its rows
//   are then [a0, a1)); flags & 2: also update the rows below the block.
// Backward: update columns [a0, a1) by rows [k, k + kb) (own pivot rows or
//   rows below the pivots), flags & 1: by atomic adds; then the triangle.
struct SolveItem {
    int32_t s, k, kb, a0, a1, tk, tkb, flags;
};

enum LaunchKind : int32_t {
    K_ASSEMBLE = 0,
    K_POTRF = 1,
    K_TRSM = 2,
    K_SYRK = 3,
    K_SELECT = 4,  // learned precision selector over the level's candidate fronts (items: SelFront)
};

struct Launch {
    int32_t kind;
    int32_t level;
    int64_t first;  // first item in the kind's item array
    int64_t count;  // number of items (= workgroups)
    int64_t red;    // K_SYRK: the last `red` items belong to fronts that may run below FP64
};

// ---------------------------------------------------------------------
// Kernel launchers (kernels.hip). All asynchronous on `st`.
struct DevPtrs {
    const FrontDev *fronts;
    const int32_t *rows;    // front rows (global permuted numbering)
    const int32_t *cmap;    // per child: positions of its CB rows in the parent
    const int64_t *cmofs;   // per supernode: offset into cmap
    const int32_t *pcol_ptr;
    const int32_t *amap;    // per permuted entry: local row in its column's front
    const double *av;       // permuted values
    double *L;
    float *L32;             // FP32 copy of the panel columns of reduced-precision fronts (null: none)
    uint16_t *L16h, *L16l;  // bf16 copies (high part, low part) of the panel columns of BF16 / split-BF16
                            // fronts in the packed layout: (row i, column t) at l16ofs + (t / 8) f 8 + 8 i + t % 8
    double *CB;
    int32_t *flag;          // [0]: smallest failing supernode + 1 (SPD); INT32_MAX when none
    const int8_t *prec;     // per supernode: Schur-update precision 0-3 (null: all FP64)
    double *junk;           // kThreads doubles: target of masked-out stores (contents meaningless)
    float4 *colstat;        // per pivot column (permuted numbering), written by the assembly when not null:
                            // 2-norm of the column's lower part, |diagonal|, max |below the diagonal|
};

void launch_map_values(hipStream_t st, const double *vals, int64_t nvals, const int64_t *slot_ptr,
                       const int64_t *slot_src, int64_t nslots, double *av, double *redbuf);
// S A S in place over the permuted slots (scale in the original numbering),
// the unscaled values into raw; redbuf[0] becomes max|S A S| (bits), unless
// it marks a non-finite input; uses redbuf[3].
void launch_scale_values(hipStream_t st, int32_t n, const int64_t *fr_ptr, const int32_t *fr_col,
                         const int64_t *fr_slot, const int32_t *perm, const double *scale, double *av, double *raw,
                         double *redbuf);
// *flag (as an unsigned 64-bit integer) becomes 1 when an entry of scale[0..n)
// is not finite and positive, else 0
void launch_check_scale(hipStream_t st, const double *scale, int32_t n, double *flag);
// iters equilibration passes over the permuted slots' values av, from s = 1,
// ping-ponging between s0 and s1 (n doubles each); *result: the final s
// (original numbering)
void launch_equilibrate(hipStream_t st, int32_t n, const int64_t *fr_ptr, const int32_t *fr_col,
                        const int64_t *fr_slot, const int32_t *perm, const double *av, int32_t iters, double *s0,
                        double *s1, double **result);
// fmax: the longest front of the launch's level; levels with fmax <=
// wave_rows take a wavefront per column, others a workgroup per column with
// LDS chunks of at most chunk_cap rows
constexpr int32_t kAsmWaveRowsMax = 4096;
void launch_assemble(hipStream_t st, const DevPtrs &d, const AsmTask *tasks, int64_t ntasks, const ChildCol *cc,
                     int32_t fmax, int32_t chunk_cap, int32_t wave_rows);
void launch_potrf(hipStream_t st, const DevPtrs &d, const PotrfItem *it, int64_t n);
void launch_trsm(hipStream_t st, const DevPtrs &d, const TrsmItem *it, int64_t n);
// n items; the reduced-precision kernels run the last n_red of them, one
// launch per precision code in red_modes (bit c: code c)
void launch_syrk(hipStream_t st, const DevPtrs &d, const SyrkItem *it, int64_t n, int64_t n_red, int ahead,
                 int red_modes);
void launch_zero_upper(hipStream_t st, const DevPtrs &d, const PotrfItem *it, int64_t n);

// Solves: y is the permuted right-hand side block, n x nrhs (ld n).
// scale (original numbering, null: none) multiplies the gathered and the
// scattered entries
void launch_gather_rhs(hipStream_t st, const int32_t *perm, const double *b, int64_t ldb, double *y,
                       int32_t n, int32_t nrhs, const double *scale = nullptr);
void launch_scatter_rhs(hipStream_t st, const int32_t *perm, const double *y, double *b, int64_t ldb,
                        int32_t n, int32_t nrhs, const double *scale = nullptr);
// One step of a sweep (nrhs <= 4): a workgroup per item.
constexpr int32_t kSolveBlock = 128;  // pivot columns per step
constexpr int32_t kSolveSeg = 512;    // rows per backward update item
// One sweep (forward or backward) as a single persistent launch: items in
// sweep order, step_of[i] their steps, step_cnt[t] the items of step t;
// an item waits until every item of the step before has finished (the
// kernel boundary of the per-step launches, as a counter). ctr holds
// 1 + steps counters (cleared by the launch); err is set if a wait ran out.
void launch_sweep(hipStream_t st, const DevPtrs &d, bool forward, const SolveItem *items, const int32_t *step_of,
                  const int32_t *step_cnt, int32_t nitems, int32_t nsteps, int32_t *ctr, int32_t *err, double *y,
                  int32_t n, int32_t nrhs);
void launch_fwd_step(hipStream_t st, const DevPtrs &d, const SolveItem *items, int64_t count, double *y,
                     int32_t n, int32_t nrhs);
void launch_bwd_step(hipStream_t st, const DevPtrs &d, const SolveItem *items, int64_t count, double *y,
                     int32_t n, int32_t nrhs);

// Refinement: full symmetric CSR in permuted numbering (slot into av).
// r = b - A x (double-double) and per column c the maxima stat[4c..4c+3] =
// max |r_i|, max_i |r_i| / ((|A||x|)_i + |b_i|), max |x_i|, max |b_i|
// (x, b, r: n x nrhs, ld n; stat: 4 nrhs doubles, cleared by the launch).
void launch_residual(hipStream_t st, int32_t n, const int64_t *rp, const int32_t *ci,
                     const int64_t *slot, const double *av, const double *x, const double *b,
                     double *r, double *stat, int32_t nrhs);
void launch_spmv(hipStream_t st, int32_t n, const int64_t *rp, const int32_t *ci, const int64_t *slot,
                 const double *av, const double *z, double *w);
void launch_rowabs(hipStream_t st, int32_t n, const int64_t *rp, const int64_t *slot, const double *av,
                   double *out);
// x = y + d
void launch_axpy(hipStream_t st, int64_t n, const double *y, const double *d, double *x);
// FGMRES building blocks (vectors of length n, columns ld n, scalars on
// the device). dots: out[i] = V_i . w for nvec columns (mode 1: out[i] +=,
// mode 2: out[i] = sqrt); part: kDotParts * nvec doubles of scratch.
constexpr int kDotParts = 256;
void launch_dots(hipStream_t st, int64_t n, const double *V, int32_t nvec, const double *w, double *part,
                 double *out, int32_t mode);
void launch_vsub(hipStream_t st, int64_t n, const double *V, int32_t nvec, const double *h, double *w);
void launch_vdiv(hipStream_t st, int64_t n, const double *w, const double *d, double *out);
void launch_lincomb(hipStream_t st, int64_t n, const double *x0, const double *Z, int32_t nvec, const double *y,
                    double *out);
void launch_reduce_max(hipStream_t st, const double *v, int64_t n, double *out);

// ---------------------------------------------------------------------
// Learned per-supernode precision selector (selector.hip). An MLP
// kSelNF -> kSelH -> kSelH -> kSelNO (tanh hidden layers) maps a front's
// features to the predicted log10 relative error its Schur update would
// have in FP32, split BF16 and BF16 (outputs 0, 1, 2). Weight layout
// (floats): W1[kSelH][kSelNF], b1[kSelH], W2[kSelH][kSelH], b2[kSelH],
// W3[kSelNO][kSelH], b3[kSelNO], then the inputs' mean[kSelNF] and
// inverse standard deviation[kSelNF].
constexpr int kSelNF = 10, kSelH = 32, kSelNO = 3;
constexpr int kSelParams = kSelH * kSelNF + kSelH + kSelH * kSelH + kSelH + kSelNO * kSelH + kSelNO;
constexpr int kSelWeights = kSelParams + 2 * kSelNF;
// output index of each precision code (1 FP32, 2 BF16, 3 split BF16)
constexpr int kSelOutOfCode[4] = {-1, 0, 2, 1};

// One front handed to the selector: rows, pivot columns (eligible columns
// for LDL^T), delayed columns received, and where its pivot columns' stats
// are (colstat index of column 0).
struct SelFront {
    int32_t s, f, p, nd;
    int64_t cs0;
    int64_t fofs;  // LDL^T: the front's square in the arena
    int64_t sofs;  // LDL^T training pairs: the square's copies in the snapshot buffer
};
struct SelPrm {
    const float *w;            // kSelWeights (null: features only, precision 0)
    float log_tol;             // a precision is taken when its predicted log10 error <= log_tol
    int32_t mask;              // allowed precision codes: bit c for code c (1 FP32, 2 BF16, 3 split BF16)
    int32_t scaled;            // scaling state of the values (feature 9)
    const unsigned long long *maxabs_bits;  // max |A| as the bits of a double
};
// features of each front into feat[s * kSelNF ..] and its precision into
// prec[s] (or, for LDL^T, BigLd::prec of big[i] when big is not null)
void launch_select(hipStream_t st, const SelFront *fr, int32_t nf, const float4 *colstat, const SelPrm &prm,
                   float *feat, int8_t *prec);
// inference over rows of raw features: Y[i * kSelNO ..]
void launch_mlp_predict(hipStream_t st, const float *w, const float *X, int32_t n, float *Y);
// one training step's gradient (mean squared error over the labels with
// mask 1), per-block partial sums into part (nblk x (kSelParams + 1): the
// last entry the block's summed squared error); returns nblk
int32_t launch_mlp_grad(hipStream_t st, const float *w, const float *X, const float *Y, const float *M, int32_t n,
                        float *part);
// Adam step over the parameters from nblk partial gradients (divided by
// the label count cnt); the loss of the step into loss[0]
void launch_mlp_adam(hipStream_t st, float *w, float *m, float *v, const float *part, int32_t nblk, float cnt,
                     int32_t step, float lr, float *loss);
// labels of the SPD candidate fronts of a factorization: for each CB tile
// item (t0 = 0, t1 = p), max |Z_64 - Z_m| over the tile (m = FP32, split
// BF16, BF16: slots 0, 1, 2) and max |S| (slot 3) into lab[4 s ..] as float
// bits (cleared by the caller); zs: n x 128 x 128 doubles of scratch
// (n <= kMaxLaunchGroups)
void launch_label_spd(hipStream_t st, const DevPtrs &d, const SyrkItem *items, int64_t n, double *zs,
                      unsigned int *lab);
// LDL^T: statistics of the eligible columns (fr[i].p of them, rows from the
// diagonal down) of the f x f squares at arena + fr[i].fofs, into
// colstat[fr[i].cs0 ..]; maxcols: the largest fr[i].p
void launch_sq_colstat(hipStream_t st, const double *arena, const SelFront *fr, int32_t nf, int32_t maxcols,
                       float4 *colstat);
// LDL^T labels: over the contribution block [p, f)^2 (lower) of each square,
// max |snap[k * stride + sofs ..] - square| for k = 0, 1, 2 (the updates in
// FP32, split BF16, BF16) into lab[4 s + k] and max |square| into
// lab[4 s + 3], as float bits; maxq: the largest f - p
void launch_label_sq(hipStream_t st, const double *arena, const double *snap, int64_t stride, const SelFront *fr,
                     int32_t nf, int32_t maxq, unsigned int *lab);

}  // namespace smfgpu
