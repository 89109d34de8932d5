// Device data of the indefinite (LDL^T) numeric phase. Front sizes grow with
// the delayed pivots children pass up, so they are known only when a
// level starts: the host lays the level out (offsets into the arenas
// below) after reading back its children's results.
#pragma once

#include "smfgpu_internal.hpp"

namespace smfgpu {

// Front s of the current factorization. Its f x f square (column-major,
// leading dimension f, lower triangle) lives in the arena at fofs from its
// assembly until its parent has assembled the trailing (f - p) square, the
// contribution block (delayed rows first). Columns [0, pstore) are copied
// to the factor store at lofs (leading dimension f) in the FactorKeep
// layout: unit L, D on the diagonal, a 2x2 block's off-diagonal in slot
// (c, c+1), zeros elsewhere above the diagonal.
struct FrontLd {
    int32_t f, ne, nd, psym;  // rows, eligible columns, delayed columns received, own pivots
    int32_t col0;             // first own pivot column (permuted numbering)
    int32_t root;             // 1: no parent (unaccepted columns become zero pivots)
    int32_t p, pstore;        // accepted pivots; stored pivot columns
    int32_t npos, nneg, nzero, n2x2;
    int32_t delayed;
    int32_t dev;              // 1: laid out on the device (its square is not in the host's pool)
    int32_t nstatic;          // pivots whose |d| static pivoting raised
    // blocked fronts: first-pass panels that stalled, partner pairs pulled,
    // failed columns rotated to the end, second-pass full searches
    int32_t stalls, pulls, rotated, searches;
    double growth;            // largest column max / |d| over its static pivots
    double flops;
    int64_t fofs;             // arena offset of the square
    int64_t lofs;             // factor-store offset of the stored columns
    int64_t rofs;             // offset of the row ids / tags
    int64_t wofs;             // offset of the panel workspace W (blocked fronts)
    int64_t dofs;             // offset of the pivots' D^-1 entries
    int64_t cmofs;            // offset of this front's CB-row map into its parent
};

struct LdltDev {
    FrontLd *fronts;          // per supernode
    double *arena;            // front squares
    double *fac;              // stored factor columns
    int32_t *rows;            // per front: f row ids (permuted numbering), pivoted order
    int8_t *tags;             // per front: pivot tags (1, 2, -1, 0), same offsets as rows
    double *work;
    double *dinv;             // 3 per pivot
    int8_t *dkind;            // 1, 2, 0 per pivot
    const int32_t *cmap;      // CB-row maps
    const int32_t *pcol_ptr;
    const int32_t *amap;      // per permuted entry: position among its front's symbolic rows
    const double *av;
    double *junk;             // target of masked-out stores of the MFMA updates (contents meaningless)
    uint16_t *xpl;            // bf16 operand planes of the large fronts' rank-kBigWork updates (BF16 / split BF16):
                              // per front at 4 wofs, planes A hi, A lo, B hi, B lo of kBigWork f entries each
};

// Large fronts (multi-workgroup blocked LDL^T): one state per front,
// advanced by a panel kernel (one workgroup per front: the blocked
// algorithm's pivot logic on a panel of nb columns), an update kernel (many
// workgroups, MFMA) and a stall kernel (partner pulls / rotation), launched
// in turn by the host until every front is done. The per-panel deferred
// updates go to the columns of a kBigOuter-wide outer block; the pivots'
// updates of the columns beyond it are held back (their unscaled columns
// kept in W) and applied in one flush when the outer block ends, before a
// stall's pulls or rotation, and before the second pass.
// BF16 and split-BF16 fronts on bf16 planes take wider outer blocks (and
// contribution-block chunks), 1024 by default: their 256 x 256 tiles then
// read and write the FP64 entries once per rank-1024 update instead of
// once per rank-512 one, which bounds them (the FP32 accumulators sum
// twice as many products).
constexpr int32_t kBigOuter = 512;
constexpr int32_t kBigOuterMax = 1024;
// W columns per front: an outer block's pivots (L*D chunks of the final
// update), then a speculative group's final entries (kGroupSpec)
constexpr int32_t kBigWork = kBigOuterMax + 64;
constexpr int32_t kGroupSpec = kBigOuterMax;
static_assert(kBigOuterMax >= kBigOuter, "W holds the pivots of an outer block");
struct BigLd {
    int32_t s;                  // supernode
    int32_t k, live, k_paired;  // next pivot position, end of the live eligible columns, pulls' guard
    int32_t pass;               // 0 first pass, 1 second pass, 2 done
    int32_t full;               // second pass: next step is a full-search pivot
    int32_t stalled;            // first pass: the last panel stalled
    int32_t ps, pl;             // the last panel
    int32_t oe;                 // outer block end
    int32_t outer;              // outer block width (kBigOuter, or kBigOuterWide on planes)
    int32_t pend0;              // pivots [pend0, k) held back for the columns >= oe; W column 0
    int32_t u_c0, u_c1, u_p0, u_p1;  // update request: columns [u_c0, u_c1), rows [u_c0, f), pivots [u_p0, u_p1)
    int32_t x_c0, x_p0, x_p1;        // flush request: columns [x_c0, ne), rows [x_c0, f), pivots [x_p0, x_p1)
    int32_t w_base;             // pivot of W column 0 for both requests
    int32_t cnt[3];
    int32_t stalls, pulls, rotated, searches;  // as FrontLd's
    int32_t spec;               // 1: the next first-pass iteration tries a speculative group
    int32_t g_apply, g_ps, g_n; // accepted group to write: columns [g_ps, g_ps + g_n)
    int32_t g_cols;             // the group's width (its diagonal region's rows)
    int32_t prec;               // precision code of the MFMA updates: 0 FP64, 1 FP32, 2 BF16, 3 split BF16
    double flops;
};
constexpr int32_t kGroupCols = 64;   // columns of a speculative group (four panels of 16)
constexpr int32_t kGroupRows = 256;  // rows per workgroup below a group's diagonal region
static_assert(kBigWork - kGroupSpec == kGroupCols, "W holds a group's speculative columns");
constexpr int64_t kGroupBuf = 192 + 64 * 16 + 2 * 64 * 64;  // group buffer per front before the per-workgroup maxima

// Static pivoting (control.hpp:333-364): sp = static_pivot_tol * max|A|
// (0: off), fallback = fallback_static_pivot_tol * max|A| for roots and
// fronts that receive more delayed columns than their width; flags as
// SMFGPU_STATIC_*.
struct StaticPrm {
    double sp, fallback;
    int32_t flags;
};
constexpr int32_t kStaticAsZero = 1, kStaticZeroStaysZero = 2, kStaticInOrder = 4;
// whether front F pivots statically, and with which value
__host__ __device__ inline double front_static_value(const StaticPrm &S, int32_t root, int32_t nd, int32_t psym) {
    return (S.fallback > 0.0 && (root != 0 || nd > psym)) ? S.fallback : S.sp;
}

// Levels of small fronts laid out on the device: per level, a layout
// kernel (each front's size from its children's delays, offsets by a bump
// of the counters below) and one workgroup per front that builds the row
// list, assembles, factors (the unblocked kernel's steps) and stores the
// front with the square in LDS. A level whose fronts do not all fit is
// refused as a whole before any of it is factored: the counters record the
// first refused level and the reason, and the launches of later levels
// return at once.
constexpr int32_t kDevFrontMax = 95;   // largest realized front of a device level (blocked_min_front - 1)
constexpr int32_t kDevSymMax = 64;     // largest symbolic front of a level laid out on the device
enum : int32_t { kDcSq = 0, kDcFac = 1, kDcRows = 2, kDcDinv = 3, kDcLevel = 4, kDcWhy = 5, kDcN = 6 };
constexpr unsigned long long kDevWhySize = 1, kDevWhyRoom = 2;  // refusal: a front too large / a buffer full
struct DevLevel {
    const FrontDev *sym;        // symbolic fronts (f, p, col0, rofs)
    const int32_t *sym_rows;    // symbolic row ids
    const int32_t *sym_cmap;    // per child: its symbolic CB rows' positions among the parent's symbolic rows
    const int64_t *sym_cmofs;
    const int32_t *child_ptr, *child_idx, *parent;
    unsigned long long *ctr;    // kDcN counters: next free square / factor / row / D^-1 offsets, refused level, reason
    int64_t cap[4];             // ends of the room for the four offsets
    int32_t fmax;               // largest realized front taken
};
void launch_dev_layout(hipStream_t st, const LdltDev &d, const DevLevel &L, const int32_t *flist, int32_t nf,
                       int32_t lv);
void launch_dev_fronts(hipStream_t st, const LdltDev &d, const DevLevel &L, const int32_t *flist, int32_t nf,
                       int32_t lv, double u, double eff_small, double small_pivot, const StaticPrm &S, bool wave);

void launch_ldlt_small(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double u, double eff_small,
                       double small_pivot, const StaticPrm &S);
void launch_ldlt_blocked(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double u,
                         double eff_small, double small_pivot, int32_t nb, const StaticPrm &S);
void launch_ldlt_assemble(hipStream_t st, const LdltDev &d, const AsmTask *tasks, int64_t ntasks, const ChildCol *cc);
// epilogue (pivot counts, D^-1) of the fronts in flist; it also stores the
// factor columns of the first n_inline fronts, launch_ldlt_store those of
// the rest (many workgroups per front)
void launch_ldlt_epilogue(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, int32_t n_inline);
void launch_ldlt_store(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, int32_t max_cols);
void launch_ldlt_diag(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double *y, int32_t n,
                      int32_t nrhs);
// the diagonal-only solve job: D from the stored columns, by division
void launch_ldlt_diag_div(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double *y, int32_t n,
                          int32_t nrhs);
void launch_front_patch(hipStream_t st, FrontLd *fronts, const FrontLd *patch, const int32_t *ids, int32_t np);
// one step of the blocked forward / backward sweep (items of all fronts of
// a level; SolveItem flags bit 1: triangle item)
void launch_ldlt_fwd_step(hipStream_t st, const LdltDev &d, const SolveItem *items, int64_t count, double *y,
                          int32_t n, int32_t nrhs);
void launch_ldlt_bwd_step(hipStream_t st, const LdltDev &d, const SolveItem *items, int64_t count, double *y,
                          int32_t n, int32_t nrhs);
// one iteration's pivot logic; with gpart set it first decides on the group
// the stats launch evaluated (nch workgroups per front)
void launch_big_panel(hipStream_t st, const LdltDev &d, BigLd *states, int32_t nbig, double u, double eff_small,
                      double small_pivot, int32_t nb, const double *gpart, int32_t nch);
// speculative group: column maxima (stats) and the accepted panels' rows (apply)
void launch_big_group_stats(hipStream_t st, const LdltDev &d, const BigLd *states, int32_t nbig, double *gpart,
                            int32_t nch);
void launch_big_group_apply(hipStream_t st, const LdltDev &d, const BigLd *states, int32_t nbig, double *gpart,
                            int32_t nch);
// nmode[m]: fronts of the level whose updates run in precision code m
// (0 FP64, 1 FP32, 2 BF16, 3 split BF16); a launch per code in use
void launch_big_update(hipStream_t st, const LdltDev &d, const BigLd *states, int32_t nbig, int32_t per_front,
                       const int32_t *nmode);
void launch_big_stall(hipStream_t st, const LdltDev &d, BigLd *states, int32_t nbig, int32_t nb);
// final contribution-block update with pivots [c0, c0 + width) of every
// front (L*D of the chunk into W first); then the fronts' results
void launch_big_cb(hipStream_t st, const LdltDev &d, const BigLd *states, int32_t nbig, int32_t c0, int32_t width,
                   int32_t per_front, const int32_t *nmode, bool wide);
void launch_big_finish(hipStream_t st, const LdltDev &d, const BigLd *states, int32_t nbig);

}  // namespace smfgpu
