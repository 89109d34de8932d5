// Device data of the indefinite (LDL^T) numeric phase. Front sizes grow with
// the delayed pivots children pass up, so they are known only when a
// level starts: the host lays the level out (offsets into the arenas
// below) after reading back its children's results.
#pragma once

#include "smfgpu_internal.hpp"

namespace smfgpu {

// Front s of the current factorization. Its f x f square (column-major,
// leading dimension f, lower triangle) lives in the arena at fofs and stays
// there: columns [0, pstore) are the stored factor (FactorKeep layout:
// unit L, D on the diagonal, a 2x2 block's off-diagonal in slot (c, c+1)),
// the trailing (f - p) square is the contribution block the parent
// assembles, delayed rows first.
struct FrontLd {
    int32_t f, ne, nd, psym;  // rows, eligible columns, delayed columns received, own pivots
    int32_t col0;             // first own pivot column (permuted numbering)
    int32_t root;             // 1: no parent (unaccepted columns become zero pivots)
    int32_t p, pstore;        // accepted pivots; stored pivot columns
    int32_t npos, nneg, nzero, n2x2;
    int32_t delayed, pad;
    double flops;
    int64_t fofs;             // arena offset of the square
    int64_t rofs;             // offset of the row ids / tags
    int64_t wofs;             // offset of the panel workspace W (blocked fronts)
    int64_t dofs;             // offset of the pivots' D^-1 entries
    int64_t cmofs;            // offset of this front's CB-row map into its parent
};

struct LdltDev {
    FrontLd *fronts;          // per supernode
    double *arena;
    int32_t *rows;            // per front: f row ids (permuted numbering), pivoted order
    int8_t *tags;             // per front: pivot tags (1, 2, -1, 0), same offsets as rows
    double *work;
    double *dinv;             // 3 per pivot
    int8_t *dkind;            // 1, 2, 0 per pivot
    const int32_t *cmap;      // CB-row maps
    const int32_t *pcol_ptr;
    const int32_t *amap;      // per permuted entry: position among its front's symbolic rows
    const double *av;
};

void launch_ldlt_small(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double u, double eff_small,
                       double small_pivot);
void launch_ldlt_blocked(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double u,
                         double eff_small, double small_pivot, int32_t nb);
void launch_ldlt_assemble(hipStream_t st, const LdltDev &d, const AsmTask *tasks, int64_t ntasks, const ChildCol *cc);
void launch_ldlt_epilogue(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf);
void launch_ldlt_solve_level(hipStream_t st, const LdltDev &d, const int32_t *flist, int32_t nf, double *y, int32_t n,
                             int32_t nrhs, int32_t phase);
void launch_front_patch(hipStream_t st, FrontLd *fronts, const FrontLd *patch, const int32_t *ids, int32_t np);

}  // namespace smfgpu
