/*
 * smfgpu: an MI355X (gfx950) numeric phase for the smf sparse symmetric
 * multifrontal solver (danlil240/opt_solver, smf 0.2.3).
 *
 * The host keeps smf's symbolic analysis; this library takes its result
 * (the AnalysisKeep fields a numeric phase reads) and runs the numeric
 * phase on one GPU: value mapping, extend-add assembly, dense Cholesky /
 * threshold LDL^T of every front, the Schur-complement update on MFMA,
 * forward / diagonal / backward solves and FP64 iterative refinement.
 *
 * Plain C ABI: plain pointers and sizes, 0-based indices, column-major dense
 * arrays, no exceptions cross it. Every call returns SMFGPU_OK (0) or a
 * negative error code (smfgpu_error_string gives the message of the last
 * failure on the calling thread).
 *
 * Entry points and the reference interface each one stands behind
 * (reference paths relative to the reference repository root):
 *
 *   smfgpu_plan_create      replaces the per-analysis set-up of
 *                           multifrontal_factor_serial
 *                           (solver/src/multifrontal_serial.cpp:1105-1320),
 *                           reading AnalysisKeep (solver/include/smf/analysis.hpp:25-144)
 *   smfgpu_factor           replaces factor_posdef / factor_indef's numeric
 *                           work: map_values_to_permuted
 *                           (solver/src/workspace.cpp:129-271), eff_small
 *                           (solver/src/factor_indef.cpp:633-636) and
 *                           multifrontal_factor_serial
 *                           (solver/include/smf/multifrontal_serial.hpp:98-105)
 *   smfgpu_export_factor    the FactorKeep layout
 *                           (solver/include/smf/factor_posdef.hpp:16-220)
 *   smfgpu_solve            solve_core (solver/src/solve.cpp:858-1075)
 *   smfgpu_refine_ex,       refine / refine_core / classical_refine /
 *   smfgpu_refine           fgmres_refine / compute_residual_and_backward_error
 *                           (solver/src/refinement.cpp:39-130, 300-724)
 */
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#define SMFGPU_API __attribute__((visibility("default")))
#else
#define SMFGPU_API
#endif

/* Return codes. */
#define SMFGPU_OK 0
#define SMFGPU_ERR_INVALID_ARG (-1)
#define SMFGPU_ERR_ALLOC (-2)   /* device or host allocation failed (smf ErrorCode::AllocationError) */
#define SMFGPU_ERR_VALUE (-3)   /* non-finite input value (smf ErrorCode::IllegalValue) */
#define SMFGPU_ERR_DEVICE (-4)  /* HIP runtime error */
#define SMFGPU_ERR_STATE (-5)   /* call not valid for this handle (e.g. solve of a failed factor) */
#define SMFGPU_ERR_UNSUPPORTED (-6)

/* Factor status, the values of smf::FactorStatus (solver/include/smf/types.hpp:230-238). */
#define SMFGPU_STATUS_SUCCESS 0
#define SMFGPU_STATUS_NOT_POSDEF 1
#define SMFGPU_STATUS_SINGULAR 2
#define SMFGPU_STATUS_MAX_PIVOT_DELAYS 3
#define SMFGPU_STATUS_INERTIA_EXCEEDED 4

/* Precision of the Schur-complement updates of large fronts (fronts of at
 * least mixed_min_front rows). Pivot tests, D, inertia, residuals and the
 * refined solution are always FP64. */
#define SMFGPU_PREC_FP64 0
#define SMFGPU_PREC_FP32 1   /* FP32 MFMA, FP32 accumulation */
#define SMFGPU_PREC_BF16 2   /* BF16 MFMA (operands rounded to bf16), FP32 accumulation */
#define SMFGPU_PREC_AUTO 3   /* per-supernode choice by the learned selector (weights from
                                smfgpu_selector_set; fronts of at least mixed_min_front rows) */
#define SMFGPU_PREC_BF16X3 4 /* split BF16: operands as hi + lo bf16 pairs, four BF16
                                MFMAs per product (lo*lo + lo*hi + hi*lo + hi*hi), FP32
                                accumulation */

/* Memory kind of value / right-hand-side pointers. */
#define SMFGPU_MEM_HOST 0
#define SMFGPU_MEM_DEVICE 1

typedef struct smfgpu_plan_s *smfgpu_plan_t;
typedef struct smfgpu_factor_s *smfgpu_factor_t;
typedef struct smfgpu_analysis_s *smfgpu_analysis_t;

/* The AnalysisKeep fields the numeric phase reads. Supernodes are numbered
 * in postorder (children before parents); children of s are
 * child_idx[child_ptr[s] .. child_ptr[s+1]) in descending order;
 * front_row_idx[front_row_ptr[s] ..) are the rows of front s, sorted, the
 * first sn_end[s] - sn_start[s] of them its pivot columns. pcol_ptr /
 * prow_idx is the permuted lower pattern the values are scattered into,
 * orig_to_slot maps each entry of the analysed matrix to its slot there
 * (-1: dropped; several entries may share a slot and are summed). */
typedef struct {
    int32_t n;
    int32_t nsup;
    const int32_t *sn_start;
    const int32_t *sn_end;
    const int32_t *sn_parent;
    const int32_t *child_ptr;
    const int32_t *child_idx;
    const int32_t *front_row_ptr;
    const int32_t *front_row_idx;
    const int32_t *pcol_ptr;
    const int32_t *prow_idx;
    int64_t n_orig;
    const int32_t *orig_to_slot;
    const int32_t *perm; /* perm[new] = old */
} smfgpu_analysis_desc;

/* Numeric options (the smf::Control fields the numeric phase reads,
 * solver/include/smf/control.hpp:331-440). */
typedef struct {
    int32_t posdef;          /* 1: Cholesky (MatrixType 3), 0: LDL^T (MatrixType 4) */
    double pivot_u;          /* threshold u (default 0.01) */
    double small_pivot;      /* zero-pivot threshold (default 1e-20) */
    int32_t panel;           /* indefinite panel width nb (default 16) */
    int32_t precision;       /* SMFGPU_PREC_* for the Schur updates */
    int32_t mixed_min_front; /* fronts with fewer rows stay FP64 (default 1024) */
    int32_t max_negative_pivots; /* < 0: off */
    int32_t value_mem;       /* SMFGPU_MEM_* of the values pointer */
    int32_t work_limit_on;   /* 1: stop LDL^T once front work would pass work_limit */
    double work_limit;       /* sum of num_elig * f^2 allowed (the scale-on-delay
                                threshold, multifrontal_serial.cpp:653-671);
                                a stopped run reports SMFGPU_STATUS_MAX_PIVOT_DELAYS */
    /* Static pivoting (LDL^T; control.hpp:333-364, factor_indef.cpp:642-645):
     * static_pivot_tol > 0: a column with no acceptable 1x1/2x2 pivot is
     * taken as a 1x1 pivot, |d| raised to static_pivot_tol * max|A|
     * (force_static_pivot, ldlt_kernel.cpp:285-320); no delays.
     * fallback_static_pivot_tol > 0: the same only in roots and in fronts
     * that receive more delayed columns than their own width, with hybrid
     * inertia (multifrontal_serial.cpp:562, 615-629). */
    double static_pivot_tol;
    double fallback_static_pivot_tol;
    int32_t static_flags;    /* SMFGPU_STATIC_* bits */
    /* Learned selector (precision SMFGPU_PREC_AUTO): each front of at least
     * mixed_min_front rows takes the cheapest precision in selector_mask
     * (bits 1 << SMFGPU_PREC_FP32 / _BF16 / _BF16X3; 0: all three) whose
     * predicted relative error of the Schur update is <= selector_tol, else
     * FP64. The prediction runs on the device between the front's assembly
     * and its factorization. */
    int32_t selector_mask;
    double selector_tol;     /* default 1e-4 */
    int32_t scaled;          /* 1: the values are scaled (the selector's scaling-state feature) */
    int32_t collect;         /* 1 (FP64 only): keep the features of the fronts of at least
                                mixed_min_front rows for smfgpu_selector_pairs */
    int32_t scale_mem;       /* SMFGPU_MEM_* of scale */
    int32_t reserved;
    /* Symmetric scaling (ScalingMethod::UserSupplied, control.hpp:148; the
     * S A S pass of factor_indef.cpp:508-559 and factor_posdef.cpp:405-421):
     * null, or n finite positive doubles indexed by the original (analysed
     * matrix's) row. The factorization then factors S A S, each entry
     * a_rc * (s_r * s_c), with max|S A S| in place of max|A| for the pivot
     * thresholds; its solves return x = S (S A S)^-1 S b (solve.cpp:1022-1067)
     * and its refinement computes residuals with the unscaled A. An invalid
     * entry (non-finite or <= 0) fails with SMFGPU_ERR_INVALID_ARG
     * (validate_positive_scale, scaling.cpp:9-19). Sets the selector's
     * scaling-state feature. smfgpu_equilibration_scale computes the
     * reference's Equilibration scaling on the device. */
    const double *scale;
} smfgpu_factor_options;

/* static_flags (control.hpp:349-364) */
#define SMFGPU_STATIC_AS_ZERO 1         /* static_pivots_as_zero */
#define SMFGPU_STATIC_ZERO_STAYS_ZERO 2 /* static_zero_stays_zero */
#define SMFGPU_STATIC_IN_ORDER 4        /* static_pivot_in_order (no search) */

/* Factor statistics (the smf::Info fields of a factorization,
 * solver/include/smf/info.hpp:24-221). */
typedef struct {
    int32_t status;          /* SMFGPU_STATUS_* */
    int32_t num_positive;
    int32_t num_negative;
    int32_t num_zero;
    int32_t delayed_pivots;
    int32_t two_by_two_pivots;
    int32_t max_front;
    int32_t fronts_reduced;  /* fronts whose Schur update ran below FP64 */
    double actual_flops;     /* kernel flops (reference count) */
    double front_work;       /* sum num_elig * f^2 */
    double max_abs;          /* max |a_ij| of the input */
    double factor_seconds;   /* device time of the numeric factorization */
    int64_t factor_entries;  /* sum f * p */
    int32_t static_pivots;   /* pivots whose |d| static pivoting raised */
    int32_t pad;
    double pivot_growth;     /* largest column max / |d| over static pivots */
    double reserved[2];
} smfgpu_factor_info;

SMFGPU_API const char *smfgpu_error_string(void);
SMFGPU_API int smfgpu_default_options(smfgpu_factor_options *opt);

/* ---- symbolic analysis on the host (stands in for smf's Solver::analyse,
 * solver/src/symbolic_analysis.cpp:434, when no smf analysis is at hand).
 * perm (new -> old) may be null for the identity; nemin: relaxed
 * amalgamation width (default 32 when <= 0). */
SMFGPU_API int smfgpu_analyse(int32_t n, const int32_t *col_ptr, const int32_t *row_idx,
                              const int32_t *perm, int32_t nemin, smfgpu_analysis_t *out);
SMFGPU_API int smfgpu_analysis_describe(smfgpu_analysis_t an, smfgpu_analysis_desc *desc);
/* out[0..5]: nsup, sum of front rows, factor entries (sum f p), max front,
 * delay-free kernel flops, depth (levels). */
SMFGPU_API int smfgpu_analysis_stats(smfgpu_analysis_t an, double *out);
SMFGPU_API void smfgpu_analysis_free(smfgpu_analysis_t *an);

/* ---- device plan (one per analysis; owns the device buffers) ---- */
SMFGPU_API int smfgpu_plan_create(const smfgpu_analysis_desc *desc, int32_t posdef,
                                  smfgpu_plan_t *out);
SMFGPU_API void smfgpu_plan_free(smfgpu_plan_t *plan);
/* out[0..7]: device bytes held, levels, kernel launches per factorization,
 * delay-free flops, factor entries, contribution-block entries, assembly
 * bytes (algorithmic), solve bytes per RHS (algorithmic). */
SMFGPU_API int smfgpu_plan_stats(smfgpu_plan_t plan, double *out);

/* Equilibration scaling on the device (ScalingMethod::Equilibration,
 * compute_equilibration_scale, solver/src/scaling.cpp:57-101): from s = 1,
 * iters passes (the reference runs 3) of s_i /= sqrt(max_j |a_ij| s_i s_j)
 * over the symmetric matrix of the analysed pattern, every row's maximum
 * taken with the previous pass's s; rows without a nonzero keep s_i. values:
 * one per entry (memory kind value_mem, as for smfgpu_factor); scale
 * receives the n factors in the original ordering (memory kind scale_mem),
 * ready for smfgpu_factor_options.scale. The same s as the reference's to
 * the bit. Non-finite values fail with SMFGPU_ERR_VALUE. */
SMFGPU_API int smfgpu_equilibration_scale(smfgpu_plan_t plan, const double *values, int64_t nvalues,
                                          int32_t value_mem, int32_t iters, double *scale, int32_t scale_mem);

/* ---- numeric factorization. values: one per entry of the analysed
 * matrix (host or device memory per opt->value_mem). The returned factor
 * stays on the device; a plan has one live factor (a new factorization
 * invalidates the previous factor handle's contents). */
SMFGPU_API int smfgpu_factor(smfgpu_plan_t plan, const double *values, int64_t nvalues,
                             const smfgpu_factor_options *opt, smfgpu_factor_t *out,
                             smfgpu_factor_info *info);
SMFGPU_API void smfgpu_factor_free(smfgpu_factor_t *f);

/* FactorKeep layout on the host: steps (4 int64 per supernode step:
 * row_offset, factor_offset, p, f), factor values (sum f p doubles),
 * row indices (sum f), pivot types (sum p int8; SPD: all 1). sizes[0..3]
 * receive the four lengths when the pointers are null. */
SMFGPU_API int smfgpu_export_factor(smfgpu_factor_t f, int64_t *sizes, int64_t *steps,
                                    double *factor_values, int32_t *row_indices,
                                    int8_t *pivot_types);

/* Per-supernode results of the last factorization, 8 int32 per supernode:
 * front rows f (with delays), accepted pivots p, positive / negative / zero
 * pivots, 2x2 pivots, columns delayed to the parent, eligible columns
 * (process_front's per-front Info terms, multifrontal_serial.cpp:864-1031).
 * Lets a caller that factors many independent systems as one forest report
 * each system's inertia and delays. */
SMFGPU_API int smfgpu_factor_front_stats(smfgpu_factor_t f, int32_t *out);
/* Blocked-path counters of the last factorization, 4 int32 per supernode:
 * first-pass panels that stalled, partner pairs pulled in front of a
 * stalled panel, failed columns rotated to the end of the live range, and
 * second-pass full searches (the stall handling of ldlt_factor_front's
 * blocked path, ldlt_kernel.cpp:1028-1131). No counterpart in the
 * reference's Info: a diagnostic of how often a front took those branches.
 * Zero for SPD plans and for fronts below the blocked size. */
SMFGPU_API int smfgpu_factor_stall_stats(smfgpu_factor_t f, int32_t *out);
/* Precision code of each supernode's Schur updates in the last
 * factorization (one int8 per supernode: SMFGPU_PREC_FP64 / _FP32 / _BF16 /
 * _BF16X3), as the static policy or the learned selector set them. */
SMFGPU_API int smfgpu_factor_precisions(smfgpu_factor_t f, int8_t *out);

/* ---- learned per-supernode precision selector. An MLP (10 features ->
 * 32 -> 32 -> 3, tanh) predicts, from cheap features of an assembled front
 * (front order, pivot columns, contribution-block share, spread of the
 * pivot-column norms, diagonal dominance (worst and mean), spread of the
 * diagonal, column norms against max|A|, delayed columns received, scaling
 * state), the log10 relative error max|dZ| / max|S| of its Schur update in
 * FP32, split BF16 and BF16. Training and inference are HIP kernels; no
 * reference counterpart (the reference factors in FP64 only). */
#define SMFGPU_SELECTOR_FEATURES 10
#define SMFGPU_SELECTOR_OUTPUTS 3
#define SMFGPU_SELECTOR_WEIGHTS 1527 /* W1 b1 W2 b2 W3 b3, input mean, input 1/std (floats) */
SMFGPU_API int smfgpu_selector_set(smfgpu_plan_t plan, const float *weights);
/* Training pairs of an FP64 factorization made with opt.collect = 1: for
 * each front of at least mixed_min_front rows with a contribution block,
 * its features (X, SMFGPU_SELECTOR_FEATURES floats) and the measured log10
 * max|Z_m - Z_64| / max|S| of its contribution-block update Z = L21 L21^T
 * run on the MFMA tiles in FP32, split BF16 and BF16 from the FP64 factor
 * (Y, 3 floats; S the FP64 contribution block), and its supernode. Returns
 * the number of pairs (at most cap; with null arrays, the number there
 * are), or a negative error code. Cholesky factors: the updates recomputed
 * from the stored factor; LDL^T factors: measured during the factorization
 * on the large fronts (the multi-workgroup path) by running each one's final
 * contribution-block update in every precision from the same square. */
SMFGPU_API int smfgpu_selector_pairs(smfgpu_factor_t f, float *X, float *Y, int32_t *fronts, int32_t cap);
/* Adam training of the MLP on n pairs on the device (kernels of the
 * library): minimises the squared error over the labels whose mask entry
 * is 1 (M null: all). weights holds the start point and receives the
 * result (its input mean / 1/std are set from X first); loss, when not
 * null, receives the mean loss of each epoch (epochs floats). */
SMFGPU_API int smfgpu_selector_train(const float *X, const float *Y, const float *M, int32_t n, int32_t epochs,
                                     float lr, float *weights, float *loss);
/* Inference kernel over n feature rows: Y (n x 3) predicted log10 errors. */
SMFGPU_API int smfgpu_selector_predict(const float *weights, const float *X, int32_t n, float *Y);

/* ---- solves: b is n x nrhs, leading dimension ldb >= n, in place.
 * job as smf::SolveJob (solver/include/smf/types.hpp:250-256; solve_core,
 * solver/src/solve.cpp:1050-1068):
 *   0 full       A x = b, b and x in the original ordering;
 *   1 forward    x = L^-1 P^T b: b original, x in the permuted ordering;
 *   2 diagonal   x = D^-1 b, both permuted (the reference's solve_diag:
 *                division by d, scaled 2x2 solve; a Cholesky factor copies);
 *   3 backward   x = P L^-T b: b permuted, x original;
 *   4 diag+back  x = P L^-T D^-1 b: b permuted, x original.
 * mem: SMFGPU_MEM_*. */
SMFGPU_API int smfgpu_solve(smfgpu_factor_t f, double *b, int32_t nrhs, int32_t ldb,
                            int32_t job, int32_t mem);

/* smfgpu_solve with b in device memory that allocates nothing: no
 * counterpart in the reference. It is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The right-hand-side work block is a
 * fixed 64 MiB device array per GPU (resolved when a plan is created on that
 * device, held until the process ends) instead of a plan buffer allocated on
 * first use; right-hand sides pass through the sweeps in groups of at most
 * min(4 (Cholesky) / 16 (LDL^T), 2^23 / n) columns. Plans with n > 2^23
 * rows are refused with SMFGPU_ERR_INVALID_ARG. Every call on one GPU
 * shares that array: calls that overlap on one GPU (from several threads or
 * plans) race. Returns after the solve has finished. */
SMFGPU_API int smfgpu_solve_noalloc(smfgpu_factor_t f, double *b, int32_t nrhs, int32_t ldb,
                                    int32_t job);

/* ---- FP64 iterative refinement against the analysed matrix (refine /
 * refine_core, solver/src/refinement.cpp:300-724). values: null for the
 * factored values, else values of the analysed pattern (one per entry, as
 * given to smfgpu_factor; memory kind value_mem) to refine against; b and x
 * are n x nrhs (leading dimensions ldb, ldx >= n; device or host per mem);
 * x is refined in place and left at the best iterate. Residuals use
 * double-double products and sums; eta_inf = ||b - A x||_inf /
 * (||A||_inf ||x||_inf + ||b||_inf), omega = max_i |b - A x|_i /
 * (|A||x| + |b|)_i (refinement.cpp:39-130), over all columns. */
#define SMFGPU_REFINE_CLASSICAL 0 /* x += M^-1 r; stop on no improvement or two steps contracting < 0.85 */
#define SMFGPU_REFINE_AUTO 1      /* classical, escalating to FGMRES on a step contracting < 0.5 */
#define SMFGPU_REFINE_FGMRES 2    /* right-preconditioned FGMRES(restart), one column at a time */
typedef struct {
    int32_t method;        /* SMFGPU_REFINE_* (Control::refine_method, default AUTO) */
    int32_t max_iters;     /* classical steps (Control::refinement_max_iters, default 5) */
    int32_t max_steps;     /* AUTO / FGMRES: solves in total (Control::refine_max_steps, default 10) */
    int32_t restart;       /* FGMRES restart length (Control::fgmres_restart, default 20) */
    double eta_target;     /* stop once eta_inf <= eta_target ... (default 1e-10) */
    double omega_target;   /* ... and omega <= omega_target (default 1e-8) */
    int32_t value_mem;     /* SMFGPU_MEM_* of values (when not null) */
    int32_t reserved[3];
} smfgpu_refine_options;
typedef struct {
    double eta, omega;             /* of the returned x */
    double eta_start, omega_start; /* of the given x */
    int32_t steps;                 /* solves taken */
    int32_t method_used;           /* SMFGPU_REFINE_CLASSICAL or SMFGPU_REFINE_FGMRES */
    int32_t reached;               /* 1: both targets met */
    int32_t stalled;               /* 1: stopped with a target unmet */
    int32_t fallback;              /* smfgpu_factor_solve: 1 when the reduced-precision factor's
                                      refinement stalled and the plain FP64 factor was used */
    int32_t pad;
    double reserved[1];
} smfgpu_refine_info;
SMFGPU_API int smfgpu_refine_default_options(smfgpu_refine_options *opt);
SMFGPU_API int smfgpu_refine_ex(smfgpu_factor_t f, const double *values, const double *b, int32_t ldb, double *x,
                                int32_t ldx, int32_t nrhs, const smfgpu_refine_options *opt, int32_t mem,
                                smfgpu_refine_info *info);
/* Shorthand: AUTO with max_iters classical steps and at most max_iters
 * solves in total, eta_target = tol, no omega target; ldb = ldx = n.
 * out[0] eta_inf, out[1] omega, out[2] steps taken. */
SMFGPU_API int smfgpu_refine(smfgpu_factor_t f, const double *values, const double *b,
                             double *x, int32_t nrhs, int32_t max_iters, double tol,
                             int32_t mem, double *out);

/* Factor, solve and refine in one call (Solver::factor_solve,
 * solver/src/solver.cpp:14-73, on an existing analysis): x = A^-1 b, b and
 * x n x nrhs with leading dimension ld (memory kind mem). When the Schur
 * updates ran below FP64 (opt->precision FP32 / BF16 / split BF16 / AUTO)
 * and refinement stops with a target unmet, the plan is refactored with
 * every front in FP64 and the solve and refinement are rerun: the plain FP64
 * factor judges the error (rinfo->fallback = 1). *out receives the factor
 * that produced x. */
SMFGPU_API int smfgpu_factor_solve(smfgpu_plan_t plan, const double *values, int64_t nvalues,
                                   const smfgpu_factor_options *opt, const double *b, double *x, int32_t nrhs,
                                   int32_t ld, const smfgpu_refine_options *ropt, int32_t mem,
                                   smfgpu_factor_t *out, smfgpu_factor_info *finfo, smfgpu_refine_info *rinfo);

/* Per-plan tuning knobs (the autotuner's parameters). Keys:
 * SMFGPU_TUNE_SOLVE_FUSE_ROWS: an SPD front whose last pivot block has at
 * most this many rows below it has them updated by the workgroup that solves
 * the block, saving a step of the sweep (default 512). */
#define SMFGPU_TUNE_SOLVE_FUSE_ROWS 1
/* SMFGPU_TUNE_SYRK_AHEAD: panel chunks (16 columns each) whose global loads
 * the SPD Schur-update kernel keeps in flight, 1 or 2 (default 2). */
#define SMFGPU_TUNE_SYRK_AHEAD 2
/* SMFGPU_TUNE_PANEL_OUTER: SPD pivot columns per trailing Schur update
 * (a multiple of 64, default 512); rebuilds the launch sequence. */
#define SMFGPU_TUNE_PANEL_OUTER 3
/* SMFGPU_TUNE_CB_SEG: SPD pivot columns per update of a contribution block,
 * which is updated once its front's pivots are all factored (default 4096;
 * FP32 Schur updates accumulate in FP32 over one segment). */
#define SMFGPU_TUNE_CB_SEG 4
/* SMFGPU_TUNE_LDLT_BIG_MIN: LDL^T fronts with at least this many rows
 * (default 256, minimum 96) run the blocked algorithm over many
 * workgroups (speculative panel groups, pivot logic per panel in one
 * workgroup, MFMA updates); smaller blocked fronts take one workgroup. */
#define SMFGPU_TUNE_LDLT_BIG_MIN 5
/* SMFGPU_TUNE_LDLT_GROUP: 1 (default) lets the large-front LDL^T path try
 * groups of four first-pass panels at once (accepted when every pivot
 * passes the first threshold test, otherwise the panel is redone with the
 * full search; the same pivots either way), 0 factors one panel at a time.
 * Groups need the panel width 16. */
#define SMFGPU_TUNE_LDLT_GROUP 6
/* SMFGPU_TUNE_LDLT_UPDATE_WG: workgroups of the large-front MFMA update
 * launches, shared by the level's large fronts (a multiple of 8, default
 * 1024; at least 64 per front). */
#define SMFGPU_TUNE_LDLT_UPDATE_WG 7
/* SMFGPU_TUNE_LDLT_DEVICE_FRONTS: LDL^T levels whose symbolic fronts all
 * have at most min(64, value) rows are laid out and factored on the device
 * (one workgroup per front, no host round trip per level); a level with a
 * front that delays grow past value rows goes to the host layout. Default
 * 95 (the largest unblocked front), 0 turns it off; at most 95. */
#define SMFGPU_TUNE_LDLT_DEVICE_FRONTS 8
/* SMFGPU_TUNE_ASM_CHUNK: SPD assembly's LDS column chunk is the level's
 * longest column up to this many rows (default 2048, in [256, 16384]):
 * longer chunks pass each child column once, shorter ones let more
 * workgroups share a CU. */
#define SMFGPU_TUNE_ASM_CHUNK 9
/* SMFGPU_TUNE_LDLT_DEVICE_WAVE: 1 (default) factors the device levels'
 * fronts of at most 48 rows with one wavefront each (many fronts per CU),
 * 0 gives every front of a device level a 256-thread workgroup. */
#define SMFGPU_TUNE_LDLT_DEVICE_WAVE 10
/* SMFGPU_TUNE_SOLVE_PERSISTENT: 1 runs each SPD sweep as one launch, a
 * workgroup per item taking its item by ticket in step order and waiting
 * on a per-step counter; 0 (default) launches every step separately. On
 * the 100^3 27-point operator the single launch measured 19.5 ms per sweep
 * against 6.8 ms: each item's device-scope release and acquire write back
 * and invalidate its XCD's L2. */
#define SMFGPU_TUNE_SOLVE_PERSISTENT 11
/* SMFGPU_TUNE_PANEL_OUTER_REDUCED: SMFGPU_TUNE_PANEL_OUTER for the
 * factorizations whose fronts may run below FP64 (default 256: their
 * trailing updates accumulate that many products in FP32). */
#define SMFGPU_TUNE_PANEL_OUTER_REDUCED 12
/* SMFGPU_TUNE_LDLT_PLANES: 1 (default) gives the large LDL^T fronts of BF16
 * and split-BF16 precision their outer-block updates (the held-back pivots'
 * flush and the contribution block's chunks) on bf16 copies of the
 * operands, written once per update, in 256 x 256 tiles; 0 rounds the FP64
 * operands inside every 128 x 128 tile. With the same outer block (512:
 * SMFGPU_TUNE_LDLT_OUTER_WIDE) both round and sum the same way and the
 * factors agree bit for bit. */
#define SMFGPU_TUNE_LDLT_PLANES 13
/* SMFGPU_TUNE_LDLT_OUTER_WIDE: outer block (and contribution-block chunk)
 * of the large LDL^T fronts updated on bf16 planes, a multiple of 64 in
 * [64, 1024] (default 1024; the other fronts use 512). Each rank-w update
 * reads and writes the FP64 entries once, and its FP32 accumulators sum w
 * products. */
#define SMFGPU_TUNE_LDLT_OUTER_WIDE 14
/* SMFGPU_TUNE_ASM_WAVE_ROWS: SPD assembly levels whose fronts all have at
 * most this many rows (in [0, 4096], default 2048; 0: none) assemble one
 * column per wavefront, four per workgroup, instead of one column per
 * workgroup: the bottom levels' many short columns are otherwise bound by
 * workgroup dispatch. The same sums in the same order either way. */
#define SMFGPU_TUNE_ASM_WAVE_ROWS 15
SMFGPU_API int smfgpu_plan_tune(smfgpu_plan_t plan, int32_t key, int64_t value);

/* Device time (ms) of the last factorization's dominant kernel family:
 * out[0] Schur-update kernel total ms, out[1] launches, out[2] flops in it,
 * out[3] assembly ms, out[4] assembly bytes, out[5] solve ms (last solve),
 * out[6] solve bytes (last solve), out[7] forward sweep ms, out[8] backward
 * sweep ms (SPD, last right-hand-side group), out[9] forward launches,
 * out[10] backward launches. Measured with HIP events on the plan's
 * stream when timing is enabled. out[11]: levels of the last LDL^T
 * factorization laid out on the device (counted with timing off too);
 * out[12..15] host wall ms of the last LDL^T factorization's phases: the
 * device levels (launches and wait), their read-back, the statistics, the
 * solve schedule. out holds 16 doubles. */
SMFGPU_API int smfgpu_set_timing(smfgpu_plan_t plan, int32_t enabled);
SMFGPU_API int smfgpu_kernel_times(smfgpu_plan_t plan, double *out);
/* SPD launch sequence of the last timed factorization: 5 doubles per launch
 * (kind 0 assembly / 1 Cholesky / 2 TRSM / 3 Schur, level, items, Schur
 * flops, ms or -1 when not timed), at most cap launches; returns the
 * number of launches (out may be null). */
SMFGPU_API int64_t smfgpu_launch_profile(smfgpu_plan_t plan, double *out, int64_t cap);

#ifdef __cplusplus
}
#endif
