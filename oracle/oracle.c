/*
 * ORACLE -- test infrastructure only. A plain-C restatement of smf's numeric
 * phase (danlil240/opt_solver, smf 0.2.3) used as the checker of the GPU
 * path: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * load it. Nothing in the product links it.
 *
 * Restated from (paths relative to the reference repository):
 *   select_pivot, two_by_two_ok      solver/src/ldlt_kernel.cpp:97-275
 *   sym_swap                         solver/src/ldlt_kernel.cpp:333-345
 *   eliminate_1x1 / inverse_2x2 / eliminate_2x2  ldlt_kernel.cpp:350-545
 *   inertia_1x1, apply_pivot         solver/src/ldlt_kernel.cpp:655-698
 *   add_inertia_2x2, scaled_det_2x2  solver/src/dense_indef.cpp:52-88, pivoting.cpp:18-28
 *   deferred_update, schur_update    solver/src/ldlt_kernel.cpp:705-933 (plain loops
 *                                    in place of BLAS: blocked fronts agree to rounding)
 *   ldlt_factor_front                solver/src/ldlt_kernel.cpp:970-1139
 *   process_front (rows, assembly, store, CB push)  solver/src/multifrontal_serial.cpp:472-1035
 *   map_values_to_permuted           solver/src/workspace.cpp:129-271
 *   eff_small                        solver/src/factor_indef.cpp:633-636
 *   solve_core forward / D / backward  solver/src/solve.cpp:858-1075,
 *                                    solve_forward.cpp, solve_backward.cpp, solve_diag.cpp:49-123
 *   compute_residual_and_backward_error  solver/src/refinement.cpp:39-130
 *
 * Not restated: static / fallback / budget static pivoting, APTP and tiles,
 * the scale-on-delay policy (the parity configuration turns it off), and
 * the work / negative-pivot limits. Compile with -ffp-contract=off, as the
 * reference compiles ldlt_kernel.cpp: the unblocked kernel (fronts below
 * blocked_min_front rows) then reproduces the reference bit for bit.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define WEAK_PIVOT_RATIO 1e-6
#define BLOCKED_MIN_FRONT 96

typedef struct {
    double u, eff_small, small_pivot;
    int nb, posdef;
} orc_params;

/* stalls / pulls / rotated / searches: instrumentation of the blocked
   path with no counterpart in the reference's Info: first-pass panels that
   stalled, partner pairs pulled in front of a stalled panel, failed columns
   rotated to the end of the live range, and second-pass full searches. */
typedef struct {
    int p;
    int pos, neg, zero;
    double flops;
    int stalls, pulls, rotated, searches;
} orc_result;

static int is_fin(double v) { return isfinite(v); }

static double max3(double a, double b, double c) {
    double m = a;
    if (m < b) m = b;
    if (m < c) m = c;
    return m;
}

static double sym_at(const double *F, size_t ld, int i, int j) {
    return (i >= j) ? F[(size_t)j * ld + i] : F[(size_t)i * ld + j];
}

static int two_by_two_ok(const double *F, size_t ld, int f, int cand, int r, const double *col,
                         const double *col_r, double m2, double off1, int ext0, const orc_params *prm,
                         double *growth) {
    const double d00 = col[cand];
    const double d10 = sym_at(F, ld, r, cand);
    const double d11 = col_r[r];
    const double s2 = max3(fabs(d00), fabs(d10), fabs(d11));
    if (!(s2 >= prm->eff_small) || !is_fin(s2)) return 0;
    const double a = d00 / s2, b = d10 / s2, c = d11 / s2;
    const double det_s = a * c - b * b;
    if (!(fabs(det_s) > prm->small_pivot)) return 0;
    double off0 = m2;
    for (int i = ext0; i < f; ++i) off0 = fmax(off0, fabs(col[i]));
    double g = 0.0;
    if (off0 > 0.0 || off1 > 0.0) {
        const double bound0 = (fabs(c) * off0 + fabs(b) * off1) / s2;
        const double bound1 = (fabs(b) * off0 + fabs(a) * off1) / s2;
        const double bm = bound0 > bound1 ? bound0 : bound1;
        if (prm->u * bm > fabs(det_s)) return 0;
        g = bm / fabs(det_s);
    }
    *growth = g;
    return 1;
}

/* kind: 0 none, 1 1x1 at a, 2 2x2 at (a, b) */
static void select_pivot(const double *F, size_t ld, int f, int k, int num_elig, int limit,
                         const orc_params *prm, int *kind, int *pa, int *pb) {
    *kind = 0;
    *pa = -1;
    *pb = -1;
    if (prm->posdef) {
        const double d = F[(size_t)k * ld + k];
        if (d > 0.0 && d <= DBL_MAX) {
            *kind = 1;
            *pa = k;
        }
        return;
    }
    if (k < limit) {
        const double *col0 = F + (size_t)k * ld;
        const double a0 = fabs(col0[k]);
        if (a0 >= prm->eff_small) {
            double colmax = 0.0;
            for (int i = k + 1; i < f; ++i) {
                const double v = fabs(col0[i]);
                if (v > colmax) colmax = v;
            }
            if (!(prm->u * colmax > a0) && !(a0 < WEAK_PIVOT_RATIO * colmax)) {
                *kind = 1;
                *pa = k;
                return;
            }
        }
    }
    for (int cand = k; cand < limit; ++cand) {
        const double *col = F + (size_t)cand * ld;
        const double a_cand = fabs(col[cand]);
        double m1 = 0.0, m2 = 0.0;
        int r_best = -1;
        for (int i = k; i < cand; ++i) {
            const double v = fabs(F[(size_t)i * ld + cand]);
            if (v > m1) {
                m2 = m1;
                m1 = v;
                r_best = i;
            } else if (v > m2) {
                m2 = v;
            }
        }
        for (int i = cand + 1; i < num_elig; ++i) {
            const double v = fabs(col[i]);
            if (v > m1) {
                m2 = m1;
                m1 = v;
                r_best = i;
            } else if (v > m2) {
                m2 = v;
            }
        }
        const int ext0 = num_elig > cand + 1 ? num_elig : cand + 1;
        int weak_1x1 = 0;
        double weak_growth = 0.0;
        if (a_cand >= prm->eff_small && !(prm->u * m1 > a_cand)) {
            int ok = 1;
            double colmax = m1;
            for (int i = ext0; i < f; ++i) {
                const double v = fabs(col[i]);
                if (prm->u * v > a_cand) {
                    ok = 0;
                    break;
                }
                colmax = fmax(colmax, v);
            }
            if (ok) {
                if (!(a_cand < WEAK_PIVOT_RATIO * colmax)) {
                    *kind = 1;
                    *pa = cand;
                    return;
                }
                weak_1x1 = 1;
                weak_growth = colmax / a_cand;
            }
        }
        if (r_best < k || r_best >= limit || m1 <= 0.0) {
            if (weak_1x1) {
                *kind = 1;
                *pa = cand;
                return;
            }
            continue;
        }
        const double *col_r = F + (size_t)r_best * ld;
        double om2 = 0.0, off1 = 0.0;
        for (int i = k; i < r_best; ++i) {
            const double v = fabs(F[(size_t)i * ld + r_best]);
            om2 = fmax(om2, v);
            if (i != cand) off1 = fmax(off1, v);
        }
        for (int i = r_best + 1; i < f; ++i) {
            const double v = fabs(col_r[i]);
            om2 = fmax(om2, v);
            if (i != cand) off1 = fmax(off1, v);
        }
        double growth2 = 0.0;
        if (k + 1 < limit && two_by_two_ok(F, ld, f, cand, r_best, col, col_r, m2, off1, ext0, prm, &growth2) &&
            (!weak_1x1 || growth2 < weak_growth)) {
            *kind = 2;
            *pa = cand;
            *pb = r_best;
            return;
        }
        if (weak_1x1) {
            *kind = 1;
            *pa = cand;
            return;
        }
        const double a_r = fabs(col_r[r_best]);
        if (a_r >= prm->eff_small && (om2 == 0.0 || a_r >= prm->u * om2)) {
            *kind = 1;
            *pa = r_best;
            return;
        }
    }
}

#define SWAPD(x, y) do { double t_ = (x); (x) = (y); (y) = t_; } while (0)

static void sym_swap(double *F, size_t ld, int f, int p, int q, int32_t *rows, double *W, int w_cols) {
    if (p == q) return;
    if (p > q) {
        int t = p;
        p = q;
        q = t;
    }
    const size_t sf = ld, wf = (size_t)f;
    SWAPD(F[p * sf + p], F[q * sf + q]);
    for (int m = 0; m < p; ++m) SWAPD(F[m * sf + p], F[m * sf + q]);
    for (int m = p + 1; m < q; ++m) SWAPD(F[p * sf + m], F[m * sf + q]);
    for (int m = q + 1; m < f; ++m) SWAPD(F[p * sf + m], F[q * sf + m]);
    int32_t tr = rows[p];
    rows[p] = rows[q];
    rows[q] = tr;
    for (int c = 0; c < w_cols; ++c) SWAPD(W[c * wf + p], W[c * wf + q]);
}

static void eliminate_1x1(double *F, int f, int k, int ucol_end, double *W, int w_col) {
    const size_t sf = (size_t)f;
    double *ck = F + k * sf;
    const double inv_d = 1.0 / ck[k];
    if (W && k + 1 < f) memcpy(W + (size_t)w_col * sf + k + 1, ck + k + 1, (size_t)(f - k - 1) * sizeof(double));
    for (int j = k + 1; j < ucol_end; ++j) {
        const double lj = ck[j] * inv_d;
        if (lj == 0.0) continue;
        double *cj = F + (size_t)j * sf;
        for (int i = j; i < f; ++i) cj[i] -= ck[i] * lj;
    }
    for (int i = k + 1; i < f; ++i) ck[i] *= inv_d;
}

static void inverse_2x2(double d00, double d10, double d11, double *i00, double *i10, double *i11) {
    const double s = max3(fabs(d00), fabs(d10), fabs(d11));
    const double a = d00 / s, b = d10 / s, c = d11 / s;
    const double inv = 1.0 / ((a * c - b * b) * s);
    *i00 = c * inv;
    *i10 = -b * inv;
    *i11 = a * inv;
}

static void eliminate_2x2(double *F, int f, int k, int ucol_end, double *W, int w_col) {
    const size_t sf = (size_t)f;
    double *c0 = F + k * sf, *c1 = F + (k + 1) * sf;
    double i00, i10, i11;
    inverse_2x2(c0[k], c0[k + 1], c1[k + 1], &i00, &i10, &i11);
    if (W && k + 2 < f) {
        memcpy(W + (size_t)w_col * sf + k + 2, c0 + k + 2, (size_t)(f - k - 2) * sizeof(double));
        memcpy(W + (size_t)(w_col + 1) * sf + k + 2, c1 + k + 2, (size_t)(f - k - 2) * sizeof(double));
    }
    for (int j = k + 2; j < ucol_end; ++j) {
        const double a0 = c0[j], a1 = c1[j];
        const double l0 = i00 * a0 + i10 * a1;
        const double l1 = i10 * a0 + i11 * a1;
        if (l0 == 0.0 && l1 == 0.0) continue;
        double *d0 = F + (size_t)j * sf;
        for (int i = j; i < f; ++i) d0[i] -= c0[i] * l0 + c1[i] * l1;
    }
    for (int i = k + 2; i < f; ++i) {
        const double a0 = c0[i], a1 = c1[i];
        c0[i] = i00 * a0 + i10 * a1;
        c1[i] = i10 * a0 + i11 * a1;
    }
}

static void inertia_1x1(double d, double small, orc_result *r) {
    if (!is_fin(d) || fabs(d) < small) ++r->zero;
    else if (d > 0.0) ++r->pos;
    else ++r->neg;
}

static double scaled_det_2x2(double d00, double d10, double d11, double *s_out) {
    const double s = max3(fabs(d00), fabs(d10), fabs(d11));
    *s_out = s;
    if (!is_fin(s) || s == 0.0) return 0.0;
    const double a = d00 / s, b = d10 / s, c = d11 / s;
    return a * c - b * b;
}

static void add_inertia_2x2(double d00, double d10, double d11, double small, orc_result *r) {
    if (d10 == 0.0) {
        inertia_1x1(d00, small, r);
        inertia_1x1(d11, small, r);
        return;
    }
    double s = 0.0;
    const double det_s = scaled_det_2x2(d00, d10, d11, &s);
    if (s == 0.0 || !is_fin(det_s)) {
        r->zero += 2;
        return;
    }
    const double tr_s = (d00 / s) + (d11 / s);
    if (det_s < -small) {
        ++r->pos;
        ++r->neg;
    } else if (fabs(det_s) <= small) {
        ++r->zero;
        if (tr_s > small) ++r->pos;
        else if (tr_s < -small) ++r->neg;
        else ++r->zero;
    } else {
        if (tr_s > 0.0) r->pos += 2;
        else if (tr_s < 0.0) r->neg += 2;
        else r->zero += 2;
    }
}

static void apply_pivot(int kind, int pa, int pb, double *F, int f, int *kp, int32_t *rows, int8_t *tags,
                        double *W, int w_cols, int ucol_end, const orc_params *prm, orc_result *res) {
    const size_t sf = (size_t)f;
    int k = *kp;
    if (kind == 1) {
        sym_swap(F, sf, f, k, pa, rows, W, w_cols);
        inertia_1x1(F[k * sf + k], prm->eff_small, res);
        tags[k] = 1;
        eliminate_1x1(F, f, k, ucol_end, W, w_cols);
        const double rem = (double)(f - (k + 1));
        res->flops += rem + rem * (rem + 1.0);
        *kp = k + 1;
        return;
    }
    int partner = pb;
    if (pa != k) {
        sym_swap(F, sf, f, k, pa, rows, W, w_cols);
        if (partner == k) partner = pa;
    }
    if (partner != k + 1) sym_swap(F, sf, f, k + 1, partner, rows, W, w_cols);
    add_inertia_2x2(F[k * sf + k], F[k * sf + k + 1], F[(k + 1) * sf + k + 1], prm->small_pivot, res);
    tags[k] = 2;
    tags[k + 1] = -1;
    eliminate_2x2(F, f, k, ucol_end, W, w_cols);
    const double rem = (double)(f - (k + 2));
    res->flops += 6.0 * rem + 2.0 * rem * (rem + 1.0);
    *kp = k + 2;
}

/* F[c0:f, c0:e] -= L[c0:f, p0:p1] W[c0:e, 0:jb]^T, lower part */
static void deferred_update(double *F, int f, int p0, int p1, int c0, int e, const double *W) {
    const int jb = p1 - p0;
    if (jb <= 0 || e - c0 <= 0) return;
    const size_t sf = (size_t)f;
    for (int j = c0; j < e; ++j) {
        double *Cj = F + (size_t)j * sf;
        for (int i = j; i < f; ++i) {
            double s = 0.0;
            for (int t = 0; t < jb; ++t) s += F[(size_t)(p0 + t) * sf + i] * W[(size_t)t * sf + j];
            Cj[i] -= s;
        }
    }
}

/* F[e:f, e:f] -= L[e:f, p0:p] D L[e:f, p0:p]^T, lower */
static void schur_update(double *F, int f, int p0, int p, int e, const int8_t *tags) {
    const int q = f - e;
    if (p <= p0 || q <= 0) return;
    const size_t sf = (size_t)f;
    double *LD = (double *)malloc((size_t)q * (size_t)(p - p0) * sizeof(double));
    for (int c = p0; c < p; ++c) {
        const double *Lc = F + (size_t)c * sf;
        double *dst = LD + (size_t)(c - p0) * q;
        if (tags[c] == 1) {
            for (int i = 0; i < q; ++i) dst[i] = Lc[e + i] * Lc[c];
        } else if (tags[c] == 2) {
            const double *Lc1 = Lc + sf;
            const double d00 = Lc[c], d10 = Lc[c + 1];
            for (int i = 0; i < q; ++i) dst[i] = Lc[e + i] * d00 + Lc1[e + i] * d10;
        } else if (tags[c] == -1) {
            const double *Lc0 = Lc - sf;
            const double d10 = Lc0[c], d11 = Lc[c];
            for (int i = 0; i < q; ++i) dst[i] = Lc0[e + i] * d10 + Lc[e + i] * d11;
        } else {
            for (int i = 0; i < q; ++i) dst[i] = 0.0;
        }
    }
    for (int jj = 0; jj < q; ++jj) {
        double *Cj = F + (size_t)(e + jj) * sf + e;
        for (int ii = jj; ii < q; ++ii) {
            double s = 0.0;
            for (int t = 0; t < p - p0; ++t) s += F[(size_t)(p0 + t) * sf + e + ii] * LD[(size_t)t * q + jj];
            Cj[ii] -= s;
        }
    }
    free(LD);
}

/* ldlt_factor_front, first = 0, threshold pivoting (no static pivots). */
int orc_ldlt_front(double *F, int f, int num_elig, int32_t *rows, const orc_params *prm, int8_t *tags,
                   orc_result *res) {
    memset(res, 0, sizeof(*res));
    int k = 0;
    const size_t sf = (size_t)f;
    if (f < BLOCKED_MIN_FRONT || prm->nb <= 1) {
        while (k < num_elig) {
            int kind, pa, pb;
            select_pivot(F, sf, f, k, num_elig, num_elig, prm, &kind, &pa, &pb);
            if (kind == 0) break;
            apply_pivot(kind, pa, pb, F, f, &k, rows, tags, NULL, 0, f, prm, res);
        }
        res->p = k;
        return k;
    }
    double *W = (double *)calloc(sf * (size_t)(prm->nb + 2), sizeof(double));
    if (!prm->posdef) {
        int live = num_elig, k_paired = -1;
        while (k < live) {
            const int ps = k;
            const int pl = live < ps + prm->nb ? live : ps + prm->nb;
            int stalled = 0;
            while (k < pl) {
                int kind, pa, pb;
                select_pivot(F, sf, f, k, num_elig, pl, prm, &kind, &pa, &pb);
                if (kind == 0) {
                    stalled = 1;
                    break;
                }
                apply_pivot(kind, pa, pb, F, f, &k, rows, tags, W, k - ps, pl, prm, res);
            }
            deferred_update(F, f, ps, k, pl, num_elig, W);
            res->stalls += stalled;
            if (stalled && k != k_paired) {
                int32_t order[128];
                int np = 0;
                const int max_pairs = (prm->nb / 2) < 64 ? prm->nb / 2 : 64;
                for (int pos = k; pos < pl && np / 2 < max_pairs; ++pos) {
                    const double *col = F + (size_t)pos * sf;
                    double in_max = 0.0;
                    for (int i = k; i < pos; ++i) in_max = fmax(in_max, fabs(F[(size_t)i * sf + pos]));
                    for (int i = pos + 1; i < pl; ++i) in_max = fmax(in_max, fabs(col[i]));
                    double out_max = 0.0;
                    int r = -1;
                    for (int i = pl; i < num_elig; ++i) {
                        const double v = fabs(col[i]);
                        if (v > out_max) {
                            out_max = v;
                            r = i;
                        }
                    }
                    if (r < 0 || !(out_max > in_max)) continue;
                    int taken = 0;
                    for (int t = 1; t < np; t += 2) taken = taken || order[t] == rows[r];
                    if (taken) continue;
                    order[np++] = rows[pos];
                    order[np++] = rows[r];
                }
                if (np > 0) {
                    res->pulls += np / 2;
                    for (int t = 0; t < np; ++t) {
                        int q = k + t;
                        while (rows[q] != order[t]) ++q;
                        sym_swap(F, sf, f, k + t, q, rows, NULL, 0);
                    }
                    if (k + np > live) live = k + np;
                    k_paired = k;
                    continue;
                }
            }
            if (stalled) {
                const int m = pl - k;
                const int mv = m < live - pl ? m : live - pl;
                for (int t = 0; t < mv; ++t) sym_swap(F, sf, f, k + t, live - mv + t, rows, NULL, 0);
                live -= m;
                res->rotated += m;
            }
        }
    }
    while (k < num_elig) {
        const int ps = k;
        const int pl = num_elig < ps + prm->nb ? num_elig : ps + prm->nb;
        int stalled = 0;
        while (k < pl) {
            int kind, pa, pb;
            select_pivot(F, sf, f, k, num_elig, pl, prm, &kind, &pa, &pb);
            if (kind == 0) {
                stalled = 1;
                break;
            }
            apply_pivot(kind, pa, pb, F, f, &k, rows, tags, W, k - ps, pl, prm, res);
        }
        deferred_update(F, f, ps, k, pl, num_elig, W);
        if (stalled && k < num_elig) {
            ++res->searches;
            const int ss = k;
            int kind, pa, pb;
            select_pivot(F, sf, f, k, num_elig, num_elig, prm, &kind, &pa, &pb);
            if (kind == 0) break;
            apply_pivot(kind, pa, pb, F, f, &k, rows, tags, W, 0, ss, prm, res);
            deferred_update(F, f, ss, k, k, num_elig, W);
        }
    }
    schur_update(F, f, 0, k, num_elig, tags);
    free(W);
    res->p = k;
    return k;
}

/* ------------------------------------------------------------------------
 * Multifrontal factorization over a given analysis.
 */
typedef struct {
    int32_t n, nsup;
    const int32_t *sn_start, *sn_end, *sn_parent, *child_ptr, *child_idx, *front_row_ptr, *front_row_idx;
    const int32_t *pcol_ptr, *prow_idx, *orig_to_slot, *perm;
    int64_t n_orig;
} orc_analysis;

typedef struct {
    int32_t n, nsup, status;
    int32_t pos, neg, zero, delayed, two_by_two, max_front;
    double flops, front_work, max_abs;
    /* FactorKeep layout */
    int64_t *steps;      /* 4 per supernode: row_offset, factor_offset, p, f */
    double *L;           /* factor values */
    int64_t nL, nrows;
    int32_t *rows;       /* solve row indices */
    int8_t *tags;        /* per stored pivot, in step order */
    int64_t ntags;
    int32_t posdef;
    int32_t stalls, pulls, rotated, searches; /* sums of the fronts' orc_result counters */
} orc_factor;

void orc_factor_free(orc_factor *o) {
    free(o->steps);
    free(o->L);
    free(o->rows);
    free(o->tags);
    memset(o, 0, sizeof(*o));
}

/* status: 0 success, 1 not positive definite, 2 singular; returns status */
int orc_factor_run(const orc_analysis *a, const double *values, double u, double small_pivot, int nb, int posdef,
                   orc_factor *out) {
    memset(out, 0, sizeof(*out));
    const int32_t n = a->n, ns = a->nsup;
    out->n = n;
    out->nsup = ns;
    out->posdef = posdef;
    const int64_t nperm = a->pcol_ptr[n];
    double *av = (double *)calloc((size_t)(nperm > 0 ? nperm : 1), sizeof(double));
    double vmax = 0.0;
    for (int64_t e = 0; e < a->n_orig; ++e) {
        if (!is_fin(values[e])) {
            free(av);
            out->status = 2;
            return 2;
        }
        vmax = fmax(vmax, fabs(values[e]));
    }
    for (int64_t e = 0; e < a->n_orig; ++e)
        if (a->orig_to_slot[e] >= 0) av[a->orig_to_slot[e]] += values[e];
    out->max_abs = vmax;
    const double eff_small = (vmax > 0.0 && vmax < small_pivot) ? small_pivot * vmax : small_pivot;
    orc_params prm = {u, eff_small, small_pivot, nb < 2 ? 2 : nb, posdef};
    /* per-supernode state */
    int32_t *cb_size = (int32_t *)calloc((size_t)ns, sizeof(int32_t));
    int32_t *cb_delayed = (int32_t *)calloc((size_t)ns, sizeof(int32_t));
    double **cb_val = (double **)calloc((size_t)ns, sizeof(double *));
    int32_t **cb_rows = (int32_t **)calloc((size_t)ns, sizeof(int32_t *));
    int32_t *pos = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    for (int32_t i = 0; i < n; ++i) pos[i] = -1;
    size_t capL = 1024, capR = 1024, capT = 1024;
    out->L = (double *)malloc(capL * sizeof(double));
    out->rows = (int32_t *)malloc(capR * sizeof(int32_t));
    out->tags = (int8_t *)malloc(capT);
    out->steps = (int64_t *)malloc((size_t)(ns > 0 ? ns : 1) * 4 * sizeof(int64_t));
    int64_t fo = 0, ro = 0, to = 0;
    int status = 0;
    for (int32_t s = 0; s < ns && status == 0; ++s) {
        const int32_t *frows = a->front_row_idx + a->front_row_ptr[s];
        const int f_sym = a->front_row_ptr[s + 1] - a->front_row_ptr[s];
        const int p_sym = a->sn_end[s] - a->sn_start[s];
        const int32_t *kids = a->child_idx + a->child_ptr[s];
        const int nk = a->child_ptr[s + 1] - a->child_ptr[s];
        int nd = 0;
        for (int c = 0; c < nk; ++c) nd += cb_delayed[kids[c]];
        int cap_rows = nd + f_sym;
        for (int c = 0; c < nk; ++c) cap_rows += cb_size[kids[c]];
        int32_t *rows = (int32_t *)malloc((size_t)(cap_rows > 0 ? cap_rows : 1) * sizeof(int32_t));
        int f = 0;
        for (int c = 0; c < nk; ++c)
            for (int t = 0; t < cb_delayed[kids[c]]; ++t) {
                rows[f] = cb_rows[kids[c]][t];
                pos[rows[f]] = f;
                ++f;
            }
        for (int t = 0; t < f_sym; ++t) {
            rows[f] = frows[t];
            pos[frows[t]] = f;
            ++f;
        }
        const int num_elig = nd + p_sym;
        for (int c = 0; c < nk; ++c)
            for (int t = cb_delayed[kids[c]]; t < cb_size[kids[c]]; ++t)
                if (pos[cb_rows[kids[c]][t]] < 0) {
                    rows[f] = cb_rows[kids[c]][t];
                    pos[rows[f]] = f;
                    ++f;
                }
        const size_t sf = (size_t)f;
        double *F = (double *)calloc(sf * sf, sizeof(double));
        for (int32_t j = a->sn_start[s]; j < a->sn_end[s]; ++j) {
            double *Fj = F + (size_t)pos[j] * sf;
            for (int32_t k = a->pcol_ptr[j]; k < a->pcol_ptr[j + 1]; ++k) Fj[pos[a->prow_idx[k]]] += av[k];
        }
        for (int c = 0; c < nk; ++c) {
            const int32_t ch = kids[c];
            const int q = cb_size[ch];
            const double *cv = cb_val[ch];
            for (int jj = 0; jj < q; ++jj) {
                const int lj = pos[cb_rows[ch][jj]];
                for (int ii = jj; ii < q; ++ii) {
                    const int li = pos[cb_rows[ch][ii]];
                    if (li >= lj) F[(size_t)lj * sf + li] += *cv++;
                    else F[(size_t)li * sf + lj] += *cv++;
                }
            }
            free(cb_val[ch]);
            free(cb_rows[ch]);
            cb_val[ch] = NULL;
            cb_rows[ch] = NULL;
        }
        int8_t *tags = (int8_t *)calloc(sf > 0 ? sf : 1, 1);
        orc_result kr;
        memset(&kr, 0, sizeof(kr));
        if (posdef) {
            /* in-order LDL^T without pivot search (ldlt_factor_front(prm.posdef)),
               then L <- L sqrt(D) as the SPD store does */
            orc_params pp = prm;
            pp.nb = 1; /* unblocked; the blocked variant differs only in rounding */
            orc_ldlt_front(F, f, num_elig, rows, &pp, tags, &kr);
        } else {
            orc_ldlt_front(F, f, num_elig, rows, &prm, tags, &kr);
        }
        const int p = kr.p;
        out->pos += kr.pos;
        out->neg += kr.neg;
        out->zero += kr.zero;
        out->flops += kr.flops;
        out->stalls += kr.stalls;
        out->pulls += kr.pulls;
        out->rotated += kr.rotated;
        out->searches += kr.searches;
        out->front_work += (double)num_elig * (double)f * (double)f;
        if (f > out->max_front) out->max_front = f;
        const int is_root = a->sn_parent[s] == -1;
        if (posdef && p < num_elig) {
            status = 1;
            free(F);
            free(tags);
            for (int t = 0; t < f; ++t) pos[rows[t]] = -1;
            free(rows);
            break;
        }
        int p_store = p;
        if (is_root && p < num_elig) {
            for (int c = p; c < num_elig; ++c) tags[c] = 0;
            out->zero += num_elig - p;
            p_store = num_elig;
        }
        /* store: columns [0, p_store) of F, f x p_store at fo */
        if ((size_t)(fo + (int64_t)sf * p_store) > capL) {
            while ((size_t)(fo + (int64_t)sf * p_store) > capL) capL *= 2;
            out->L = (double *)realloc(out->L, capL * sizeof(double));
        }
        double *Ls = out->L + fo;
        if (posdef) {
            for (int c = 0; c < p_store; ++c) {
                double *Lc = F + (size_t)c * sf;
                const double sd = sqrt(Lc[c]);
                for (int i = 0; i < c; ++i) Lc[i] = 0.0;
                Lc[c] = sd;
                for (int i = c + 1; i < f; ++i) Lc[i] *= sd;
            }
        } else {
            double d10 = 0.0;
            for (int c = 0; c < p_store; ++c) {
                double *Lc = F + (size_t)c * sf;
                const int8_t tag = tags[c];
                if (tag == 0) {
                    for (int i = c + 1; i < f; ++i) Lc[i] = 0.0;
                    Lc[c] = 0.0;
                } else if (tag == 2) {
                    ++out->two_by_two;
                    d10 = Lc[c + 1];
                    Lc[c + 1] = 0.0;
                }
                for (int i = 0; i < c; ++i) Lc[i] = 0.0;
                if (tag == -1) Lc[c - 1] = d10;
            }
        }
        memcpy(Ls, F, sf * (size_t)p_store * sizeof(double));
        out->steps[4 * s + 0] = ro;
        out->steps[4 * s + 1] = fo;
        out->steps[4 * s + 2] = p_store;
        out->steps[4 * s + 3] = f;
        if ((size_t)(ro + f) > capR) {
            while ((size_t)(ro + f) > capR) capR *= 2;
            out->rows = (int32_t *)realloc(out->rows, capR * sizeof(int32_t));
        }
        memcpy(out->rows + ro, rows, sf * sizeof(int32_t));
        if ((size_t)(to + p_store) > capT) {
            while ((size_t)(to + p_store) > capT) capT *= 2;
            out->tags = (int8_t *)realloc(out->tags, capT);
        }
        memcpy(out->tags + to, tags, (size_t)p_store);
        to += p_store;
        fo += (int64_t)sf * p_store;
        ro += f;
        /* contribution block: F[p:f, p:f] packed lower, delayed rows first */
        const int q = f - p;
        if (!is_root && q > 0) {
            double *cv = (double *)malloc((size_t)q * (size_t)(q + 1) / 2 * sizeof(double));
            int32_t *cr = (int32_t *)malloc((size_t)q * sizeof(int32_t));
            size_t at = 0;
            for (int jj = 0; jj < q; ++jj)
                for (int ii = jj; ii < q; ++ii) cv[at++] = F[(size_t)(p + jj) * sf + p + ii];
            memcpy(cr, rows + p, (size_t)q * sizeof(int32_t));
            cb_val[s] = cv;
            cb_rows[s] = cr;
            cb_size[s] = q;
            cb_delayed[s] = num_elig - p;
            out->delayed += num_elig - p;
        }
        for (int t = 0; t < f; ++t) pos[rows[t]] = -1;
        free(F);
        free(tags);
        free(rows);
    }
    if (status == 0 && !posdef && out->zero > 0) status = 2;
    out->nL = fo;
    out->nrows = ro;
    out->ntags = to;
    for (int32_t s = 0; s < ns; ++s) {
        free(cb_val[s]);
        free(cb_rows[s]);
    }
    free(cb_val);
    free(cb_rows);
    free(cb_size);
    free(cb_delayed);
    free(pos);
    free(av);
    out->status = status;
    return status;
}

/* Full solve x = P (L D L^T)^{-1} P^T b in place (b: n x nrhs, ld n).
   Steps in supernode order; rows of each step in solve_row_indices. */
int orc_solve(const orc_factor *o, const int32_t *perm, double *b, int nrhs) {
    const int32_t n = o->n;
    double *w = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    for (int r = 0; r < nrhs; ++r) {
        double *bc = b + (size_t)r * n;
        for (int32_t i = 0; i < n; ++i) w[i] = bc[perm[i]];
        /* forward: L^{-1} (unit lower for LDL^T, Cholesky L for SPD) */
        for (int32_t s = 0; s < o->nsup; ++s) {
            const int64_t ro = o->steps[4 * s], fo = o->steps[4 * s + 1];
            const int p = (int)o->steps[4 * s + 2], f = (int)o->steps[4 * s + 3];
            const double *L = o->L + fo;
            const int32_t *rows = o->rows + ro;
            for (int c = 0; c < p; ++c) {
                double xc = w[rows[c]];
                if (o->posdef) xc /= L[(size_t)c * f + c];
                w[rows[c]] = xc;
                for (int i = c + 1; i < f; ++i) w[rows[i]] -= L[(size_t)c * f + i] * xc;
            }
        }
        /* diagonal (LDL^T): D inline, 2x2 off-diagonal in slot (c, c+1) */
        if (!o->posdef) {
            int64_t to = 0;
            for (int32_t s = 0; s < o->nsup; ++s) {
                const int64_t ro = o->steps[4 * s], fo = o->steps[4 * s + 1];
                const int p = (int)o->steps[4 * s + 2], f = (int)o->steps[4 * s + 3];
                const double *L = o->L + fo;
                const int32_t *rows = o->rows + ro;
                for (int c = 0; c < p;) {
                    const double d00 = L[(size_t)c * f + c];
                    if (c + 1 < p && L[(size_t)(c + 1) * f + c] != 0.0) {
                        const double d10 = L[(size_t)(c + 1) * f + c], d11 = L[(size_t)(c + 1) * f + c + 1];
                        double sc = 0.0;
                        const double det_s = scaled_det_2x2(d00, d10, d11, &sc);
                        double e0 = 0, e1 = 0, e2 = 0;
                        if (!(sc == 0.0 || !is_fin(det_s) || det_s == 0.0)) {
                            e0 = (d11 / sc) / det_s / sc;
                            e1 = -(d10 / sc) / det_s / sc;
                            e2 = (d00 / sc) / det_s / sc;
                        }
                        const double r0 = w[rows[c]], r1 = w[rows[c + 1]];
                        w[rows[c]] = e0 * r0 + e1 * r1;
                        w[rows[c + 1]] = e1 * r0 + e2 * r1;
                        c += 2;
                    } else {
                        const double di = (d00 == 0.0 || !is_fin(d00)) ? 0.0 : 1.0 / d00;
                        w[rows[c]] *= di;
                        c += 1;
                    }
                }
                to += p;
            }
        }
        /* backward: L^{-T} */
        for (int32_t s = o->nsup - 1; s >= 0; --s) {
            const int64_t ro = o->steps[4 * s], fo = o->steps[4 * s + 1];
            const int p = (int)o->steps[4 * s + 2], f = (int)o->steps[4 * s + 3];
            const double *L = o->L + fo;
            const int32_t *rows = o->rows + ro;
            for (int c = p - 1; c >= 0; --c) {
                double xc = w[rows[c]];
                for (int i = c + 1; i < f; ++i) xc -= L[(size_t)c * f + i] * w[rows[i]];
                if (o->posdef) xc /= L[(size_t)c * f + c];
                w[rows[c]] = xc;
            }
        }
        for (int32_t i = 0; i < n; ++i) bc[perm[i]] = w[i];
    }
    free(w);
    return 0;
}

/* compute_residual_and_backward_error over the lower CSC (n, cp, ri, v) of the
   analysed matrix (duplicates summed into canonical entries first, as
   refine does), long double products with Neumaier sums. out: eta_inf,
   omega (maximum over the right-hand sides). */
typedef struct {
    long double sum, c;
} neu;

static void neu_add(neu *a, long double v) {
    const long double t = a->sum + v;
    if (fabsl(a->sum) >= fabsl(v)) a->c += (a->sum - t) + v;
    else a->c += (v - t) + a->sum;
    a->sum = t;
}

void orc_backward_error(int32_t n, const int32_t *cp, const int32_t *ri, const double *v, const double *b,
                        const double *x, int nrhs, double *eta_out, double *omega_out) {
    /* canonical entries: sort each column's rows, sum duplicates (long double) */
    const int64_t nnz = cp[n];
    int64_t *ord = (int64_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t));
    for (int64_t k = 0; k < nnz; ++k) ord[k] = k;
    /* insertion sort inside columns (columns are short) */
    for (int32_t j = 0; j < n; ++j)
        for (int64_t k = cp[j] + 1; k < cp[j + 1]; ++k) {
            const int64_t key = ord[k];
            int64_t m = k - 1;
            while (m >= cp[j] && ri[ord[m]] > ri[key]) {
                ord[m + 1] = ord[m];
                --m;
            }
            ord[m + 1] = key;
        }
    int32_t *er = (int32_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
    int32_t *ec = (int32_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
    long double *ev = (long double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(long double));
    int64_t ne = 0;
    for (int32_t j = 0; j < n; ++j)
        for (int64_t k = cp[j]; k < cp[j + 1]; ++k) {
            const int32_t i = ri[ord[k]];
            if (i < j || i >= n) continue;
            if (ne > 0 && er[ne - 1] == i && ec[ne - 1] == j) ev[ne - 1] += (long double)v[ord[k]];
            else {
                er[ne] = i;
                ec[ne] = j;
                ev[ne] = (long double)v[ord[k]];
                ++ne;
            }
        }
    /* ||A||_inf: max row sum of |A| (symmetric) */
    long double *rs = (long double *)calloc((size_t)(n > 0 ? n : 1), sizeof(long double));
    for (int64_t e = 0; e < ne; ++e) {
        rs[er[e]] += fabsl(ev[e]);
        if (er[e] != ec[e]) rs[ec[e]] += fabsl(ev[e]);
    }
    long double normA = 0.0L;
    for (int32_t i = 0; i < n; ++i) normA = rs[i] > normA ? rs[i] : normA;
    double eta = 0.0, omega = 0.0;
    neu *ax = (neu *)malloc((size_t)(n > 0 ? n : 1) * sizeof(neu));
    neu *aax = (neu *)malloc((size_t)(n > 0 ? n : 1) * sizeof(neu));
    for (int r = 0; r < nrhs; ++r) {
        const double *bc = b + (size_t)r * n, *xc = x + (size_t)r * n;
        long double nb = 0, nx = 0;
        int bad = 0;
        for (int32_t i = 0; i < n; ++i) {
            if (!is_fin(bc[i]) || !is_fin(xc[i])) bad = 1;
            nb = fabsl((long double)bc[i]) > nb ? fabsl((long double)bc[i]) : nb;
            nx = fabsl((long double)xc[i]) > nx ? fabsl((long double)xc[i]) : nx;
        }
        if (bad) {
            eta = omega = INFINITY;
            break;
        }
        memset(ax, 0, (size_t)n * sizeof(neu));
        memset(aax, 0, (size_t)n * sizeof(neu));
        for (int64_t e = 0; e < ne; ++e) {
            const int32_t i = er[e], j = ec[e];
            const long double val = ev[e], a = fabsl(val);
            neu_add(&ax[i], val * (long double)xc[j]);
            neu_add(&aax[i], a * fabsl((long double)xc[j]));
            if (i != j) {
                neu_add(&ax[j], val * (long double)xc[i]);
                neu_add(&aax[j], a * fabsl((long double)xc[i]));
            }
        }
        long double rinf = 0, om = 0;
        for (int32_t i = 0; i < n; ++i) {
            const long double res = (long double)bc[i] - (ax[i].sum + ax[i].c);
            const long double ar = fabsl(res);
            rinf = ar > rinf ? ar : rinf;
            const long double den = (aax[i].sum + aax[i].c) + fabsl((long double)bc[i]);
            long double oi = 0;
            if (den > 0) oi = ar / den;
            else if (ar != 0) oi = INFINITY;
            om = oi > om ? oi : om;
        }
        const long double ed = normA * nx + nb;
        const long double et = ed > 0 ? rinf / ed : (rinf == 0 ? 0.0L : (long double)INFINITY);
        if ((double)et > eta) eta = (double)et;
        if ((double)om > omega) omega = (double)om;
    }
    *eta_out = eta;
    *omega_out = omega;
    free(ord);
    free(er);
    free(ec);
    free(ev);
    free(rs);
    free(ax);
    free(aax);
}
