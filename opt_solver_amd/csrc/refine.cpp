// A16: iterative refinement against the analysed matrix, on the device
// (the reference's refine / refine_core, solver/src/refinement.cpp:300-724).
//
// Every vector lives on the device in the permuted numbering of the
// factorization's CSR; the residual and the per-column maxima of the
// backward errors are reduced on the device (k_residual), so one
// refinement step moves 4 nrhs doubles to the host. Methods:
//  - classical: x <- x_best + M^-1 r, keeping the best iterate; stop at the
//    targets, on a step that does not improve the score, or after two steps
//    that contract it by less than 0.85 (classical_refine,
//    refinement.cpp:317-396);
//  - Auto: the same steps, escalating to FGMRES when a step contracts the
//    score by less than 0.5 or the worse step is reached (and, on a
//    statically pivoted factor, when min(3, cap) steps leave the targets
//    unmet);
//  - FGMRES(restart), right-preconditioned by the factor, one column at a
//    time, restarted from the best iterate with its true residual
//    (fgmres_column, refinement.cpp:451-601). Orthogonalisation is classical
//    Gram-Schmidt applied twice (CGS2, two dot-product passes per basis
//    vector on the device) instead of the reference's modified
//    Gram-Schmidt: the same orthogonality to working precision with
//    O(1) launches per basis vector. The small Hessenberg least-squares
//    problem (Givens rotations) is solved on the host.
// The score of an iterate is max(eta / eta_target, omega / omega_target)
// (metric_score, refinement.cpp:283-285).
#include "smfgpu_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace smfgpu {
namespace {

constexpr double kMaxGrowth = 1e8;  // FGMRES: largest accepted ||x|| over the reference scale (refinement.cpp:449)

struct Ctx {
    smfgpu_plan_s *P;
    hipStream_t st;
    int32_t n;
    int32_t nrhs;
    const double *av;   // values of the residual's matrix (permuted slots)
    double norm_a;      // ||A||_inf
    double *b;          // permuted right-hand sides, n x nrhs
    double *r;          // residual of the last evaluated iterate
    double *stat;       // 4 nrhs maxima
    double *tmp;        // n x nrhs: original-order staging of the solves
    double te, to;      // targets

    double score(double eta, double om) const { return std::max(eta / te, om / to); }
    bool reached(double eta, double om) const { return eta <= te && om <= to; }

    // r = b - A x for columns [c0, c0 + nc) (x, b, r at column c0), eta and
    // omega over those columns; false when a value is not finite
    int eval(const double *x, const double *bb, double *rr, int32_t nc, double &eta, double &om) {
        launch_residual(st, n, P->d_fr_ptr.p, P->d_fr_col.p, P->d_fr_slot.p, av, x, bb, rr, stat, nc);
        std::vector<double> h(4 * static_cast<size_t>(nc));
        HIP_OK(hipMemcpyAsync(h.data(), stat, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        eta = 0.0;
        om = 0.0;
        for (int32_t c = 0; c < nc; ++c) {
            const double rmax = h[4 * c], omc = h[4 * c + 1], xmax = h[4 * c + 2], bmax = h[4 * c + 3];
            const double den = norm_a * xmax + bmax;
            const double e = den > 0.0 ? rmax / den : (rmax == 0.0 ? 0.0 : INFINITY);
            eta = std::max(eta, e);
            om = std::max(om, omc);
            if (!std::isfinite(e) || !std::isfinite(omc) || !std::isfinite(xmax)) {
                eta = om = INFINITY;
                return SMFGPU_OK;
            }
        }
        return SMFGPU_OK;
    }

    // z = M^-1 v for nc permuted columns (the solve takes the original
    // order: scatter, solve, gather)
    int precond(const double *v, double *z, int32_t nc) {
        launch_scatter_rhs(st, P->d_perm.p, v, tmp, n, n, nc);
        const int rc = solve_device(P, tmp, n, tmp, n, nc, 0);
        if (rc) return rc;
        launch_gather_rhs(st, P->d_perm.p, tmp, n, z, n, nc);
        return SMFGPU_OK;
    }
};

struct Outcome {
    double eta, om;
    int steps = 0;
    bool escalate = false;
};

// classical steps from x (n x nrhs, residual in c.r): xbest / xc are two
// device vectors, the result is left in xbest
int classical(Ctx &c, double *&xbest, double *&xc, double *dx, double eta, double om, int cap, bool escalate,
              bool static_factor, Outcome &out) {
    const size_t nn = static_cast<size_t>(c.n) * c.nrhs;
    out.eta = eta;
    out.om = om;
    double best = c.score(eta, om);
    int stagnated = 0;
    const int static_check = std::min(3, cap);
    for (int it = 1; it <= cap; ++it) {
        out.steps = it;
        int rc = c.precond(c.r, dx, c.nrhs);
        if (rc) return rc;
        launch_axpy(c.st, static_cast<int64_t>(nn), xbest, dx, xc);
        double ce, co;
        if ((rc = c.eval(xc, c.b, c.r, c.nrhs, ce, co))) return rc;
        const double cs = c.score(ce, co);
        if (cs < best) {
            const double ratio = cs / best;
            std::swap(xbest, xc);
            out.eta = ce;
            out.om = co;
            best = cs;
            if (c.reached(ce, co)) return SMFGPU_OK;
            if (escalate) {
                if (ratio > 0.5 || (static_factor && it >= static_check)) {
                    out.escalate = true;
                    return SMFGPU_OK;
                }
            } else if (ratio > 0.85) {
                if (++stagnated >= 2) return SMFGPU_OK;
            } else {
                stagnated = 0;
            }
        } else {
            // the candidate is no better: keep the best iterate and stop
            out.escalate = escalate && !c.reached(out.eta, out.om);
            return SMFGPU_OK;
        }
    }
    return SMFGPU_OK;
}

// FGMRES(restart) on column col of x (device, n), budget solves; x is left at
// the best iterate, eta / om its backward errors
int fgmres_column(Ctx &c, double *x, const double *bcol, int budget, int restart, double &eta, double &om,
                  int &used) {
    const int32_t n = c.n;
    const size_t un = static_cast<size_t>(n);
    used = 0;
    int rc;
    // residual of the start in c.r (column 0 of the scratch)
    if ((rc = c.eval(x, bcol, c.r, 1, eta, om))) return rc;
    if (!std::isfinite(eta) || budget <= 0 || c.reached(eta, om)) return SMFGPU_OK;
    // reference scale of ||x|| (refinement.cpp:476-484)
    double xref = 0.0, bmax = 0.0;
    {
        double h[4];
        HIP_OK(hipMemcpyAsync(h, c.stat, 4 * sizeof(double), hipMemcpyDeviceToHost, c.st));
        HIP_OK(hipStreamSynchronize(c.st));
        xref = h[2];
        bmax = h[3];
    }
    if (c.norm_a > 0.0) xref = std::max(xref, bmax / c.norm_a);
    const double x_limit = kMaxGrowth * xref;
    const int m_max = std::max(1, std::min(restart, budget));
    const size_t mm = static_cast<size_t>(m_max);
    // device vectors: V (m + 1), Z (m), x0, xc, rc (candidate residual), w = V_{j+1}
    smfgpu_plan_s *P = c.P;
    const size_t need = (2 * mm + 5) * un;
    if (P->d_rvec.n < need && (rc = P->d_rvec.alloc(need))) return rc;
    double *V = P->d_rvec.p;
    double *Z = V + (mm + 1) * un;
    double *x0 = Z + mm * un;
    double *xc = x0 + un;
    double *rcand = xc + un;
    double *rbest = rcand + un;
    if (P->d_h.n < 2 * mm + 8 && (rc = P->d_h.alloc(2 * mm + 8))) return rc;
    double *dh = P->d_h.p;  // [0, m+1): h column, [m+1]: norm, then y
    if (P->d_part.n < static_cast<size_t>(kDotParts) * (mm + 1) &&
        (rc = P->d_part.alloc(static_cast<size_t>(kDotParts) * (mm + 1))))
        return rc;
    HIP_OK(hipMemcpyAsync(x0, x, un * sizeof(double), hipMemcpyDeviceToDevice, c.st));
    HIP_OK(hipMemcpyAsync(rbest, c.r, un * sizeof(double), hipMemcpyDeviceToDevice, c.st));
    std::vector<double> H((mm + 1) * mm), cs(mm), sn(mm), g(mm + 1), y(mm), hcol(mm + 2);
    auto h = [&](size_t i, size_t j) -> double & { return H[j * (mm + 1) + i]; };
    double best = c.score(eta, om);
    bool done = false;
    while (used < budget && !done) {
        // beta = ||r_best||, v_0 = r_best / beta
        launch_dots(c.st, n, rbest, 1, rbest, P->d_part.p, dh + mm + 1, 2);
        double beta = 0.0;
        HIP_OK(hipMemcpyAsync(&beta, dh + mm + 1, sizeof(double), hipMemcpyDeviceToHost, c.st));
        HIP_OK(hipStreamSynchronize(c.st));
        if (!(beta > 0.0) || !std::isfinite(beta)) break;
        launch_vdiv(c.st, n, rbest, dh + mm + 1, V);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        const int m = std::min(m_max, budget - used);
        bool improved = false, stop = false;
        for (int j = 0; j < m; ++j) {
            double *vj = V + static_cast<size_t>(j) * un;
            double *zj = Z + static_cast<size_t>(j) * un;
            ++used;
            if ((rc = c.precond(vj, zj, 1))) return rc;
            double *w = V + static_cast<size_t>(j + 1) * un;
            launch_spmv(c.st, n, P->d_fr_ptr.p, P->d_fr_col.p, P->d_fr_slot.p, c.av, zj, w);
            // CGS2: h = V^T w, w -= V h, twice; then ||w||
            launch_dots(c.st, n, V, j + 1, w, P->d_part.p, dh, 0);
            launch_vsub(c.st, n, V, j + 1, dh, w);
            launch_dots(c.st, n, V, j + 1, w, P->d_part.p, dh + mm + 2, 0);
            launch_vsub(c.st, n, V, j + 1, dh + mm + 2, w);
            launch_dots(c.st, n, w, 1, w, P->d_part.p, dh + mm + 1, 2);
            HIP_OK(hipMemcpyAsync(hcol.data(), dh, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, c.st));
            std::vector<double> h2(j + 1);
            double hnext = 0.0;
            HIP_OK(hipMemcpyAsync(h2.data(), dh + mm + 2, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, c.st));
            HIP_OK(hipMemcpyAsync(&hnext, dh + mm + 1, sizeof(double), hipMemcpyDeviceToHost, c.st));
            HIP_OK(hipStreamSynchronize(c.st));
            for (int i = 0; i <= j; ++i) h(i, j) = hcol[i] + h2[i];
            if (!std::isfinite(hnext)) {
                stop = true;
                break;
            }
            // previous rotations on column j, then the one zeroing h(j+1, j)
            for (int i = 0; i < j; ++i) {
                const double a = h(i, j), b2 = h(i + 1, j);
                h(i, j) = cs[i] * a + sn[i] * b2;
                h(i + 1, j) = -sn[i] * a + cs[i] * b2;
            }
            const double diag = h(j, j);
            const double rho = std::hypot(diag, hnext);
            if (rho == 0.0) {
                stop = true;
                break;
            }
            cs[j] = diag / rho;
            sn[j] = hnext / rho;
            h(j, j) = rho;
            h(j + 1, j) = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            // y of the (j+1)-column problem, candidate x0 + Z y, true residual
            const int k = j + 1;
            for (int i = k - 1; i >= 0; --i) {
                double s = g[i];
                for (int l = i + 1; l < k; ++l) s -= h(i, l) * y[l];
                y[i] = s / h(i, i);
            }
            HIP_OK(hipMemcpyAsync(dh + mm + 2, y.data(), k * sizeof(double), hipMemcpyHostToDevice, c.st));
            launch_lincomb(c.st, n, x0, Z, k, dh + mm + 2, xc);
            double ce, co;
            if ((rc = c.eval(xc, bcol, rcand, 1, ce, co))) return rc;
            double xcmax = 0.0;
            HIP_OK(hipMemcpyAsync(&xcmax, c.stat + 2, sizeof(double), hipMemcpyDeviceToHost, c.st));
            HIP_OK(hipStreamSynchronize(c.st));
            if (xcmax <= x_limit && std::isfinite(ce)) {
                const double sc = c.score(ce, co);
                if (sc < best) {
                    best = sc;
                    eta = ce;
                    om = co;
                    HIP_OK(hipMemcpyAsync(x, xc, un * sizeof(double), hipMemcpyDeviceToDevice, c.st));
                    std::swap(rbest, rcand);
                    improved = true;
                    if (c.reached(eta, om)) {
                        done = true;
                        break;
                    }
                }
            }
            if (hnext <= 0.0) break;  // the Krylov space is invariant
            launch_vdiv(c.st, n, w, dh + mm + 1, w);
        }
        if (stop || !improved) break;
        HIP_OK(hipMemcpyAsync(x0, x, un * sizeof(double), hipMemcpyDeviceToDevice, c.st));
    }
    HIP_OK(hipStreamSynchronize(c.st));
    return SMFGPU_OK;
}

}  // namespace

int refine_run(smfgpu_factor_s *f, const double *values, int32_t value_mem, const double *b, int64_t ldb, double *x,
               int64_t ldx, int32_t nrhs, const smfgpu_refine_options &o, int32_t mem, smfgpu_refine_info &info) {
    smfgpu_plan_s *P = f->plan;
    hipStream_t st = P->stream;
    const int32_t n = P->n;
    const size_t un = static_cast<size_t>(n), nn = un * nrhs;
    std::memset(&info, 0, sizeof(info));
    int rc;
    // the residual's matrix: the factored values (unscaled when the factor
    // is of S A S), or others mapped the same way
    const double *av = P->scaled ? P->d_av_raw.p : P->d_av.p;
    double norm_a = P->norm_a;
    bool have_norm = P->norm_a_valid;
    if (values) {
        if (P->d_av_ref.n < static_cast<size_t>(std::max<int64_t>(P->nnz_perm, 1)) &&
            (rc = P->d_av_ref.alloc(std::max<int64_t>(P->nnz_perm, 1))))
            return rc;
        const double *vd = values;
        if (value_mem == SMFGPU_MEM_HOST) {
            if (P->d_vals.n < static_cast<size_t>(P->n_orig) && (rc = P->d_vals.alloc(P->n_orig))) return rc;
            HIP_OK(hipMemcpyAsync(P->d_vals.p, values, P->n_orig * sizeof(double), hipMemcpyHostToDevice, st));
            vd = P->d_vals.p;
        }
        launch_map_values(st, vd, P->n_orig, P->d_slot_ptr.p, P->d_slot_src.p, P->nnz_perm, P->d_av_ref.p,
                          P->d_red.p + 2);
        av = P->d_av_ref.p;
        have_norm = false;
    }
    // vectors: xbest, xc, dx (n x nrhs each), then b, r, tmp
    if (P->d_x.n < 3 * nn && (rc = P->d_x.alloc(3 * nn))) return rc;
    if (P->d_bb.n < 3 * nn && (rc = P->d_bb.alloc(3 * nn))) return rc;
    if (P->d_stat.n < 4 * static_cast<size_t>(nrhs) && (rc = P->d_stat.alloc(4 * static_cast<size_t>(nrhs)))) return rc;
    double *xbest = P->d_x.p, *xc = xbest + nn, *dx = xc + nn;
    Ctx c{P, st, n, nrhs, av, 0.0, P->d_bb.p, P->d_bb.p + nn, P->d_stat.p, P->d_bb.p + 2 * nn, o.eta_target,
          o.omega_target};
    if (!have_norm) {
        if (P->d_r.n < un && (rc = P->d_r.alloc(un))) return rc;
        launch_rowabs(st, n, P->d_fr_ptr.p, P->d_fr_slot.p, av, P->d_r.p);
        launch_reduce_max(st, P->d_r.p, n, P->d_red.p + 1);
        HIP_OK(hipMemcpyAsync(&norm_a, P->d_red.p + 1, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (!values) {
            P->norm_a = norm_a;
            P->norm_a_valid = true;
        }
    }
    c.norm_a = norm_a;
    // permuted copies of b and x (columns ld n)
    const hipMemcpyKind kin = mem == SMFGPU_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind kout = mem == SMFGPU_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    HIP_OK(hipMemcpy2DAsync(c.tmp, un * sizeof(double), b, ldb * sizeof(double), un * sizeof(double), nrhs, kin, st));
    launch_gather_rhs(st, P->d_perm.p, c.tmp, n, c.b, n, nrhs);
    HIP_OK(hipMemcpy2DAsync(c.tmp, un * sizeof(double), x, ldx * sizeof(double), un * sizeof(double), nrhs, kin, st));
    launch_gather_rhs(st, P->d_perm.p, c.tmp, n, xbest, n, nrhs);
    double eta, om;
    if ((rc = c.eval(xbest, c.b, c.r, nrhs, eta, om))) return rc;
    info.eta_start = eta;
    info.omega_start = om;
    info.method_used = SMFGPU_REFINE_CLASSICAL;
    if (!std::isfinite(eta)) {
        info.eta = eta;
        info.omega = om;
        info.stalled = 1;
        return set_error(SMFGPU_ERR_VALUE, "refine: non-finite right-hand side or solution");
    }
    Outcome out;
    out.eta = eta;
    out.om = om;
    const bool static_factor = f->info.static_pivots > 0;
    if (!c.reached(eta, om) && o.method != SMFGPU_REFINE_FGMRES && o.max_iters > 0) {
        const bool esc = o.method == SMFGPU_REFINE_AUTO;
        const int cap = esc ? std::min(o.max_iters, o.max_steps) : o.max_iters;
        if ((rc = classical(c, xbest, xc, dx, eta, om, cap, esc, static_factor, out))) return rc;
    }
    int steps = out.steps;
    const bool fg = o.method == SMFGPU_REFINE_FGMRES || (o.method == SMFGPU_REFINE_AUTO && out.escalate);
    if (fg && !c.reached(out.eta, out.om) && steps < o.max_steps) {
        info.method_used = SMFGPU_REFINE_FGMRES;
        double e_all = 0.0, o_all = 0.0;
        int most = 0;
        for (int32_t col = 0; col < nrhs; ++col) {
            double e1, o1;
            int used = 0;
            if ((rc = fgmres_column(c, xbest + col * un, c.b + col * un, o.max_steps - steps, o.restart, e1, o1,
                                    used)))
                return rc;
            most = std::max(most, used);
            e_all = std::max(e_all, e1);
            o_all = std::max(o_all, o1);
        }
        steps += most;
        out.eta = e_all;
        out.om = o_all;
    }
    launch_scatter_rhs(st, P->d_perm.p, xbest, c.tmp, n, n, nrhs);
    HIP_OK(hipMemcpy2DAsync(x, ldx * sizeof(double), c.tmp, un * sizeof(double), un * sizeof(double), nrhs, kout, st));
    HIP_OK(hipStreamSynchronize(st));
    if (const int rc = sweep_check(P)) return rc;
    info.eta = out.eta;
    info.omega = out.om;
    info.steps = steps;
    info.reached = c.reached(out.eta, out.om) ? 1 : 0;
    info.stalled = info.reached ? 0 : 1;
    return SMFGPU_OK;
}

}  // namespace smfgpu

using namespace smfgpu;

extern "C" {

SMFGPU_API int smfgpu_refine_default_options(smfgpu_refine_options *o) {
    if (!o) return set_error(SMFGPU_ERR_INVALID_ARG, "null options");
    std::memset(o, 0, sizeof(*o));
    o->method = SMFGPU_REFINE_AUTO;
    o->max_iters = 5;
    o->max_steps = 10;
    o->restart = 20;
    o->eta_target = 1e-10;
    o->omega_target = 1e-8;
    o->value_mem = SMFGPU_MEM_HOST;
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_refine_ex(smfgpu_factor_t f, const double *values, const double *b, int32_t ldb, double *x,
                                int32_t ldx, int32_t nrhs, const smfgpu_refine_options *opt, int32_t mem,
                                smfgpu_refine_info *info) {
    if (!f || !b || !x) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    if (!f->plan) return set_error(SMFGPU_ERR_STATE, "refine: the factor's plan was freed");
    smfgpu_plan_s *P = f->plan;
    if (P->live != f || !f->valid) return set_error(SMFGPU_ERR_STATE, "refine: factor not usable");
    if (nrhs <= 0 || ldb < P->n || ldx < P->n) return set_error(SMFGPU_ERR_INVALID_ARG, "refine: bad nrhs/ldb/ldx");
    smfgpu_refine_options o;
    if (opt) o = *opt;
    else smfgpu_refine_default_options(&o);
    if (o.method < SMFGPU_REFINE_CLASSICAL || o.method > SMFGPU_REFINE_FGMRES || o.max_iters < 0 ||
        o.max_steps < 0 || o.restart < 1 || !(o.eta_target > 0.0) || !(o.omega_target > 0.0))
        return set_error(SMFGPU_ERR_INVALID_ARG, "refine: bad options");
    smfgpu_refine_info inf{};
    try {
        const int rc = refine_run(f, values, o.value_mem, b, ldb, x, ldx, nrhs, o, mem, inf);
        if (info) *info = inf;
        return rc;
    } catch (const std::bad_alloc &) {
        return set_error(SMFGPU_ERR_ALLOC, "refine: host allocation failed");
    }
}

SMFGPU_API int smfgpu_refine(smfgpu_factor_t f, const double *values, const double *b, double *x, int32_t nrhs,
                             int32_t max_iters, double tol, int32_t mem, double *out) {
    smfgpu_refine_options o;
    smfgpu_refine_default_options(&o);
    o.method = SMFGPU_REFINE_AUTO;
    o.max_iters = std::max(0, max_iters);
    o.max_steps = std::max(0, max_iters);
    o.eta_target = tol > 0.0 ? tol : 1e-300;
    o.omega_target = std::numeric_limits<double>::max();
    o.value_mem = mem;
    if (!f) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    const int32_t n = f->plan ? f->plan->n : 0;
    smfgpu_refine_info inf{};
    const int rc = smfgpu_refine_ex(f, values, b, n, x, n, nrhs, &o, mem, &inf);
    if (rc == SMFGPU_OK && out) {
        out[0] = inf.eta;
        out[1] = inf.omega;
        out[2] = inf.steps;
    }
    return rc;
}

}  // extern "C"
