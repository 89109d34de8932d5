// Runs the same matrices through smf's own C API (solver/src/smf_c.cpp,
// CPU, compiled from the reference sources) and through the GPU adapter's
// (smf_gpu_c.cpp): inertia, solution and backward error of each. Prints
// one JSON object per case and exits 0 when every check holds:
//   - for the KKTs the same inertia as smf's, Sylvester's (n_x, m, 0); for
//     the Cholesky factor (n, 0, 0);
//   - backward error eta = ||A x - b|| / (||A|| ||x|| + ||b||) (inf-norms,
//     long double) <= 1e-14 on the GPU path;
//   - GPU and CPU solutions within 1e-8 of each other (relative), and a
//     refactorization on the adapter's cached plan as accurate, its
//     solution within 1e-10 of the first.
#include "smf_gpu_c.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

namespace {

struct Csc {
    int n = 0;
    std::vector<int> cp, ri;
    std::vector<double> v;
};

// lower CSC from (row, col, value) triplets with row >= col, sorted rows per
// column, duplicates summed
Csc from_triplets(int n, std::vector<std::vector<std::pair<int, double>>> &cols) {
    Csc A;
    A.n = n;
    A.cp.assign(n + 1, 0);
    for (int j = 0; j < n; ++j) {
        auto &c = cols[j];
        std::sort(c.begin(), c.end(), [](auto &a, auto &b) { return a.first < b.first; });
        int last = -1;
        for (auto &e : c) {
            if (e.first == last) {
                A.v.back() += e.second;
            } else {
                A.ri.push_back(e.first);
                A.v.push_back(e.second);
                last = e.first;
            }
        }
        A.cp[j + 1] = static_cast<int>(A.ri.size());
    }
    return A;
}

// configuration 1: the 7-point Laplacian on an m^3 grid (diagonal 6,
// off-diagonal -1)
Csc laplace7(int m) {
    const int n = m * m * m;
    std::vector<std::vector<std::pair<int, double>>> cols(n);
    auto id = [m](int x, int y, int z) { return (z * m + y) * m + x; };
    for (int z = 0; z < m; ++z)
        for (int y = 0; y < m; ++y)
            for (int x = 0; x < m; ++x) {
                const int i = id(x, y, z);
                cols[i].push_back({i, 6.0});
                if (x + 1 < m) cols[i].push_back({id(x + 1, y, z), -1.0});
                if (y + 1 < m) cols[i].push_back({id(x, y + 1, z), -1.0});
                if (z + 1 < m) cols[i].push_back({id(x, y, z + 1), -1.0});
            }
    return from_triplets(n, cols);
}

// configuration 2's generator at a smaller size: [[H + Sigma, A^T], [A,
// -1e-8 I]], H the 5-point Laplacian of a g x g grid, Sigma log-uniform in
// [1e-6, 1e6], A with 8 entries N(0, 1) per row at uniform columns
Csc kkt(int g, int m, unsigned seed) {
    const int nx = g * g, n = nx + m;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> ue(-6.0, 6.0);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::uniform_int_distribution<int> uc(0, nx - 1);
    std::vector<std::vector<std::pair<int, double>>> cols(n);
    for (int y = 0; y < g; ++y)
        for (int x = 0; x < g; ++x) {
            const int i = y * g + x;
            cols[i].push_back({i, 4.0 + std::pow(10.0, ue(rng))});
            if (x + 1 < g) cols[i].push_back({i + 1, -1.0});
            if (y + 1 < g) cols[i].push_back({i + g, -1.0});
        }
    for (int r = 0; r < m; ++r) {
        for (int t = 0; t < 8; ++t) cols[uc(rng)].push_back({nx + r, nd(rng)});
        cols[nx + r].push_back({nx + r, -1e-8});
    }
    return from_triplets(n, cols);
}

std::vector<double> matvec(const Csc &A, const std::vector<double> &x, int nrhs) {
    std::vector<double> y(static_cast<size_t>(A.n) * nrhs, 0.0);
    for (int r = 0; r < nrhs; ++r)
        for (int j = 0; j < A.n; ++j)
            for (int k = A.cp[j]; k < A.cp[j + 1]; ++k) {
                const int i = A.ri[k];
                y[i + static_cast<size_t>(r) * A.n] += A.v[k] * x[j + static_cast<size_t>(r) * A.n];
                if (i != j) y[j + static_cast<size_t>(r) * A.n] += A.v[k] * x[i + static_cast<size_t>(r) * A.n];
            }
    return y;
}

// the largest column eta of the block
double eta(const Csc &A, const std::vector<double> &x, const std::vector<double> &b, int nrhs) {
    const int n = A.n;
    std::vector<long double> rowsum(n, 0.0L);
    for (int j = 0; j < n; ++j)
        for (int k = A.cp[j]; k < A.cp[j + 1]; ++k) {
            rowsum[A.ri[k]] += std::fabs(static_cast<long double>(A.v[k]));
            if (A.ri[k] != j) rowsum[j] += std::fabs(static_cast<long double>(A.v[k]));
        }
    long double anorm = 0.0L;
    for (long double s : rowsum) anorm = std::max(anorm, s);
    double worst = 0.0;
    for (int r = 0; r < nrhs; ++r) {
        std::vector<long double> res(n, 0.0L);
        const size_t o = static_cast<size_t>(r) * n;
        for (int j = 0; j < n; ++j)
            for (int k = A.cp[j]; k < A.cp[j + 1]; ++k) {
                const int i = A.ri[k];
                res[i] += static_cast<long double>(A.v[k]) * x[o + j];
                if (i != j) res[j] += static_cast<long double>(A.v[k]) * x[o + i];
            }
        long double rn = 0.0L, xn = 0.0L, bn = 0.0L;
        for (int i = 0; i < n; ++i) {
            rn = std::max(rn, std::fabs(res[i] - b[o + i]));
            xn = std::max(xn, std::fabs(static_cast<long double>(x[o + i])));
            bn = std::max(bn, std::fabs(static_cast<long double>(b[o + i])));
        }
        worst = std::max(worst, static_cast<double>(rn / (anorm * xn + bn)));
    }
    return worst;
}

bool run_case(const char *name, const Csc &A, int mtype, int nrhs, int exp_pos, int exp_neg) {
    smf_csc_t csc{A.n, A.cp.data(), A.ri.data(), A.v.data()};
    std::mt19937_64 rng(0);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<double> xs(static_cast<size_t>(A.n) * nrhs);
    for (double &t : xs) t = u(rng);
    const std::vector<double> b = matvec(A, xs, nrhs);
    // smf's C API on the CPU
    smf_analysis_t ca = nullptr;
    smf_factor_t cf = nullptr;
    smf_inertia_t ci{};
    std::vector<double> xc = b;
    int rc = smf_analyse(&csc, &ca);
    if (rc == SMF_OK) rc = smf_factor(ca, A.v.data(), mtype, &cf);
    if (rc == SMF_OK) rc = smf_solve(ca, cf, nrhs, xc.data(), SMF_SOLVE_FULL);
    if (rc == SMF_OK) rc = smf_inertia(cf, &ci);
    const int rc_cpu = rc;
    // the same calls through the GPU adapter
    smf_gpu_analysis_t ga = nullptr;
    smf_gpu_factor_t gf = nullptr;
    smf_inertia_t gi{};
    std::vector<double> xg = b;
    rc = smf_gpu_analyse(&csc, &ga);
    if (rc == SMF_OK) rc = smf_gpu_factor(ga, A.v.data(), mtype, &gf);
    if (rc == SMF_OK) rc = smf_gpu_solve(ga, gf, nrhs, xg.data(), SMF_SOLVE_FULL);
    if (rc == SMF_OK) rc = smf_gpu_inertia(gf, &gi);
    const int rc_gpu = rc;
    // a second factorization on the same adapter plan (the IPM pattern)
    int rc_re = SMF_OK;
    double diff_re = 0.0, e_re = 1.0;
    if (rc_gpu == SMF_OK) {
        smf_gpu_factor_t gf2 = nullptr;
        std::vector<double> x2 = b;
        rc_re = smf_gpu_factor(ga, A.v.data(), mtype, &gf2);
        if (rc_re == SMF_OK) rc_re = smf_gpu_solve(ga, gf2, nrhs, x2.data(), SMF_SOLVE_FULL);
        for (size_t i = 0; i < x2.size(); ++i) diff_re = std::max(diff_re, std::fabs(x2[i] - xg[i]));
        if (rc_re == SMF_OK) e_re = eta(A, x2, b, nrhs);
        smf_gpu_free_factor(&gf2);
    }
    double e_cpu = rc_cpu == SMF_OK ? eta(A, xc, b, nrhs) : 1.0;
    double e_gpu = rc_gpu == SMF_OK ? eta(A, xg, b, nrhs) : 1.0;
    double dx = 0.0, xn = 0.0;
    for (size_t i = 0; i < xg.size(); ++i) {
        dx = std::max(dx, std::fabs(xg[i] - xc[i]));
        xn = std::max(xn, std::fabs(xc[i]));
    }
    const double rel = xn > 0.0 ? dx / xn : dx;
    bool ok = rc_cpu == SMF_OK && rc_gpu == SMF_OK && rc_re == SMF_OK;
    // smf_factor reports no counts for a Cholesky factor (factor_posdef
    // leaves Info's inertia at zero); the adapter's are the pivots' signs
    if (mtype == SMF_INDEF) ok = ok && ci.pos == gi.pos && ci.neg == gi.neg && ci.zero == gi.zero;
    ok = ok && gi.pos == exp_pos && gi.neg == exp_neg && gi.zero == 0;
    // a refactorization on the cached plan: as accurate, and the same
    // solution to rounding (atomic row updates may add in another order)
    const double rel_re = xn > 0.0 ? diff_re / xn : diff_re;
    ok = ok && e_gpu <= 1e-14 && rel <= 1e-8 && e_re <= 1e-14 && rel_re <= 1e-10;
    std::printf("{\"case\": \"%s\", \"n\": %d, \"nrhs\": %d, \"rc_cpu\": %d, \"rc_gpu\": %d, "
                "\"inertia_cpu\": [%d, %d, %d], \"inertia_gpu\": [%d, %d, %d], \"eta_cpu\": %.3e, "
                "\"eta_gpu\": %.3e, \"rel_diff\": %.3e, \"refactor_eta\": %.3e, \"refactor_rel_diff\": %.3e, "
                "\"ok\": %s}\n",
                name, A.n, nrhs, rc_cpu, rc_gpu, ci.pos, ci.neg, ci.zero, gi.pos, gi.neg, gi.zero, e_cpu, e_gpu,
                rel, e_re, rel_re, ok ? "true" : "false");
    std::fflush(stdout);
    smf_free_factor(&cf);
    smf_free_analysis(&ca);
    smf_gpu_free_factor(&gf);
    smf_gpu_free_analysis(&ga);
    return ok;
}

}  // namespace

int main() {
    bool ok = true;
    const Csc L = laplace7(20);  // configuration 1
    ok = run_case("laplace7_20", L, SMF_POSDEF, 1, L.n, 0) && ok;
    const Csc K1 = kkt(30, 180, 1);
    ok = run_case("kkt_30", K1, SMF_INDEF, 1, 900, 180) && ok;
    const Csc K2 = kkt(60, 1152, 2);
    ok = run_case("kkt_60_3rhs", K2, SMF_INDEF, 3, 3600, 1152) && ok;
    return ok ? 0 : 1;
}
