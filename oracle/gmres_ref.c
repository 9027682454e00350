/*
 * CPU restatement of the reference's linear solve — TEST INFRASTRUCTURE ONLY
 * (checker and timed CPU baseline; never part of the product path).
 *
 * LinearGmres::solve_matrix (src/linear/linear_gmres.rs:76-84) runs
 * faer_gmres::restarted_gmres(A, b, x, restart, max_outer, tol, Jacobi) in FP64
 * (faer_gmres 0.5.0, not vendored; its published algorithm is restarted GMRES
 * with a Jacobi preconditioner).  This is that algorithm in plain C, single
 * threaded like every reference example (LinearGmres::new(1), quickstart.rs:109):
 * CSR SpMV, right Jacobi preconditioning, modified Gram-Schmidt Arnoldi, Givens
 * rotations, restart after m steps, stop on ||b - A x|| / ||b|| <= tol.
 *
 * Pinning: test_oracle_cpu.py checks it against scipy SuperLU on oracle-assembled
 * systems and checks the returned residual against an independent recomputation.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Loops below run on OpenMP threads inside orc_gmres_threads (omp_set_num_threads) and on
 * one thread in orc_gmres, whose arithmetic stays sequential. */
static void spmv(int64_t n, const int64_t* rp, const int64_t* col, const double* val, const double* x, double* y) {
#pragma omp parallel for schedule(static) if (n > 4096)
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * x[col[k]];
        y[i] = s;
    }
}

static double nrm2(int64_t n, const double* x) {
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s) if (n > 4096)
    for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
    return sqrt(s);
}

/* One thread unless called through orc_gmres_threads. */
__attribute__((constructor)) static void orc_single_thread(void) { omp_set_num_threads(1); }

/* Returns 0 on convergence, 1 if not converged within max_outer cycles or
 * max_steps Arnoldi steps (max_steps <= 0: unlimited).  iters: Arnoldi steps done;
 * rel: final true relative residual. */
int orc_gmres(int64_t n, const int64_t* rp, const int64_t* col, const double* val, const double* b, double* x,
              int m, int max_outer, double tol, int64_t max_steps, int64_t* iters, double* rel) {
    double* dinv = malloc(n * sizeof(double));
    double* r = malloc(n * sizeof(double));
    double* w = malloc(n * sizeof(double));
    double* z = malloc(n * sizeof(double));
    double* V = malloc((size_t)n * (m + 1) * sizeof(double));
    double* H = calloc((size_t)(m + 1) * m, sizeof(double));
    double* cs = malloc(m * sizeof(double));
    double* sn = malloc(m * sizeof(double));
    double* g = malloc((m + 1) * sizeof(double));
    double* y = malloc(m * sizeof(double));
    int status = 1;
    int64_t steps = 0;
    for (int64_t i = 0; i < n; ++i) {
        double d = 0.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
            if (col[k] == i) d += val[k];
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
    const double bn = nrm2(n, b);
    double relr = 0.0;
    for (int outer = 0; outer <= max_outer; ++outer) {
        spmv(n, rp, col, val, x, r);
#pragma omp parallel for schedule(static) if (n > 4096)
        for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
        double beta = nrm2(n, r);
        relr = bn > 0.0 ? beta / bn : beta;
        if (relr <= tol || beta == 0.0) { status = 0; break; }
        if (outer == max_outer || (max_steps > 0 && steps >= max_steps)) break;
#pragma omp parallel for schedule(static) if (n > 4096)
        for (int64_t i = 0; i < n; ++i) V[i] = r[i] / beta;
        for (int k = 0; k <= m; ++k) g[k] = 0.0;
        g[0] = beta;
        int k = 0;
        for (int j = 0; j < m; ++j) {
            double* vj = V + (size_t)j * n;
#pragma omp parallel for schedule(static) if (n > 4096)
            for (int64_t i = 0; i < n; ++i) z[i] = dinv[i] * vj[i];
            spmv(n, rp, col, val, z, w);
            for (int l = 0; l <= j; ++l) {  /* modified Gram-Schmidt */
                const double* vl = V + (size_t)l * n;
                double h = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : h) if (n > 4096)
                for (int64_t i = 0; i < n; ++i) h += w[i] * vl[i];
#pragma omp parallel for schedule(static) if (n > 4096)
                for (int64_t i = 0; i < n; ++i) w[i] -= h * vl[i];
                H[(size_t)j * (m + 1) + l] = h;
            }
            double hn = nrm2(n, w);
            H[(size_t)j * (m + 1) + j + 1] = hn;
            if (hn > 0.0) {
#pragma omp parallel for schedule(static) if (n > 4096)
                for (int64_t i = 0; i < n; ++i) V[(size_t)(j + 1) * n + i] = w[i] / hn;
            }
            for (int l = 0; l < j; ++l) {
                double a = H[(size_t)j * (m + 1) + l], c = H[(size_t)j * (m + 1) + l + 1];
                H[(size_t)j * (m + 1) + l] = cs[l] * a + sn[l] * c;
                H[(size_t)j * (m + 1) + l + 1] = -sn[l] * a + cs[l] * c;
            }
            double a = H[(size_t)j * (m + 1) + j], c = H[(size_t)j * (m + 1) + j + 1];
            double rr = hypot(a, c);
            cs[j] = rr > 0.0 ? a / rr : 1.0;
            sn[j] = rr > 0.0 ? c / rr : 0.0;
            H[(size_t)j * (m + 1) + j] = rr;
            H[(size_t)j * (m + 1) + j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            k = j + 1;
            ++steps;
            if (fabs(g[j + 1]) <= 0.999 * tol * bn || hn == 0.0) break;
            if (max_steps > 0 && steps >= max_steps) break;
        }
        for (int i = k - 1; i >= 0; --i) {
            double s = g[i];
            for (int l = i + 1; l < k; ++l) s -= H[(size_t)l * (m + 1) + i] * y[l];
            y[i] = s / H[(size_t)i * (m + 1) + i];
        }
#pragma omp parallel for schedule(static) if (n > 4096)
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) s += V[(size_t)l * n + i] * y[l];
            x[i] += dinv[i] * s;
        }
    }
    *iters = steps;
    *rel = relr;
    free(dinv); free(r); free(w); free(z); free(V); free(H); free(cs); free(sn); free(g); free(y);
    return status;
}

/* The same solve with its vector loops on `nthreads` OpenMP threads (the reference's faer
 * solvers take a thread count, LinearGmres::new(nthreads), linear_gmres.rs:45); used only to
 * time the CPU baseline on all host cores.  orc_gmres itself runs on one thread. */
int orc_gmres_threads(int64_t n, const int64_t* rp, const int64_t* col, const double* val, const double* b, double* x,
                      int m, int max_outer, double tol, int64_t max_steps, int nthreads, int64_t* iters, double* rel) {
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
    int st = orc_gmres(n, rp, col, val, b, x, m, max_outer, tol, max_steps, iters, rel);
    omp_set_num_threads(1);
    return st;
}
