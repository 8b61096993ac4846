/*
 * usct_oracle.c -- FP64 CPU restatement of the acoustic leapfrog solver and its
 * exact discrete-adjoint FWI gradient.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product path (openusct_amd) never calls it and has no CPU fallback.
 *
 * Semantics restated (reference file:line):
 *   laplacian      simulation/ringfwi/solver.py:45-62 (interior only, border 0,
 *                  summation order: centre term first, then axes 0..ndim-1,
 *                  "+1" neighbour before "-1")
 *   sponge         simulation/ringfwi/solver.py:22-42 (mask built by the caller),
 *                  applied as p_new *= sponge (fwi.py:68-69) and
 *                  lam_k *= sponge before the residual injection (fwi.py:313-314)
 *   forward        simulation/ringfwi/fwi.py:34-81   (lap[src] += w[n-1] before
 *                  p_new = 2 p_cur - p_prev + S*lap; rec[n] = p^n, rec[0] = 0)
 *   forward_fmc    simulation/ringfwi/fwi.py:102-119 (one shot per src element,
 *                  every element receives)
 *   l2 residual    simulation/ringfwi/fwi.py:155-157
 *   gcn residual   simulation/ringfwi/fwi.py:158-167
 *   gradient       simulation/ringfwi/fwi.py:266-332 (k = nt-1..1, residual at
 *                  k = 0 never injected, duplicate receivers accumulate,
 *                  d2p = U[k] - 2U[k-1] + (U[k-2] if k >= 2 else 0), g /= m)
 *   Born forward   simulation/ringfwi/fwi.py:351-378 (dp^n = 2dp^{n-1} - dp^{n-2}
 *                  + S*L(dp^{n-1}) - (v/m)(U^n - 2U^{n-1} + U^{n-2}), times the sponge;
 *                  traces dp^n at the receivers, row 0 = 0)
 *   J^T apply      simulation/ringfwi/fwi.py:381-403 (the gradient's adjoint sweep
 *                  with the Born traces as residual)
 *   GN H v         simulation/ringfwi/fwi.py:406-431 (sum over transmits)
 *   loop structure and OpenMP placement follow software/core/libuap.hpp:29-224
 *   (rows/planes split across threads, serial receiver/source updates).
 *
 * Floating-point order matches the Python reference expression by expression
 * (compile with -ffp-contract=off), so forward traces agree with it to the last
 * few ulps.  The misfit J uses Neumaier-compensated summation.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static int64_t cells_of(const int *dims, int ndim) {
    int64_t n = 1;
    for (int d = 0; d < ndim; ++d) n *= dims[d];
    return n;
}

/* Interior-only second-order Laplacian (solver.py:45-62). */
static void lap_fp64(const double *p, double *out, const int *dims, int ndim,
                     double inv_h2) {
    const int64_t N = cells_of(dims, ndim);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) out[i] = 0.0;
    if (ndim == 2) {
        const int ny = dims[0], nx = dims[1];
        #pragma omp parallel for schedule(static)
        for (int iy = 1; iy < ny - 1; ++iy) {
            for (int ix = 1; ix < nx - 1; ++ix) {
                const int64_t c = (int64_t)iy * nx + ix;
                double acc = -4.0 * p[c];
                acc = acc + p[c + nx];
                acc = acc + p[c - nx];
                acc = acc + p[c + 1];
                acc = acc + p[c - 1];
                out[c] = acc * inv_h2;
            }
        }
    } else {
        const int nz = dims[0], ny = dims[1], nx = dims[2];
        const int64_t sz = (int64_t)ny * nx;
        #pragma omp parallel for schedule(static)
        for (int iz = 1; iz < nz - 1; ++iz) {
            for (int iy = 1; iy < ny - 1; ++iy) {
                for (int ix = 1; ix < nx - 1; ++ix) {
                    const int64_t c = (int64_t)iz * sz + (int64_t)iy * nx + ix;
                    double acc = -6.0 * p[c];
                    acc = acc + p[c + sz];
                    acc = acc + p[c - sz];
                    acc = acc + p[c + nx];
                    acc = acc + p[c - nx];
                    acc = acc + p[c + 1];
                    acc = acc + p[c - 1];
                    out[c] = acc * inv_h2;
                }
            }
        }
    }
}

static int check_args(const int *dims, int ndim, int nt, const int64_t *lin, int n) {
    if (ndim != 2 && ndim != 3) return -1;
    for (int d = 0; d < ndim; ++d) if (dims[d] < 1) return -1;
    if (nt < 1) return -1;
    const int64_t N = cells_of(dims, ndim);
    for (int j = 0; j < n; ++j) if (lin[j] < 0 || lin[j] >= N) return -2;
    return 0;
}

/*
 * One transmit event (fwi.py:34-81) with a point source at src_lin.
 * rec  : nt * n_rec (rec[n*n_rec + j] = p^n[rec_lin[j]]), may be NULL
 * hist : nt * N     (hist[n*N + i] = p^n[i]),              may be NULL
 * sponge: N multiplicative mask or NULL.
 */
int oracle_forward(const double *m, const int *dims, int ndim, double h, double dt,
                   int nt, int64_t src_lin, const double *wavelet,
                   const int64_t *rec_lin, int n_rec, const double *sponge,
                   double *rec, double *hist) {
    if (check_args(dims, ndim, nt, rec_lin, n_rec)) return -1;
    const int64_t N = cells_of(dims, ndim);
    if (src_lin < 0 || src_lin >= N) return -2;
    const double inv_h2 = 1.0 / (h * h);
    const double dt2 = dt * dt;
    double *S = (double *)malloc(sizeof(double) * N);
    double *a = (double *)calloc(N, sizeof(double));
    double *b = (double *)calloc(N, sizeof(double));
    double *c = (double *)calloc(N, sizeof(double));
    double *lap = (double *)calloc(N, sizeof(double));
    if (!S || !a || !b || !c || !lap) {
        free(S); free(a); free(b); free(c); free(lap);
        return -3;
    }
    for (int64_t i = 0; i < N; ++i) S[i] = dt2 / m[i];
    double *p_prev = a, *p_cur = b, *p_new = c;
    if (hist) memset(hist, 0, sizeof(double) * N);
    if (rec) for (int j = 0; j < n_rec; ++j) rec[j] = 0.0;

    for (int n = 1; n < nt; ++n) {
        lap_fp64(p_cur, lap, dims, ndim, inv_h2);
        lap[src_lin] += wavelet[n - 1];
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            double v = 2.0 * p_cur[i] - p_prev[i] + S[i] * lap[i];
            if (sponge) v *= sponge[i];
            p_new[i] = v;
        }
        double *t = p_prev; p_prev = p_cur; p_cur = p_new; p_new = t;
        if (hist) {
            double *row = hist + (int64_t)n * N;
            #pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < N; ++i) row[i] = p_cur[i];
        }
        if (rec) {
            double *row = rec + (int64_t)n * n_rec;
            for (int j = 0; j < n_rec; ++j) row[j] = p_cur[rec_lin[j]];
        }
    }
    free(S); free(a); free(b); free(c); free(lap);
    return 0;
}

/* data: n_tx * nt * n_rec, fwi.py:102-119. */
int oracle_forward_fmc(const double *m, const int *dims, int ndim, double h, double dt,
                       int nt, const int64_t *tx_lin, int n_tx, const int64_t *rec_lin,
                       int n_rec, const double *wavelet, const double *sponge,
                       double *data) {
    for (int t = 0; t < n_tx; ++t) {
        int rc = oracle_forward(m, dims, ndim, h, dt, nt, tx_lin[t], wavelet, rec_lin,
                                n_rec, sponge, data + (int64_t)t * nt * n_rec, NULL);
        if (rc) return rc;
    }
    return 0;
}

/* Neumaier compensated accumulator. */
typedef struct { double s, c; } kacc;
static inline void kadd(kacc *a, double x) {
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}

/*
 * Misfit and adjoint source of one transmit's traces (nt x n_rec, time-major).
 * misfit_type 0 = l2 (fwi.py:155-157), 1 = gcn (fwi.py:158-167).
 */
static double adjoint_residual(const double *dsyn, const double *dobs, int nt, int n_rec,
                               int misfit_type, double *res) {
    const int64_t total = (int64_t)nt * n_rec;
    if (misfit_type == 0) {
        kacc J = {0.0, 0.0};
        for (int64_t k = 0; k < total; ++k) {
            res[k] = dsyn[k] - dobs[k];
            kadd(&J, res[k] * res[k]);
        }
        return 0.5 * (J.s + J.c);
    }
    const double eps = 1e-12;
    double *ns = (double *)calloc(n_rec, sizeof(double));
    double *no = (double *)calloc(n_rec, sizeof(double));
    double *cc = (double *)calloc(n_rec, sizeof(double));
    for (int k = 0; k < nt; ++k)
        for (int j = 0; j < n_rec; ++j) {
            const double s = dsyn[(int64_t)k * n_rec + j], o = dobs[(int64_t)k * n_rec + j];
            ns[j] += s * s;
            no[j] += o * o;
        }
    for (int j = 0; j < n_rec; ++j) {
        ns[j] = sqrt(ns[j]) + eps;
        no[j] = sqrt(no[j]) + eps;
    }
    for (int k = 0; k < nt; ++k)
        for (int j = 0; j < n_rec; ++j)
            cc[j] += (dsyn[(int64_t)k * n_rec + j] / ns[j]) * (dobs[(int64_t)k * n_rec + j] / no[j]);
    double J = 0.0;
    for (int j = 0; j < n_rec; ++j) J += 1.0 - cc[j];
    for (int k = 0; k < nt; ++k)
        for (int j = 0; j < n_rec; ++j) {
            const int64_t q = (int64_t)k * n_rec + j;
            res[q] = -(dobs[q] / no[j] - cc[j] * (dsyn[q] / ns[j])) / ns[j];
        }
    free(ns); free(no); free(cc);
    return J;
}

/*
 * Adjoint sweep of one transmit (fwi.py:310-326): lam^k for k = nt-1..1 with the
 * residual res (nt x n_rec) injected at the receivers, imaging
 * grad += lam^k (U^k - 2U^{k-1} + U^{k-2}).  Work arrays l1, l2, lk, tmp, lap: N.
 */
static void adjoint_image(const double *S, const double *U, const double *res, const int *dims,
                          int ndim, double inv_h2, int nt, const int64_t *rec_lin, int n_rec,
                          const double *sponge, double *l1, double *l2, double *lk, double *tmp,
                          double *lap, double *grad) {
    const int64_t N = cells_of(dims, ndim);
    memset(l1, 0, sizeof(double) * N);
    memset(l2, 0, sizeof(double) * N);
    for (int k = nt - 1; k >= 1; --k) {
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) tmp[i] = S[i] * l1[i];
        lap_fp64(tmp, lap, dims, ndim, inv_h2);
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            double v = 2.0 * l1[i] - l2[i] + lap[i];
            if (sponge) v *= sponge[i];
            lk[i] = v;
        }
        const double *rr = res + (int64_t)k * n_rec;
        for (int j = 0; j < n_rec; ++j) lk[rec_lin[j]] -= rr[j];
        const double *Uk = U + (int64_t)k * N;
        const double *Uk1 = U + (int64_t)(k - 1) * N;
        const double *Uk2 = k >= 2 ? U + (int64_t)(k - 2) * N : NULL;
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            double d2p = Uk[i] - 2.0 * Uk1[i] + (Uk2 ? Uk2[i] : 0.0);
            grad[i] += lk[i] * d2p;
        }
        double *t2 = l2; l2 = l1; l1 = lk; lk = t2;
    }
}

/*
 * Misfit and exact discrete-adjoint gradient w.r.t. m (fwi.py:266-332).
 * dobs: n_tx * nt * n_rec; grad: N; J returned through *J_out.
 */
int oracle_misfit_and_gradient(const double *m, const int *dims, int ndim, double h,
                               double dt, int nt, const int64_t *tx_lin, int n_tx,
                               const int64_t *rec_lin, int n_rec, const double *wavelet,
                               const double *dobs, int misfit_type, const double *sponge,
                               double *grad, double *J_out) {
    if (check_args(dims, ndim, nt, rec_lin, n_rec)) return -1;
    if (check_args(dims, ndim, nt, tx_lin, n_tx)) return -1;
    if (misfit_type != 0 && misfit_type != 1) return -4;
    const int64_t N = cells_of(dims, ndim);
    const double inv_h2 = 1.0 / (h * h);
    const double dt2 = dt * dt;
    double *S = (double *)malloc(sizeof(double) * N);
    double *U = (double *)malloc(sizeof(double) * N * (int64_t)nt);
    double *dsyn = (double *)malloc(sizeof(double) * (int64_t)nt * n_rec);
    double *res = (double *)malloc(sizeof(double) * (int64_t)nt * n_rec);
    double *l1 = (double *)malloc(sizeof(double) * N);
    double *l2 = (double *)malloc(sizeof(double) * N);
    double *lk = (double *)malloc(sizeof(double) * N);
    double *tmp = (double *)malloc(sizeof(double) * N);
    double *lap = (double *)malloc(sizeof(double) * N);
    if (!S || !U || !dsyn || !res || !l1 || !l2 || !lk || !tmp || !lap) {
        free(S); free(U); free(dsyn); free(res); free(l1); free(l2); free(lk);
        free(tmp); free(lap);
        return -3;
    }
    for (int64_t i = 0; i < N; ++i) S[i] = dt2 / m[i];
    for (int64_t i = 0; i < N; ++i) grad[i] = 0.0;
    kacc J = {0.0, 0.0};
    int rc = 0;
    for (int t = 0; t < n_tx && !rc; ++t) {
        rc = oracle_forward(m, dims, ndim, h, dt, nt, tx_lin[t], wavelet, rec_lin, n_rec,
                            sponge, dsyn, U);
        if (rc) break;
        kadd(&J, adjoint_residual(dsyn, dobs + (int64_t)t * nt * n_rec, nt, n_rec,
                                  misfit_type, res));
        adjoint_image(S, U, res, dims, ndim, inv_h2, nt, rec_lin, n_rec, sponge, l1, l2, lk, tmp,
                      lap, grad);
    }
    if (!rc) for (int64_t i = 0; i < N; ++i) grad[i] /= m[i];
    *J_out = J.s + J.c;
    free(S); free(U); free(dsyn); free(res); free(l1); free(l2); free(lk); free(tmp);
    free(lap);
    return rc;
}

/*
 * Born forward of one transmit (fwi.py:351-378) given its background history U
 * (nt x N): dp^n = 2dp^{n-1} - dp^{n-2} + S*L(dp^{n-1}) - vm (U^n - 2U^{n-1} + U^{n-2}),
 * times the sponge; dd[n*n_rec + j] = dp^n at receiver j (row 0 = 0).
 * Work arrays a, b, c, lap: N.
 */
static void born_traces(const double *S, const double *vm, const double *U, const int *dims,
                        int ndim, double inv_h2, int nt, const int64_t *rec_lin, int n_rec,
                        const double *sponge, double *a, double *b, double *c, double *lap,
                        double *dd) {
    const int64_t N = cells_of(dims, ndim);
    double *dp_prev = a, *dp_cur = b, *dp_new = c;
    memset(dp_prev, 0, sizeof(double) * N);
    memset(dp_cur, 0, sizeof(double) * N);
    for (int j = 0; j < n_rec; ++j) dd[j] = 0.0;
    for (int n = 1; n < nt; ++n) {
        lap_fp64(dp_cur, lap, dims, ndim, inv_h2);
        const double *Un = U + (int64_t)n * N;
        const double *Un1 = U + (int64_t)(n - 1) * N;
        const double *Un2 = n >= 2 ? U + (int64_t)(n - 2) * N : NULL;
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            const double d2U = Un[i] - 2.0 * Un1[i] + (Un2 ? Un2[i] : 0.0);
            double w = 2.0 * dp_cur[i] - dp_prev[i] + S[i] * lap[i] - vm[i] * d2U;
            if (sponge) w *= sponge[i];
            dp_new[i] = w;
        }
        double *t3 = dp_prev; dp_prev = dp_cur; dp_cur = dp_new; dp_new = t3;
        double *row = dd + (int64_t)n * n_rec;
        for (int j = 0; j < n_rec; ++j) row[j] = dp_cur[rec_lin[j]];
    }
}

/*
 * Gauss-Newton Hessian-vector product (J^T J) v, summed over transmits
 * (fwi.py:406-431): per transmit the background forward with history, the Born
 * forward of v (fwi.py:351-378) and the adjoint sweep with the Born traces as
 * residual (fwi.py:381-403); hv = sum / m.
 */
int oracle_gn_hessian_vector(const double *m, const int *dims, int ndim, double h, double dt,
                             int nt, const int64_t *tx_lin, int n_tx, const int64_t *rec_lin,
                             int n_rec, const double *wavelet, const double *v,
                             const double *sponge, double *hv) {
    if (check_args(dims, ndim, nt, rec_lin, n_rec)) return -1;
    if (check_args(dims, ndim, nt, tx_lin, n_tx)) return -1;
    const int64_t N = cells_of(dims, ndim);
    const double inv_h2 = 1.0 / (h * h);
    const double dt2 = dt * dt;
    double *S = (double *)malloc(sizeof(double) * N);
    double *vm = (double *)malloc(sizeof(double) * N);
    double *U = (double *)malloc(sizeof(double) * N * (int64_t)nt);
    double *dd = (double *)calloc((size_t)nt * n_rec, sizeof(double));
    double *a = (double *)malloc(sizeof(double) * N);
    double *b = (double *)malloc(sizeof(double) * N);
    double *c = (double *)malloc(sizeof(double) * N);
    double *l1 = (double *)malloc(sizeof(double) * N);
    double *l2 = (double *)malloc(sizeof(double) * N);
    double *lk = (double *)malloc(sizeof(double) * N);
    double *tmp = (double *)malloc(sizeof(double) * N);
    double *lap = (double *)malloc(sizeof(double) * N);
    if (!S || !vm || !U || !dd || !a || !b || !c || !l1 || !l2 || !lk || !tmp || !lap) {
        free(S); free(vm); free(U); free(dd); free(a); free(b); free(c); free(l1); free(l2);
        free(lk); free(tmp); free(lap);
        return -3;
    }
    for (int64_t i = 0; i < N; ++i) {
        S[i] = dt2 / m[i];
        vm[i] = v[i] / m[i];
        hv[i] = 0.0;
    }
    int rc = 0;
    for (int t = 0; t < n_tx && !rc; ++t) {
        rc = oracle_forward(m, dims, ndim, h, dt, nt, tx_lin[t], wavelet, NULL, 0, sponge, NULL,
                            U);
        if (rc) break;
        born_traces(S, vm, U, dims, ndim, inv_h2, nt, rec_lin, n_rec, sponge, a, b, c, lap, dd);
        adjoint_image(S, U, dd, dims, ndim, inv_h2, nt, rec_lin, n_rec, sponge, l1, l2, lk, tmp,
                      lap, hv);
    }
    if (!rc) for (int64_t i = 0; i < N; ++i) hv[i] /= m[i];
    free(S); free(vm); free(U); free(dd); free(a); free(b); free(c); free(l1); free(l2); free(lk);
    free(tmp); free(lap);
    return rc;
}

/* Born traces J v of every transmit (fwi.py:351-378); dd: n_tx * nt * n_rec. */
int oracle_born_fmc(const double *m, const int *dims, int ndim, double h, double dt, int nt,
                    const int64_t *tx_lin, int n_tx, const int64_t *rec_lin, int n_rec,
                    const double *wavelet, const double *v, const double *sponge, double *dd) {
    if (check_args(dims, ndim, nt, rec_lin, n_rec)) return -1;
    if (check_args(dims, ndim, nt, tx_lin, n_tx)) return -1;
    const int64_t N = cells_of(dims, ndim);
    const double inv_h2 = 1.0 / (h * h);
    const double dt2 = dt * dt;
    double *S = (double *)malloc(sizeof(double) * N);
    double *vm = (double *)malloc(sizeof(double) * N);
    double *U = (double *)malloc(sizeof(double) * N * (int64_t)nt);
    double *a = (double *)malloc(sizeof(double) * N);
    double *b = (double *)malloc(sizeof(double) * N);
    double *c = (double *)malloc(sizeof(double) * N);
    double *lap = (double *)malloc(sizeof(double) * N);
    if (!S || !vm || !U || !a || !b || !c || !lap) {
        free(S); free(vm); free(U); free(a); free(b); free(c); free(lap);
        return -3;
    }
    for (int64_t i = 0; i < N; ++i) {
        S[i] = dt2 / m[i];
        vm[i] = v[i] / m[i];
    }
    int rc = 0;
    for (int t = 0; t < n_tx && !rc; ++t) {
        rc = oracle_forward(m, dims, ndim, h, dt, nt, tx_lin[t], wavelet, NULL, 0, sponge, NULL,
                            U);
        if (!rc)
            born_traces(S, vm, U, dims, ndim, inv_h2, nt, rec_lin, n_rec, sponge, a, b, c, lap,
                        dd + (int64_t)t * nt * n_rec);
    }
    free(S); free(vm); free(U); free(a); free(b); free(c); free(lap);
    return rc;
}

/* Full pressure history of one transmit (fwi.py:92-99); hist: nt * N. */
int oracle_simulate_wavefield(const double *m, const int *dims, int ndim, double h,
                              double dt, int nt, int64_t src_lin, const double *wavelet,
                              const double *sponge, double *hist) {
    return oracle_forward(m, dims, ndim, h, dt, nt, src_lin, wavelet, NULL, 0, sponge,
                          NULL, hist);
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
