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
 *   footprints     simulation/ringfwi/fwi.py:74-76 (receiver group = weighted sum in
 *                  footprint order), :84-89 (transmit drives each cell with
 *                  wavelet * w_i), :316-319 (adjoint spreads res * w_i);
 *                  groups from geometry.py:212-278 (built by the caller)
 *   time windows   simulation/ringfwi/fwi.py:255-262 (misfit of W*dsyn vs W*dobs),
 *                  :298-301 (adjoint source times W again)
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
 * Elements as groups of grid cells with weights, in CSR form: group g covers
 * cells[off[g] .. off[g+1]) with weights w[...] (geometry.py:212-278 footprints;
 * a point element is one cell of weight 1).  On transmit every cell of the group
 * is driven with wavelet * w_i (fwi.py:84-89); on receive a group records
 * sum_i w_i p[cell_i], summed in footprint order from 0 (fwi.py:74-76), and the
 * adjoint subtracts res * w_i at each cell (fwi.py:316-319).  For point elements
 * these reduce exactly to the point forms (x * 1 = x, 0 + x = x).
 */
typedef struct {
    int n;
    const int64_t *off, *cells;
    const double *w;
} groups;

static int check_groups(const groups *g, int64_t N) {
    if (g->n < 0 || g->off[0] != 0) return -2;
    for (int j = 0; j < g->n; ++j)
        if (g->off[j + 1] < g->off[j]) return -2;
    for (int64_t i = 0; i < g->off[g->n]; ++i)
        if (g->cells[i] < 0 || g->cells[i] >= N) return -2;
    return 0;
}

/* Point groups over lin[0..n): off (n+1) and w (n) are allocated here. */
static int point_groups(const int64_t *lin, int n, groups *g, int64_t **off, double **w) {
    *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    *w = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!*off || !*w) {
        free(*off); free(*w);
        return -3;
    }
    for (int j = 0; j <= n; ++j) (*off)[j] = j;
    for (int j = 0; j < n; ++j) (*w)[j] = 1.0;
    g->n = n;
    g->off = *off;
    g->cells = lin;
    g->w = *w;
    return 0;
}

/* Samples of every group at one level: row[j] = sum_i w_i p[cell_i]. */
static void sample_groups(const groups *rec, const double *p, double *row) {
    for (int j = 0; j < rec->n; ++j) {
        double acc = 0.0;
        for (int64_t i = rec->off[j]; i < rec->off[j + 1]; ++i) acc = acc + rec->w[i] * p[rec->cells[i]];
        row[j] = acc;
    }
}

/*
 * One transmit event (fwi.py:34-81) driven at src_cells with weights src_w.
 * rec  : nt * rec->n (row 0 = 0), may be NULL (with rec == NULL)
 * hist : nt * N     (hist[n*N + i] = p^n[i]), may be NULL
 * sponge: N multiplicative mask or NULL.
 */
static int forward_g(const double *m, const int *dims, int ndim, double h, double dt, int nt,
                     const int64_t *src_cells, const double *src_w, int64_t n_src,
                     const double *wavelet, const groups *recg, const double *sponge,
                     double *rec, double *hist) {
    const int64_t N = cells_of(dims, ndim);
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
    if (rec) for (int j = 0; j < recg->n; ++j) rec[j] = 0.0;

    for (int n = 1; n < nt; ++n) {
        lap_fp64(p_cur, lap, dims, ndim, inv_h2);
        for (int64_t i = 0; i < n_src; ++i) lap[src_cells[i]] += wavelet[n - 1] * src_w[i];
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
        if (rec) sample_groups(recg, p_cur, rec + (int64_t)n * recg->n);
    }
    free(S); free(a); free(b); free(c); free(lap);
    return 0;
}

/* Point-source forward with point receivers (kept for the simulate_wavefield path). */
int oracle_forward(const double *m, const int *dims, int ndim, double h, double dt,
                   int nt, int64_t src_lin, const double *wavelet,
                   const int64_t *rec_lin, int n_rec, const double *sponge,
                   double *rec, double *hist) {
    if (check_args(dims, ndim, nt, rec_lin, n_rec)) return -1;
    const int64_t N = cells_of(dims, ndim);
    if (src_lin < 0 || src_lin >= N) return -2;
    groups g;
    int64_t *off;
    double *w;
    if (point_groups(rec_lin, n_rec, &g, &off, &w)) return -3;
    const double one = 1.0;
    int rc = forward_g(m, dims, ndim, h, dt, nt, &src_lin, &one, 1, wavelet, &g, sponge,
                       rec, hist);
    free(off); free(w);
    return rc;
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
 * Windowed misfit of one transmit (fwi.py:298-301): the residual of W*dsyn
 * against W*dobs, times W again (the chain rule through the window).  W = NULL
 * is the unwindowed misfit.  ws, wo: work arrays of nt*n_rec.
 */
static double windowed_residual(const double *dsyn, const double *dobs, const double *W, int nt,
                                int n_rec, int misfit_type, double *res, double *ws, double *wo) {
    if (!W) return adjoint_residual(dsyn, dobs, nt, n_rec, misfit_type, res);
    const int64_t total = (int64_t)nt * n_rec;
    for (int64_t q = 0; q < total; ++q) {
        ws[q] = dsyn[q] * W[q];
        wo[q] = dobs[q] * W[q];
    }
    const double J = adjoint_residual(ws, wo, nt, n_rec, misfit_type, res);
    for (int64_t q = 0; q < total; ++q) res[q] = res[q] * W[q];
    return J;
}

/*
 * Adjoint sweep of one transmit (fwi.py:303-326): lam^k for k = nt-1..1 with the
 * residual res (nt x rec->n) injected over the receiver groups, imaging
 * grad += lam^k (U^k - 2U^{k-1} + U^{k-2}).  Work arrays l1, l2, lk, tmp, lap: N.
 */
static void adjoint_image(const double *S, const double *U, const double *res, const int *dims,
                          int ndim, double inv_h2, int nt, const groups *rec,
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
        const double *rr = res + (int64_t)k * rec->n;
        for (int j = 0; j < rec->n; ++j)
            for (int64_t i = rec->off[j]; i < rec->off[j + 1]; ++i)
                lk[rec->cells[i]] -= rr[j] * rec->w[i];
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

#define TX_ARGS(t) (tx->cells + tx->off[t]), (tx->w + tx->off[t]), (tx->off[(t) + 1] - tx->off[t])

/*
 * Misfit and exact discrete-adjoint gradient w.r.t. m (fwi.py:266-332) over
 * element groups, with optional time windows tw (n_tx * nt * n_rec, NULL = none).
 * grad = NULL computes the misfit only (fwi.py:246-263).
 */
static int misfit_gradient_g(const double *m, const int *dims, int ndim, double h, double dt,
                             int nt, const groups *tx, const groups *rec, const double *wavelet,
                             const double *dobs, const double *tw, int misfit_type,
                             const double *sponge, double *grad, double *J_out) {
    if (misfit_type != 0 && misfit_type != 1) return -4;
    const int64_t N = cells_of(dims, ndim);
    const int n_rec = rec->n;
    const int64_t T = (int64_t)nt * n_rec;
    const double inv_h2 = 1.0 / (h * h);
    const double dt2 = dt * dt;
    double *S = (double *)malloc(sizeof(double) * N);
    double *U = grad ? (double *)malloc(sizeof(double) * N * (int64_t)nt) : NULL;
    double *dsyn = (double *)malloc(sizeof(double) * T);
    double *res = (double *)malloc(sizeof(double) * T);
    double *ws = (double *)malloc(sizeof(double) * T);
    double *wo = (double *)malloc(sizeof(double) * T);
    double *l1 = (double *)malloc(sizeof(double) * N);
    double *l2 = (double *)malloc(sizeof(double) * N);
    double *lk = (double *)malloc(sizeof(double) * N);
    double *tmp = (double *)malloc(sizeof(double) * N);
    double *lap = (double *)malloc(sizeof(double) * N);
    if (!S || (grad && !U) || !dsyn || !res || !ws || !wo || !l1 || !l2 || !lk || !tmp || !lap) {
        free(S); free(U); free(dsyn); free(res); free(ws); free(wo); free(l1); free(l2);
        free(lk); free(tmp); free(lap);
        return -3;
    }
    for (int64_t i = 0; i < N; ++i) S[i] = dt2 / m[i];
    if (grad) for (int64_t i = 0; i < N; ++i) grad[i] = 0.0;
    kacc J = {0.0, 0.0};
    int rc = 0;
    for (int t = 0; t < tx->n && !rc; ++t) {
        rc = forward_g(m, dims, ndim, h, dt, nt, TX_ARGS(t), wavelet, rec, sponge, dsyn, U);
        if (rc) break;
        kadd(&J, windowed_residual(dsyn, dobs + (int64_t)t * T, tw ? tw + (int64_t)t * T : NULL,
                                   nt, n_rec, misfit_type, res, ws, wo));
        if (grad)
            adjoint_image(S, U, res, dims, ndim, inv_h2, nt, rec, sponge, l1, l2, lk, tmp, lap,
                          grad);
    }
    if (!rc && grad) for (int64_t i = 0; i < N; ++i) grad[i] /= m[i];
    *J_out = J.s + J.c;
    free(S); free(U); free(dsyn); free(res); free(ws); free(wo); free(l1); free(l2); free(lk);
    free(tmp); free(lap);
    return rc;
}

/*
 * Born forward of one transmit (fwi.py:350-378) given its background history U
 * (nt x N): dp^n = 2dp^{n-1} - dp^{n-2} + S*L(dp^{n-1}) - vm (U^n - 2U^{n-1} + U^{n-2}),
 * times the sponge; dd[n*rec->n + j] = receiver group j's sample of dp^n (row 0 = 0).
 * Work arrays a, b, c, lap: N.
 */
static void born_traces(const double *S, const double *vm, const double *U, const int *dims,
                        int ndim, double inv_h2, int nt, const groups *rec,
                        const double *sponge, double *a, double *b, double *c, double *lap,
                        double *dd) {
    const int64_t N = cells_of(dims, ndim);
    double *dp_prev = a, *dp_cur = b, *dp_new = c;
    memset(dp_prev, 0, sizeof(double) * N);
    memset(dp_cur, 0, sizeof(double) * N);
    for (int j = 0; j < rec->n; ++j) dd[j] = 0.0;
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
        sample_groups(rec, dp_cur, dd + (int64_t)n * rec->n);
    }
}

/*
 * Gauss-Newton Hessian-vector product (J^T J) v summed over transmits
 * (fwi.py:410-431) when hv != NULL, or the Born traces J v of every transmit
 * into dd_all (n_tx * nt * n_rec) when dd_all != NULL: per transmit the
 * background forward with history, the Born forward of v (fwi.py:350-378) and
 * the adjoint sweep with the Born traces as residual (fwi.py:382-407); hv = sum / m.
 */
static int gn_g(const double *m, const int *dims, int ndim, double h, double dt, int nt,
                const groups *tx, const groups *rec, const double *wavelet, const double *v,
                const double *sponge, double *hv, double *dd_all) {
    const int64_t N = cells_of(dims, ndim);
    const int n_rec = rec->n;
    const double inv_h2 = 1.0 / (h * h);
    const double dt2 = dt * dt;
    double *S = (double *)malloc(sizeof(double) * N);
    double *vm = (double *)malloc(sizeof(double) * N);
    double *U = (double *)malloc(sizeof(double) * N * (int64_t)nt);
    double *dd = (double *)calloc((size_t)nt * n_rec + 1, sizeof(double));
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
        if (hv) hv[i] = 0.0;
    }
    int rc = 0;
    for (int t = 0; t < tx->n && !rc; ++t) {
        rc = forward_g(m, dims, ndim, h, dt, nt, TX_ARGS(t), wavelet, rec, sponge, NULL, U);
        if (rc) break;
        double *out = dd_all ? dd_all + (int64_t)t * nt * n_rec : dd;
        born_traces(S, vm, U, dims, ndim, inv_h2, nt, rec, sponge, a, b, c, lap, out);
        if (hv)
            adjoint_image(S, U, out, dims, ndim, inv_h2, nt, rec, sponge, l1, l2, lk, tmp, lap,
                          hv);
    }
    if (!rc && hv) for (int64_t i = 0; i < N; ++i) hv[i] /= m[i];
    free(S); free(vm); free(U); free(dd); free(a); free(b); free(c); free(l1); free(l2); free(lk);
    free(tmp); free(lap);
    return rc;
}

/* Element groups from caller CSR arrays (checked against the grid). */
#define GROUP_PARAMS(p)                                                                   \
    int n_##p, const int64_t *p##_off, const int64_t *p##_cells, const double *p##_w
#define MAKE_GROUP(p) groups p = {n_##p, p##_off, p##_cells, p##_w}

static int check_grid(const int *dims, int ndim, int nt) {
    if (ndim != 2 && ndim != 3) return -1;
    for (int d = 0; d < ndim; ++d) if (dims[d] < 1) return -1;
    return nt < 1 ? -1 : 0;
}

/* data: n_tx * nt * n_rec (fwi.py:102-119) over element groups. */
int oracle_forward_fmc_g(const double *m, const int *dims, int ndim, double h, double dt, int nt,
                         GROUP_PARAMS(tx), GROUP_PARAMS(rec), const double *wavelet,
                         const double *sponge, double *data) {
    MAKE_GROUP(tx);
    MAKE_GROUP(rec);
    if (check_grid(dims, ndim, nt)) return -1;
    const int64_t N = cells_of(dims, ndim);
    if (check_groups(&tx, N) || check_groups(&rec, N)) return -2;
    for (int t = 0; t < tx.n; ++t) {
        int rc = forward_g(m, dims, ndim, h, dt, nt, tx.cells + tx.off[t], tx.w + tx.off[t],
                           tx.off[t + 1] - tx.off[t], wavelet, &rec, sponge,
                           data + (int64_t)t * nt * rec.n, NULL);
        if (rc) return rc;
    }
    return 0;
}

/* Misfit (grad = NULL) or misfit and gradient over element groups, time windows tw. */
int oracle_misfit_and_gradient_g(const double *m, const int *dims, int ndim, double h, double dt,
                                 int nt, GROUP_PARAMS(tx), GROUP_PARAMS(rec),
                                 const double *wavelet, const double *dobs, const double *tw,
                                 int misfit_type, const double *sponge, double *grad,
                                 double *J_out) {
    MAKE_GROUP(tx);
    MAKE_GROUP(rec);
    if (check_grid(dims, ndim, nt)) return -1;
    const int64_t N = cells_of(dims, ndim);
    if (check_groups(&tx, N) || check_groups(&rec, N)) return -2;
    return misfit_gradient_g(m, dims, ndim, h, dt, nt, &tx, &rec, wavelet, dobs, tw, misfit_type,
                             sponge, grad, J_out);
}

/* GN Hessian-vector product (hv) or Born traces (dd) over element groups. */
int oracle_gn_g(const double *m, const int *dims, int ndim, double h, double dt, int nt,
                GROUP_PARAMS(tx), GROUP_PARAMS(rec), const double *wavelet, const double *v,
                const double *sponge, double *hv, double *dd) {
    MAKE_GROUP(tx);
    MAKE_GROUP(rec);
    if (check_grid(dims, ndim, nt)) return -1;
    const int64_t N = cells_of(dims, ndim);
    if (check_groups(&tx, N) || check_groups(&rec, N)) return -2;
    return gn_g(m, dims, ndim, h, dt, nt, &tx, &rec, wavelet, v, sponge, hv, dd);
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
