/* libusct from C, no Python: the problem of the reference's native example
 * (software/core/example_main.cpp:19-73) -- a 41 x 41 grid at 1 mm with a
 * 2600 m/s flaw in 3000 m/s, a 12-element ring of radius 16 mm, a 400 kHz
 * Ricker wavelet, nt = 300 -- through the C ABI of include/usct.h:
 * full-matrix-capture forward, then misfit and gradient from the uniform
 * starting model, printing the checksums the reference's example prints.
 *
 * Beyond the checksums it checks what the mathematics fixes:
 *   - reciprocity: data[i][t][j] = data[j][t][i] for point elements without a
 *     sponge (to rounding);
 *   - the observed data equal to the starting model's own traces give J = 0
 *     and a zero gradient, exactly;
 *   - the stateless call and the plan give the same traces, bit for bit;
 *   - an element outside the grid is refused with USCT_EINVAL and a message;
 *   - the libuap-typed entry points (host arrays, int cells) give the same
 *     traces bit for bit and the same misfit and gradient.
 * With an argument it also writes, as float64 in this order, the wavelet (nt),
 * the flaw model m (41 x 41), the element indices (12), the traces
 * (12 x nt x 12), the gradient (41 x 41) and J to that file, for a comparison
 * with the reference's core on the same inputs.
 *
 * Build (tests/test_c_example.py does this):
 *   gcc -std=c11 -O2 -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -Iinclude \
 *       tests/c_example/usct_example.c -Lopenusct_amd -lusct -L/opt/rocm/lib \
 *       -lamdhip64 -lm -Wl,-rpath,<repo>/openusct_amd -o usct_example
 * Exit status 0 = every check passed.
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "usct.h"

#define N_EL 12

static int failures = 0;

static void check(int ok, const char *what) {
    printf("%s %s\n", ok ? "ok  " : "FAIL", what);
    if (!ok) failures++;
}

#define HIPC(call)                                                              \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));          \
            return 2;                                                           \
        }                                                                       \
    } while (0)

#define USCT(call)                                                              \
    do {                                                                        \
        int s_ = (call);                                                        \
        if (s_ != USCT_OK) {                                                    \
            fprintf(stderr, "%s: status %d: %s\n", #call, s_, usct_last_error()); \
            return 3;                                                           \
        }                                                                       \
    } while (0)

int main(int argc, char **argv) {
    const double PI = 3.14159265358979323846;
    const int n = 41, nt = 300;
    const int dims[2] = {n, n};
    const double h = 1.0e-3, dt = 1.5e-7, f0 = 0.4e6;
    const size_t N = (size_t)n * n;
    const size_t ND = (size_t)N_EL * nt * N_EL;

    double *wavelet = malloc(sizeof(double) * nt), *m = malloc(sizeof(double) * N),
           *m0 = malloc(sizeof(double) * N), *data = malloc(sizeof(double) * ND),
           *data2 = malloc(sizeof(double) * ND), *grad = malloc(sizeof(double) * N);
    if (!wavelet || !m || !m0 || !data || !data2 || !grad) return 2;
    /* Ricker wavelet, t0 = 1/f0 (example_main.cpp:28-33) */
    for (int i = 0; i < nt; ++i) {
        const double t = i * dt - 1.0 / f0;
        const double a = (PI * f0 * t) * (PI * f0 * t);
        wavelet[i] = (1.0 - 2.0 * a) * exp(-a);
    }
    /* uniform 3000 m/s with a 4 mm, 2600 m/s flaw (example_main.cpp:35-44) */
    const double cx = (n - 1) * h / 2.0;
    for (int iy = 0; iy < n; ++iy)
        for (int ix = 0; ix < n; ++ix) {
            const double r = hypot(ix * h - 0.60 * (n - 1) * h, iy * h - 0.50 * (n - 1) * h);
            const double c = r <= 0.004 ? 2600.0 : 3000.0;
            m[(size_t)iy * n + ix] = 1.0 / (c * c);
            m0[(size_t)iy * n + ix] = 1.0 / (3000.0 * 3000.0);
        }
    /* 12 elements on a 16 mm ring, linear indices (example_main.cpp:46-53) */
    int64_t elem[N_EL];
    for (int k = 0; k < N_EL; ++k) {
        const double th = 2.0 * PI * k / N_EL;
        const long ix = lround((cx + 0.016 * cos(th)) / h);
        const long iy = lround((cx + 0.016 * sin(th)) / h);
        elem[k] = (int64_t)iy * n + ix;
    }

    double *d_m, *d_m0, *d_w, *d_data, *d_data0, *d_grad, *d_J;
    HIPC(hipMalloc((void **)&d_m, sizeof(double) * N));
    HIPC(hipMalloc((void **)&d_m0, sizeof(double) * N));
    HIPC(hipMalloc((void **)&d_w, sizeof(double) * nt));
    HIPC(hipMalloc((void **)&d_data, sizeof(double) * ND));
    HIPC(hipMalloc((void **)&d_data0, sizeof(double) * ND));
    HIPC(hipMalloc((void **)&d_grad, sizeof(double) * N));
    HIPC(hipMalloc((void **)&d_J, sizeof(double)));
    HIPC(hipMemcpy(d_m, m, sizeof(double) * N, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d_m0, m0, sizeof(double) * N, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d_w, wavelet, sizeof(double) * nt, hipMemcpyHostToDevice));

    /* forward, stateless form (uap::forward_fmc's argument order) */
    USCT(usct_forward_fmc_direct(d_m, dims, 2, h, dt, nt, elem, N_EL, elem, N_EL, d_w, d_data,
                                 USCT_METHOD_BATCHED, NULL));
    HIPC(hipDeviceSynchronize());
    HIPC(hipMemcpy(data, d_data, sizeof(double) * ND, hipMemcpyDeviceToHost));
    double s = 0.0, amax = 0.0;
    for (size_t i = 0; i < ND; ++i) {
        s += fabs(data[i]);
        if (fabs(data[i]) > amax) amax = fabs(data[i]);
    }
    printf("forward: sum|data| = %.6e\n", s);

    /* the same forward and the gradient through a plan */
    usct_plan *plan = NULL;
    USCT(usct_plan_create(&plan, 2, dims, h, dt, nt, elem, N_EL, elem, N_EL, USCT_METHOD_BATCHED,
                          0, 0, USCT_FLAG_AUTOTUNE));
    USCT(usct_forward_fmc(plan, d_m, d_w, NULL, d_data0, NULL));
    HIPC(hipDeviceSynchronize());
    HIPC(hipMemcpy(data2, d_data0, sizeof(double) * ND, hipMemcpyDeviceToHost));
    check(memcmp(data, data2, sizeof(double) * ND) == 0, "plan and stateless traces bit for bit");
    USCT(usct_misfit_and_gradient(plan, d_m0, d_w, NULL, d_data, USCT_MISFIT_L2, d_J, d_grad, NULL));
    HIPC(hipDeviceSynchronize());
    double J = 0.0;
    HIPC(hipMemcpy(&J, d_J, sizeof(double), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(grad, d_grad, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (argc > 1) {
        FILE *f = fopen(argv[1], "wb");
        double el[N_EL];
        for (int k = 0; k < N_EL; ++k) el[k] = (double)elem[k];
        if (!f || fwrite(wavelet, sizeof(double), nt, f) != (size_t)nt ||
            fwrite(m, sizeof(double), N, f) != N || fwrite(el, sizeof(double), N_EL, f) != N_EL ||
            fwrite(data, sizeof(double), ND, f) != ND || fwrite(grad, sizeof(double), N, f) != N ||
            fwrite(&J, sizeof(double), 1, f) != 1 || fclose(f) != 0) {
            fprintf(stderr, "cannot write %s\n", argv[1]);
            return 2;
        }
    }
    double gs = 0.0;
    int finite = 1;
    for (size_t i = 0; i < N; ++i) {
        gs += fabs(grad[i]);
        finite &= isfinite(grad[i]) != 0;
    }
    printf("misfit J = %.6e, sum|grad| = %.6e\n", J, gs);
    check(J > 0.0 && gs > 0.0 && finite, "misfit and gradient of the flaw data are positive and finite");

    /* the libuap interface itself: host arrays and int cells, exactly the
     * argument types of uap::forward_fmc / uap::misfit_and_gradient */
    {
        int elem32[N_EL];
        for (int k = 0; k < N_EL; ++k) elem32[k] = (int)elem[k];
        double *data3 = malloc(sizeof(double) * ND), *grad3 = malloc(sizeof(double) * N);
        if (!data3 || !grad3) return 2;
        USCT(usct_uap_forward_fmc(m, dims, 2, h, dt, nt, elem32, N_EL, elem32, N_EL, wavelet, data3));
        check(memcmp(data, data3, sizeof(double) * ND) == 0,
              "usct_uap_forward_fmc (host arrays, int cells) traces bit for bit");
        const double J3 = usct_uap_misfit_and_gradient(m0, dims, 2, h, dt, nt, elem32, N_EL, elem32,
                                                       N_EL, wavelet, data, grad3, 0);
        double dg = 0.0, ng = 0.0;
        for (size_t i = 0; i < N; ++i) {
            dg += (grad3[i] - grad[i]) * (grad3[i] - grad[i]);
            ng += grad[i] * grad[i];
        }
        printf("usct_uap_misfit_and_gradient: J = %.6e, |J - J_plan| / J = %.2e, gradient rel L2 "
               "%.2e\n", J3, fabs(J3 - J) / J, sqrt(dg / ng));
        check(fabs(J3 - J) <= 1e-9 * J && sqrt(dg) <= 1e-6 * sqrt(ng),
              "usct_uap_misfit_and_gradient equals the plan's misfit and gradient");
        int badc[2] = {elem32[0], (int)N};
        const double Jbad = usct_uap_misfit_and_gradient(m0, dims, 2, h, dt, nt, badc, 2, badc, 2,
                                                         wavelet, data, grad3, 0);
        check(isnan(Jbad) && strlen(usct_last_error()) > 0,
              "usct_uap_misfit_and_gradient returns NaN and a message for a cell outside the grid");
        free(data3);
        free(grad3);
    }

    /* reciprocity of the traces */
    double rmax = 0.0;
    for (int i = 0; i < N_EL; ++i)
        for (int j = 0; j < N_EL; ++j)
            for (int t = 0; t < nt; ++t) {
                const double a = data[((size_t)i * nt + t) * N_EL + j];
                const double b = data[((size_t)j * nt + t) * N_EL + i];
                if (fabs(a - b) > rmax) rmax = fabs(a - b);
            }
    printf("reciprocity: max |d_ij - d_ji| / max |d| = %.3e\n", rmax / amax);
    check(amax > 0.0 && rmax <= 1e-6 * amax, "reciprocity data[i][t][j] = data[j][t][i]");

    /* observed data = the starting model's own traces: J = 0, gradient 0 */
    USCT(usct_forward_fmc(plan, d_m0, d_w, NULL, d_data0, NULL));
    USCT(usct_misfit_and_gradient(plan, d_m0, d_w, NULL, d_data0, USCT_MISFIT_L2, d_J, d_grad, NULL));
    HIPC(hipDeviceSynchronize());
    HIPC(hipMemcpy(&J, d_J, sizeof(double), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(grad, d_grad, sizeof(double) * N, hipMemcpyDeviceToHost));
    int zero = J == 0.0;
    for (size_t i = 0; i < N; ++i) zero &= grad[i] == 0.0;
    check(zero, "own traces as data: J = 0 and gradient = 0");
    USCT(usct_plan_destroy(plan));

    /* an element outside the grid */
    int64_t bad[2] = {elem[0], (int64_t)N};
    usct_plan *p2 = NULL;
    const int rc = usct_plan_create(&p2, 2, dims, h, dt, nt, bad, 2, bad, 2, USCT_METHOD_BATCHED, 0,
                                    0, 0);
    printf("bad element: status %d, \"%s\"\n", rc, usct_last_error());
    check(rc == USCT_EINVAL && p2 == NULL && strlen(usct_last_error()) > 0,
          "an element outside the grid is refused");

    hipFree(d_m); hipFree(d_m0); hipFree(d_w); hipFree(d_data); hipFree(d_data0);
    hipFree(d_grad); hipFree(d_J);
    free(wavelet); free(m); free(m0); free(data); free(data2); free(grad);
    printf(failures ? "libusct native example FAILED\n" : "libusct native example OK\n");
    return failures ? 1 : 0;
}
