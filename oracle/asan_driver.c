/*
 * Sanitizer run of the CPU oracle (test infrastructure only): builds the
 * oracle in the same translation unit under AddressSanitizer and UBSan and
 * drives every exported entry point once on tiny grids -- surface, crosswell
 * and wells + surface receivers, orders 4 and 8, float and double -- so that
 * an out-of-bounds tap, record or residual index, or a leak, ends the run.
 * `make -C oracle asan-check` builds and runs it.
 */
#include <stdio.h>

#include "elastic_oracle.c"

static int run_case(int acq_type, int sdo)
{
    const int nz = 30, nx = 34, npml = 4, ns = 2, nt = 40, nr = 10;
    orc_grid g = {nz, nx, sdo, npml, 1.125, -0.041667, 0.0, 0.0, 4e-4, 5.0, 5.0};
    if (sdo == 4) { g.c1 = 1.195955; g.c2 = -0.079498; g.c3 = 0.009763; g.c4 = -0.000697; }
    const long n = (long)nz * nx;
    int srcx[2] = {8, 20}, srcz[2] = {6, 6};
    orc_acq q = {ns, nr, acq_type == 2 ? nr - 4 : nr, 2, 5, 6, srcx, srcz, nt, 1, nt,
                 acq_type, acq_type == 2 ? 2 : 0, 5, 6, nx - 6, 6};
    if (acq_type == 0) { q.rec_right_col = nx - 6; q.rec_right_row0 = 5; }
    float *lam = calloc(n, sizeof(float)), *mu = calloc(n, sizeof(float));
    float *bu = calloc(n, sizeof(float)), *dpx = calloc(n, sizeof(float));
    float *dpz = calloc(n, sizeof(float));
    double *lamd = calloc(n, sizeof(double)), *mud = calloc(n, sizeof(double));
    double *bud = calloc(n, sizeof(double)), *dpxd = calloc(n, sizeof(double));
    double *dpzd = calloc(n, sizeof(double));
    for (long k = 0; k < n; ++k) {
        lam[k] = 4e9f; mu[k] = 2e9f; bu[k] = 1.0f / 2000.0f;
        dpx[k] = (k % nx < npml || k % nx >= nx - npml) ? 30.0f : 0.0f;
        dpz[k] = (k / nx < npml || k / nx >= nz - npml) ? 30.0f : 0.0f;
        lamd[k] = lam[k]; mud[k] = mu[k]; bud[k] = bu[k]; dpxd[k] = dpx[k]; dpzd[k] = dpz[k];
    }
    const int ninj = nt;
    float *src = calloc((size_t)ns * ninj * 5, sizeof(float));
    double *srcd = calloc((size_t)ns * ninj * 5, sizeof(double));
    for (int s = 0; s < ns; ++s)
        for (int t = 0; t < 10; ++t) {
            src[((long)s * ninj + t) * 5 + 2] = src[((long)s * ninj + t) * 5 + 3] = (float)(t - 5);
            srcd[((long)s * ninj + t) * 5 + 2] = srcd[((long)s * ninj + t) * 5 + 3] = t - 5;
        }
    int chp[3] = {1, 20, nt - 1};
    const size_t nrec = (size_t)5 * ns * nt * nr;
    float *seis = calloc(nrec, sizeof(float)), *W = calloc((size_t)ns * 3 * 5 * n, sizeof(float));
    float *G = calloc(6 * n, sizeof(float));
    double *res = calloc(nrec, sizeof(double)), *g3 = calloc(3 * n, sizeof(double));
    double *seisd = calloc(nrec, sizeof(double));
    int rc = orc_forward_f(&g, &q, lam, mu, bu, dpx, dpz, src, chp, 3, seis, W);
    rc |= orc_gradient_f(&g, &q, lam, mu, bu, dpx, dpz, chp, 3, W, seis, G);
    rc |= orc_forward_d(&g, &q, lamd, mud, bud, dpxd, dpzd, srcd, chp, 3, seisd, NULL);
    for (size_t k = 0; k < nrec; ++k) res[k] = seisd[k];
    rc |= orc_exact_gradient_d(&g, &q, lamd, mud, bud, dpxd, dpzd, srcd, res, g3);
    long cells[10];
    rc |= orc_receiver_cells(&g, &q, cells);
    double *A = calloc(5 * n, sizeof(double));
    rc |= orc_inject_residual_d(&g, &q, A, res, 1, nt - 1, 0);
    rc |= orc_inject_residual_d(&g, &q, A, res, 0, 1, 1);
    free(A);
    free(lam); free(mu); free(bu); free(dpx); free(dpz);
    free(lamd); free(mud); free(bud); free(dpxd); free(dpzd);
    free(src); free(srcd); free(seis); free(W); free(G); free(res); free(g3); free(seisd);
    return rc;
}

int main(void)
{
    int rc = 0;
    for (int a = 0; a <= 2; ++a)
        for (int sdo = 2; sdo <= 4; sdo += 2) {
            const int r = run_case(a, sdo);
            printf("acq_type %d sdo %d: %s\n", a, sdo, r ? "error" : "ok");
            rc |= r;
        }
    printf(rc ? "asan-check: FAILED\n" : "asan-check: clean\n");
    return rc;
}
