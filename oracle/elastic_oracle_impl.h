/*
 * Templated body of the CPU oracle.  Included twice by elastic_oracle.c with
 *   R = float  (SFX(x) = x##_f): float storage, double intermediates wherever the
 *              reference's OpenCL expressions promote to double (unsuffixed
 *              literals on an fp64-capable device);
 *   R = double (SFX(x) = x##_d): FP64 storage and arithmetic, same constants.
 *
 * TEST INFRASTRUCTURE ONLY: nothing in the product links or calls this.
 *
 * Field layout: one contiguous block of 5 arrays [vx, vz, txx, tzz, txz], each
 * row-major [nz*nx] (x fastest), as the reference's buffers
 * (src/PyFWI/wave_propagation.py:211-223).  Taps use flat indices i*nx+j+k like
 * the reference's macros (wave_propagation.py:452-472); a flat index outside
 * [0, nz*nx) reads as zero (the reference reads past the buffer there).
 */

static inline R SFX(tap)(const orc_grid *g, int kind, const R *f, const R *a,
                         const R *b, long k)
{
    long n = (long)g->nz * g->nx;
    if (k < 0 || k >= n) return (R)0;
    switch (kind) {
    case TAP_F:   return f[k];
    case TAP_FA:  return f[k] * a[k];                 /* e.g. Avx*rho_b   */
    default:      return (a[k] + 2 * b[k]) * f[k];    /* (lam+2mu)*A      */
    }
}

/* Staggered first derivative at flat index c along stride s (1: x, nx: z).
 * plus=0: backward-staggered pairs (c,c-s),(c+s,c-2s),(c+2s,c-3s),(c+3s,c-4s);
 * plus=1: forward-staggered pairs (c+s,c),(c+2s,c-s),(c+3s,c-2s),(c+4s,c-3s).
 * Each pair difference is taken in storage precision, weighted by the double
 * coefficients, summed left to right and divided by h; the result is rounded to
 * storage precision, as the reference's `float Dxx = (...)/dx;` declarations do
 * (src/PyFWI/elastic.cl:168-172, 248-252, 385-393).  For sdo=2 the two
 * zero-weight pairs are skipped, which leaves the sum unchanged. */
static inline R SFX(der)(const orc_grid *g, int kind, const R *f, const R *a,
                         const R *b, long c, long s, int plus, double h)
{
    const double cc[4] = {g->c1, g->c2, g->c3, g->c4};
    double acc = 0.0;
    for (int k = 0; k < g->sdo; ++k) {
        long ia = plus ? (k + 1) : k;
        long ib = plus ? -k : -(k + 1);
        R d = SFX(tap)(g, kind, f, a, b, c + ia * s) - SFX(tap)(g, kind, f, a, b, c + ib * s);
        acc += cc[k] * d;
    }
    return (R)(acc / h);
}

#define DXM(F)  SFX(der)(g, TAP_F, F, 0, 0, c, 1, 0, g->dx)
#define DXP(F)  SFX(der)(g, TAP_F, F, 0, 0, c, 1, 1, g->dx)
#define DZM(F)  SFX(der)(g, TAP_F, F, 0, 0, c, g->nx, 0, g->dz)
#define DZP(F)  SFX(der)(g, TAP_F, F, 0, 0, c, g->nx, 1, g->dz)

/* Active-cell loop of the stencil kernels: rows and columns sdo..N-sdo
 * inclusive (src/PyFWI/elastic.cl:163-166). */
#define FOR_ACTIVE(lo_i, hi_i, lo_j, hi_j)                                  \
    _Pragma("omp parallel for schedule(static)")                            \
    for (int i = (lo_i); i <= (hi_i); ++i)                                  \
        for (int j = (lo_j); j <= (hi_j); ++j)

/* ---------------------------------------------------------------- forward */

/* velocity half step for vx; coeff=+1 forward, -1 time reversed
 * (src/PyFWI/elastic.cl:155-185) */
static void SFX(vel_x)(const orc_grid *g, int coeff, R *vx, const R *txx,
                       const R *txz, const R *bu, const R *dxp, const R *dzp)
{
    const double dt = g->dt;
    FOR_ACTIVE(g->sdo, g->nz - g->sdo, g->sdo, g->nx - g->sdo) {
        long c = (long)i * g->nx + j;
        R dtxx = DXM(txx), dtxz = DZM(txz);
        R damp = dxp[c] + dzp[c];
        if (coeff == 1 || ORC_MUTANT == 4) {
            vx[c] += coeff * dt * (bu[c] * (dtxx + dtxz));
            vx[c] += coeff * dt * (-coeff * damp * vx[c]);
        } else {
            vx[c] += coeff * dt * (damp * vx[c]);
            vx[c] += coeff * dt * (bu[c] * (dtxx + dtxz));
        }
    }
}

/* (src/PyFWI/elastic.cl:193-226) */
static void SFX(vel_z)(const orc_grid *g, int coeff, R *vz, const R *tzz,
                       const R *txz, const R *bu, const R *dxp, const R *dzp)
{
    const double dt = g->dt;
    FOR_ACTIVE(g->sdo, g->nz - g->sdo, g->sdo, g->nx - g->sdo) {
        long c = (long)i * g->nx + j;
        R dtxz = DXP(txz), dtzz = DZP(tzz);
        R damp = dzp[c] + dxp[c];
        if (coeff == 1) {
            vz[c] += coeff * dt * (bu[c] * (dtzz + dtxz));
            vz[c] += coeff * dt * (-damp * vz[c]);
        } else {
            vz[c] += coeff * dt * (damp * vz[c]);
            vz[c] += coeff * dt * (bu[c] * (dtzz + dtxz));
        }
    }
}

/* normal stresses (src/PyFWI/elastic.cl:235-270) */
static void SFX(tau_n)(const orc_grid *g, int coeff, const R *vx, const R *vz,
                       R *txx, R *tzz, const R *lam, const R *mu,
                       const R *dxp, const R *dzp)
{
    const double dt = g->dt;
    FOR_ACTIVE(g->sdo, g->nz - g->sdo, g->sdo, g->nx - g->sdo) {
        long c = (long)i * g->nx + j;
        R exx = DXP(vx), ezz = DZM(vz);
        R m = lam[c] + 2 * mu[c];
        R damp = dxp[c] + dzp[c];
        if (coeff == 1) {
            txx[c] += coeff * dt * (m * exx + lam[c] * ezz);
            tzz[c] += coeff * dt * (m * ezz + lam[c] * exx);
            txx[c] += coeff * dt * (-damp * txx[c]);
            tzz[c] += coeff * dt * (-damp * tzz[c]);
        } else {
            txx[c] += coeff * dt * (damp * txx[c]);
            tzz[c] += coeff * dt * (damp * tzz[c]);
            txx[c] += coeff * dt * (m * exx + lam[c] * ezz);
            tzz[c] += coeff * dt * (m * ezz + lam[c] * exx);
        }
    }
}

/* shear stress (src/PyFWI/elastic.cl:277-308) */
static void SFX(tau_s)(const orc_grid *g, int coeff, const R *vx, const R *vz,
                       R *txz, const R *mu, const R *dxp, const R *dzp)
{
    const double dt = g->dt;
    FOR_ACTIVE(g->sdo, g->nz - g->sdo, g->sdo, g->nx - g->sdo) {
        long c = (long)i * g->nx + j;
        R gxz = DXM(vz) + DZP(vx);
        R damp = dxp[c] + dzp[c];
        if (coeff == 1) {
            txz[c] += coeff * dt * (mu[c] * gxz);
            txz[c] += coeff * dt * (-damp * txz[c]);
        } else {
            txz[c] += coeff * dt * (damp * txz[c]);
            txz[c] += coeff * dt * (mu[c] * gxz);
        }
    }
}

/* ---------------------------------------------------- adjoint (pyfwi mode) */

/* adjoint stresses from adjoint velocities (src/PyFWI/elastic.cl:372-411) */
static void SFX(adj_tau)(const orc_grid *g, const R *avx, const R *avz, R *atxx,
                         R *atzz, R *atxz, const R *bu, const R *dxp,
                         const R *dzp)
{
    const double dt = g->dt;
    FOR_ACTIVE(g->sdo, g->nz - g->sdo, g->sdo, g->nx - g->sdo) {
        long c = (long)i * g->nx + j;
        R dxp_bvx = SFX(der)(g, TAP_FA, avx, bu, 0, c, 1, 1, g->dx);
        R dzm_bvz = SFX(der)(g, TAP_FA, avz, bu, 0, c, g->nx, 0, g->dz);
        atxx[c] += -dt * (ORC_MUT(1, dzp[c], dxp[c]) * atxx[c] + dxp_bvx);
        atzz[c] += -dt * (dzp[c] * atzz[c] + dzm_bvz);
        R dzp_bvx = SFX(der)(g, TAP_FA, avx, bu, 0, c, g->nx, 1, g->dz);
        R dxm_bvz = SFX(der)(g, TAP_FA, avz, bu, 0, c, 1, 0, g->dx);
        atxz[c] += -dt * (ORC_MUT(11, dxp[c], dxp[c] + dzp[c]) * atxz[c] + (dzp_bvx + dxm_bvz));
    }
}

/* adjoint velocities from adjoint stresses (src/PyFWI/elastic.cl:419-468) */
static void SFX(adj_v)(const orc_grid *g, R *avx, R *avz, const R *atxx,
                       const R *atzz, const R *atxz, const R *lam, const R *mu,
                       const R *dxp, const R *dzp)
{
    const double dt = g->dt;
    FOR_ACTIVE(g->sdo, g->nz - g->sdo, g->sdo, g->nx - g->sdo) {
        long c = (long)i * g->nx + j;
        R a1 = SFX(der)(g, TAP_FM, atxx, lam, mu, c, 1, 0, g->dx);
        R a2 = SFX(der)(g, TAP_FA, atzz, lam, 0, c, 1, 0, g->dx);
        R a3 = SFX(der)(g, TAP_FA, atxz, mu, 0, c, g->nx, 0, g->dz);
        R b1 = SFX(der)(g, TAP_FM, atzz, lam, mu, c, g->nx, 1, g->dz);
        R b2 = SFX(der)(g, TAP_FA, atxx, lam, 0, c, g->nx, 1, g->dz);
        R b3 = SFX(der)(g, TAP_FA, atxz, mu, 0, c, 1, 1, g->dx);
        avx[c] += -dt * (ORC_MUT(2, dxp[c] + dzp[c], 2 * dxp[c] + dzp[c]) * avx[c] + (a1 + a2 + a3));
        avz[c] += -dt * (ORC_MUT(3, 2 * dxp[c] + dzp[c], dxp[c] + 2 * dzp[c]) * avz[c] + (b2 + b1 + b3));
    }
}

/* imaging condition + squared-energy preconditioners, interior only
 * (src/PyFWI/elastic.cl:470-526).  The shear term takes the z-derivative of
 * vz where the name says vx (elastic.cl:485-486); reproduced.
 * G = [gmu, gmu_pre, glam, glam_pre, grho, grho_pre]. */
static void SFX(grad)(const orc_grid *g, const R *F, const R *A, const R *bu,
                      R *G)
{
    const long n = (long)g->nz * g->nx;
    const R *vx = F, *vz = F + n, *txx = F + 2 * n, *tzz = F + 3 * n, *txz = F + 4 * n;
    const R *avx = A, *avz = A + n, *atxx = A + 2 * n, *atzz = A + 3 * n, *atxz = A + 4 * n;
    R *gmu = G, *gmu_p = G + n, *glam = G + 2 * n, *glam_p = G + 3 * n;
    R *grho = G + 4 * n, *grho_p = G + 5 * n;
    const double dt = g->dt;
    const int lo = g->npml + g->sdo;
    FOR_ACTIVE(lo, g->nz - g->sdo - g->npml, lo, g->nx - g->sdo - g->npml) {
        long c = (long)i * g->nx + j;
        R dzp_vz = ORC_MUT(6, DZP(vx), DZP(vz)), dxm_vz = DXM(vz), dxp_vx = DXP(vx), dzm_vz = DZM(vz);
        R dzp_tzz = DZP(tzz), dxm_txx = DXM(txx), dxp_txz = DXP(txz), dzm_txz = DZM(txz);
        gmu[c] += ORC_MUT(5, 1, 2) * dt * (atxx[c] * dxp_vx + atzz[c] * dzm_vz) + dt * atxz[c] * (dzp_vz + dxm_vz);
        gmu_p[c] += ORC_MUT(8, 2, 4) * dt * dt * (txx[c] * txx[c] * dxp_vx * dxp_vx + tzz[c] * tzz[c] * dzm_vz * dzm_vz)
                    + dt * dt * txz[c] * txz[c] * (dzp_vz * dzp_vz + dxm_vz * dxm_vz);
        glam[c] += dt * (atxx[c] + atzz[c]) * (dxp_vx + dzm_vz);
        glam_p[c] += dt * dt * (txx[c] * txx[c] + ORC_MUT(9, txz[c] * txz[c], tzz[c] * tzz[c])) * (dxp_vx * dxp_vx + dzm_vz * dzm_vz);
        grho[c] += -(dt * ORC_MUT(7, bu[c], bu[c] * bu[c])) * (avx[c] * (dxm_txx + dzm_txz) + avz[c] * (dzp_tzz + dxp_txz));
        grho_p[c] += dt * dt * (ORC_MUT(10, vz[c] * vz[c], vx[c] * vx[c]) * (dxm_txx * dxm_txx + dzm_txz * dzm_txz)
                                + ORC_MUT(10, vx[c] * vx[c], vz[c] * vz[c]) * (dzp_tzz * dzp_tzz + dxp_txz * dxp_txz));
    }
}

/* --------------------------------------------------- composite step helpers */

void SFX(orc_fwd_step)(const orc_grid *g, R *F, const R *lam, const R *mu,
                       const R *bu, const R *dxp, const R *dzp)
{
    const long n = (long)g->nz * g->nx;
    SFX(vel_x)(g, 1, F, F + 2 * n, F + 4 * n, bu, dxp, dzp);
    SFX(vel_z)(g, 1, F + n, F + 3 * n, F + 4 * n, bu, dxp, dzp);
    SFX(tau_n)(g, 1, F, F + n, F + 2 * n, F + 3 * n, lam, mu, dxp, dzp);
    SFX(tau_s)(g, 1, F, F + n, F + 4 * n, mu, dxp, dzp);
}

/* stresses first, then velocities (used only to state the transpose that the
 * pyfwi adjoint step realises) */
void SFX(orc_fwd_step_sv)(const orc_grid *g, R *F, const R *lam, const R *mu,
                          const R *bu, const R *dxp, const R *dzp)
{
    const long n = (long)g->nz * g->nx;
    SFX(tau_n)(g, 1, F, F + n, F + 2 * n, F + 3 * n, lam, mu, dxp, dzp);
    SFX(tau_s)(g, 1, F, F + n, F + 4 * n, mu, dxp, dzp);
    SFX(vel_x)(g, 1, F, F + 2 * n, F + 4 * n, bu, dxp, dzp);
    SFX(vel_z)(g, 1, F + n, F + 3 * n, F + 4 * n, bu, dxp, dzp);
}

/* time-reversed step in the reference's order txz, (txx,tzz), vz, vx
 * (src/PyFWI/wave_propagation.py:780-812) */
void SFX(orc_rev_step)(const orc_grid *g, R *F, const R *lam, const R *mu,
                       const R *bu, const R *dxp, const R *dzp)
{
    const long n = (long)g->nz * g->nx;
    SFX(tau_s)(g, -1, F, F + n, F + 4 * n, mu, dxp, dzp);
    SFX(tau_n)(g, -1, F, F + n, F + 2 * n, F + 3 * n, lam, mu, dxp, dzp);
    SFX(vel_z)(g, -1, F + n, F + 3 * n, F + 4 * n, bu, dxp, dzp);
    SFX(vel_x)(g, -1, F, F + 2 * n, F + 4 * n, bu, dxp, dzp);
}

void SFX(orc_adj_step)(const orc_grid *g, R *A, const R *lam, const R *mu,
                       const R *bu, const R *dxp, const R *dzp)
{
    const long n = (long)g->nz * g->nx;
    SFX(adj_tau)(g, A, A + n, A + 2 * n, A + 3 * n, A + 4 * n, bu, dxp, dzp);
    SFX(adj_v)(g, A, A + n, A + 2 * n, A + 3 * n, A + 4 * n, lam, mu, dxp, dzp);
}

void SFX(orc_grad_acc)(const orc_grid *g, const R *F, const R *A, const R *bu,
                       R *G)
{
    SFX(grad)(g, F, A, bu, G);
}

/* ------------------------------------------- sparse source and receivers */

/* record the 5 fields at the receivers into row `row`, then add the source
 * sample to vx, vz, txx, tzz at the source cell (elastic_surface.cl:7-41,
 * elastic_crosswell.cl:1-33, elastic.cl:43-98).  cell: orc_rec_cells; a cell
 * outside the grid records 0.  seis: [5][nt_ext][nr] for this shot. */
static void SFX(inj_rec)(const orc_grid *g, const orc_acq *q, const long *cell, R *F,
                         int sx, int sz, const R *sv, R *seis, int row)
{
    const long n = (long)g->nz * g->nx;
    for (int ir = 0; row < q->nt_ext && ir < q->nr; ++ir) {
        const long k = cell[ir];
        if (k == ORC_NO_CELL) continue;
        for (int f = 0; f < 5; ++f)
            seis[((long)f * q->nt_ext + row) * q->nr + ir] =
                (k >= 0 && k < n) ? F[orc_rec_field(q, f) * n + k] : (R)0;
    }
    long c = (long)sz * g->nx + sx;
    for (int f = 0; f < 4; ++f) F[f * n + c] += sv[f];
}

/* add residual row `row` to the adjoint fields at the receivers, component f
 * into adjoint field f (elastic_surface.cl:48-72, elastic_crosswell.cl:35-58,
 * elastic.cl:316-363).  res: [5][ns][nt_ext][nr]. */
static void SFX(adj_inj)(const orc_grid *g, const orc_acq *q, const long *cell, R *A,
                         const R *res, int s, int row)
{
    const long n = (long)g->nz * g->nx;
    for (int ir = 0; ir < q->nr; ++ir) {
        const long k = cell[ir];
        if (k < 0 || k >= n) continue;
        for (int f = 0; f < 5; ++f)
            A[f * n + k] += res[(((long)f * q->ns + s) * q->nt_ext + row) * q->nr + ir];
    }
}

/* transpose of the gather of inj_rec: component f's residual goes to the
 * field that component recorded (the exact adjoint; differs from adj_inj
 * only for the crosswell vz record). */
static void SFX(rec_adj)(const orc_grid *g, const orc_acq *q, const long *cell, R *A,
                         const R *res, int s, int row)
{
    const long n = (long)g->nz * g->nx;
    for (int ir = 0; ir < q->nr; ++ir) {
        const long k = cell[ir];
        if (k < 0 || k >= n) continue;
        for (int f = 0; f < 5; ++f)
            A[orc_rec_field(q, f) * n + k] +=
                res[(((long)f * q->ns + s) * q->nt_ext + row) * q->nr + ir];
    }
}

/* Exported for tests: residual row `row` of shot s added to A [5][n] by the
 * reference's injection (exact = 0) or by the transpose of the gather
 * (exact = 1).  res: [5][ns][nt_ext][nr]. */
int SFX(orc_inject_residual)(const orc_grid *g, const orc_acq *q, R *A, const R *res,
                             int s, int row, int exact)
{
    long *cell = (long *)malloc((size_t)(q->nr > 0 ? q->nr : 1) * sizeof(long));
    if (!cell) return 1;
    orc_rec_cells(g, q, cell);
    if (exact) SFX(rec_adj)(g, q, cell, A, res, s, row);
    else SFX(adj_inj)(g, q, cell, A, res, s, row);
    free(cell);
    return 0;
}

/* ------------------------------------------------------------ host loops */

/* Forward modelling of all shots (src/PyFWI/wave_propagation.py:625-733).
 *   src  [ns][ninj][5] source samples (w[k]*component, zero past the wavelet)
 *   seis [5][ns][nt_ext][nr] output, rows never written stay as given
 *   W    [ns][nchp][5][n] snapshot of the state after step chp[c], or NULL */
int SFX(orc_forward)(const orc_grid *g, const orc_acq *q, const R *lam,
                     const R *mu, const R *bu, const R *dxp, const R *dzp,
                     const R *src, const int *chp, int nchp, R *seis, R *W)
{
    const long n = (long)g->nz * g->nx;
    const int ninj = (q->nt + q->dt_scale - 1) / q->dt_scale;
    R *F = (R *)calloc(5 * n, sizeof(R));
    R *sb = (R *)calloc((size_t)5 * q->nt_ext * q->nr, sizeof(R));
    long *cell = (long *)malloc((size_t)(q->nr > 0 ? q->nr : 1) * sizeof(long));
    if (!F || !sb || !cell) { free(F); free(sb); free(cell); return 1; }
    orc_rec_cells(g, q, cell);
    for (int s = 0; s < q->ns; ++s) {
        memset(F, 0, 5 * n * sizeof(R));
        int chpc = 0, t_src = 0;
        for (int t = 0; t < q->nt; ++t) {
            if (t % q->dt_scale == 0) {
                if (t_src < ninj)
                    SFX(inj_rec)(g, q, cell, F, q->srcx[s], q->srcz[s],
                                 src + ((long)s * ninj + t_src) * 5, sb, t_src);
                t_src++;
            }
            SFX(orc_fwd_step)(g, F, lam, mu, bu, dxp, dzp);
            int hit = 0;
            for (int c = 0; c < nchp; ++c) hit |= (chp[c] == t);
            if (hit) {
                if (W) memcpy(W + ((long)s * nchp + chpc) * 5 * n, F, 5 * n * sizeof(R));
                chpc++;
            }
        }
        for (int f = 0; f < 5; ++f)
            memcpy(seis + (((long)f * q->ns + s) * q->nt_ext) * q->nr,
                   sb + (long)f * q->nt_ext * q->nr,
                   (size_t)q->nt_ext * q->nr * sizeof(R));
    }
    free(F); free(sb); free(cell);
    return 0;
}

/* Adjoint-state gradient as the reference computes it
 * (src/PyFWI/wave_propagation.py:735-856): per shot, reconstruct the forward
 * state by restoring snapshots at chp[] and reversing single steps elsewhere,
 * inject residual rows, adjoint stress then adjoint velocity, correlate.
 *   res [5][ns][nt_ext][nr];  G [6][n] accumulated over shots (not zeroed here) */
int SFX(orc_gradient)(const orc_grid *g, const orc_acq *q, const R *lam,
                      const R *mu, const R *bu, const R *dxp, const R *dzp,
                      const int *chp, int nchp, const R *W, const R *res, R *G)
{
    const long n = (long)g->nz * g->nx;
    R *F = (R *)calloc(5 * n, sizeof(R));
    R *A = (R *)calloc(5 * n, sizeof(R));
    long *cell = (long *)malloc((size_t)(q->nr > 0 ? q->nr : 1) * sizeof(long));
    if (!F || !A || !cell) { free(F); free(A); free(cell); return 1; }
    orc_rec_cells(g, q, cell);
    for (int s = 0; s < q->ns; ++s) {
        memset(F, 0, 5 * n * sizeof(R));
        memset(A, 0, 5 * n * sizeof(R));
        int chpc = nchp - 1;
        int t_src = q->nt_ext;
        for (int t = q->nt - 1; t > 0; --t) {
            if (ORC_MUTANT != 12 && chpc >= 0 && t == chp[chpc]) {
                memcpy(F, W + ((long)s * nchp + chpc) * 5 * n, 5 * n * sizeof(R));
                chpc--;
            } else {
                SFX(orc_rev_step)(g, F, lam, mu, bu, dxp, dzp);
            }
            if (t % q->dt_scale == 0) {
                if (t_src - 1 >= 0) SFX(adj_inj)(g, q, cell, A, res, s, t_src - 1);
                t_src--;
            }
            SFX(orc_adj_step)(g, A, lam, mu, bu, dxp, dzp);
            SFX(grad)(g, F, A, bu, G);
        }
    }
    free(F); free(A); free(cell);
    return 0;
}

/* ------------------------------------------ exact discrete adjoint (FD-checkable)
 *
 * Transpose of orc_fwd_step with the source and the receiver sampling, for
 *   J = sum_k sum_r res[k][r] * seis[k][r]   (seis = sampled S_t, t = k*dt_scale)
 * so the result is dJ/d(lam, mu, b) over the whole padded grid.  With
 * res = seis - obs this is the gradient of 0.5*||seis - obs||^2.
 * Uses: D+^T = -D- and D-^T = -D+ for fields that vanish outside the active
 * cells; damping factor (1 - dt*d) transposes to itself.
 *   g3 [3][n] = [g_lam, g_mu, g_b], accumulated over shots (not zeroed here). */
int SFX(orc_exact_gradient)(const orc_grid *g, const orc_acq *q, const R *lam,
                            const R *mu, const R *bu, const R *dxp,
                            const R *dzp, const R *src, const R *res, R *g3)
{
    const long n = (long)g->nz * g->nx;
    const int nt = q->nt;
    const int ninj = (nt + q->dt_scale - 1) / q->dt_scale;
    const double dt = g->dt;
    /* P[t] = state after the injection of step t (input of the update),
     * V[t] = velocities after the velocity half step of step t. */
    R *P = (R *)malloc((size_t)nt * 5 * n * sizeof(R));
    R *V = (R *)malloc((size_t)nt * 2 * n * sizeof(R));
    R *F = (R *)calloc(5 * n, sizeof(R));
    R *A = (R *)calloc(5 * n, sizeof(R));
    R *Q = (R *)calloc(5 * n, sizeof(R));
    R *sb = (R *)calloc((size_t)5 * q->nt_ext * q->nr, sizeof(R));
    long *cell = (long *)malloc((size_t)(q->nr > 0 ? q->nr : 1) * sizeof(long));
    if (!P || !V || !F || !A || !Q || !sb || !cell) {
        free(P); free(V); free(F); free(A); free(Q); free(sb); free(cell); return 1;
    }
    orc_rec_cells(g, q, cell);
    R *glam = g3, *gmu = g3 + n, *gb = g3 + 2 * n;
    const int lo = g->sdo, hiz = g->nz - g->sdo, hix = g->nx - g->sdo;
    for (int s = 0; s < q->ns; ++s) {
        memset(F, 0, 5 * n * sizeof(R));
        int t_src = 0;
        for (int t = 0; t < nt; ++t) {
            if (t % q->dt_scale == 0) {
                if (t_src < ninj)
                    SFX(inj_rec)(g, q, cell, F, q->srcx[s], q->srcz[s],
                                 src + ((long)s * ninj + t_src) * 5, sb, t_src);
                t_src++;
            }
            memcpy(P + (size_t)t * 5 * n, F, 5 * n * sizeof(R));
            SFX(vel_x)(g, 1, F, F + 2 * n, F + 4 * n, bu, dxp, dzp);
            SFX(vel_z)(g, 1, F + n, F + 3 * n, F + 4 * n, bu, dxp, dzp);
            memcpy(V + (size_t)t * 2 * n, F, 2 * n * sizeof(R));
            SFX(tau_n)(g, 1, F, F + n, F + 2 * n, F + 3 * n, lam, mu, dxp, dzp);
            SFX(tau_s)(g, 1, F, F + n, F + 4 * n, mu, dxp, dzp);
        }
        memset(A, 0, 5 * n * sizeof(R));
        R *avx = A, *avz = A + n, *atxx = A + 2 * n, *atzz = A + 3 * n, *atxz = A + 4 * n;
        R *qxx = Q + 2 * n, *qzz = Q + 3 * n, *qxz = Q + 4 * n, *px = Q, *pz = Q + n;
        for (int t = nt - 1; t >= 0; --t) {
            const R *vx = V + (size_t)t * 2 * n, *vz = vx + n;
            const R *Pt = P + (size_t)t * 5 * n;
            const R *txx = Pt + 2 * n, *tzz = Pt + 3 * n, *txz = Pt + 4 * n;
            /* adjoint of the stress half step */
            for (long c = 0; c < 5 * n; ++c) Q[c] = 0;
            FOR_ACTIVE(lo, hiz, lo, hix) {
                long c = (long)i * g->nx + j;
                R keep = 1 - dt * (dxp[c] + dzp[c]);
                qxx[c] = atxx[c] * keep; qzz[c] = atzz[c] * keep; qxz[c] = atxz[c] * keep;
            }
            FOR_ACTIVE(lo, hiz, lo, hix) {
                long c = (long)i * g->nx + j;
                R exx = DXP(vx), ezz = DZM(vz), gxz = DXM(vz) + DZP(vx);
                glam[c] += dt * (qxx[c] + qzz[c]) * (exx + ezz);
                gmu[c] += dt * (2 * qxx[c] * exx + 2 * qzz[c] * ezz + qxz[c] * gxz);
                /* Av += dt * D^T(C q) */
                R sx = SFX(der)(g, TAP_FM, qxx, lam, mu, c, 1, 0, g->dx)
                     + SFX(der)(g, TAP_FA, qzz, lam, 0, c, 1, 0, g->dx)
                     + SFX(der)(g, TAP_FA, qxz, mu, 0, c, g->nx, 0, g->dz);
                R sz = SFX(der)(g, TAP_FA, qxx, lam, 0, c, g->nx, 1, g->dz)
                     + SFX(der)(g, TAP_FM, qzz, lam, mu, c, g->nx, 1, g->dz)
                     + SFX(der)(g, TAP_FA, qxz, mu, 0, c, 1, 1, g->dx);
                avx[c] -= dt * sx;
                avz[c] -= dt * sz;
            }
            /* adjoint of the velocity half step */
            FOR_ACTIVE(lo, hiz, lo, hix) {
                long c = (long)i * g->nx + j;
                R keep = 1 - dt * (dxp[c] + dzp[c]);
                px[c] = avx[c] * keep; pz[c] = avz[c] * keep;
            }
            FOR_ACTIVE(lo, hiz, lo, hix) {
                long c = (long)i * g->nx + j;
                gb[c] += dt * (px[c] * (DXM(txx) + DZM(txz)) + pz[c] * (DZP(tzz) + DXP(txz)));
                R rxx = SFX(der)(g, TAP_FA, px, bu, 0, c, 1, 1, g->dx);
                R rzz = SFX(der)(g, TAP_FA, pz, bu, 0, c, g->nx, 0, g->dz);
                R rxz = SFX(der)(g, TAP_FA, px, bu, 0, c, g->nx, 1, g->dz)
                      + SFX(der)(g, TAP_FA, pz, bu, 0, c, 1, 0, g->dx);
                atxx[c] = qxx[c] - dt * rxx;
                atzz[c] = qzz[c] - dt * rzz;
                atxz[c] = qxz[c] - dt * rxz;
                avx[c] = px[c];
                avz[c] = pz[c];
            }
            /* adjoint of the receiver sampling of S_t */
            if (t % q->dt_scale == 0) {
                int row = t / q->dt_scale;
                if (row < q->nt_ext) SFX(rec_adj)(g, q, cell, A, res, s, row);
            }
        }
    }
    free(P); free(V); free(F); free(A); free(Q); free(sb); free(cell);
    return 0;
}

#undef DXM
#undef DXP
#undef DZM
#undef DZP
#undef FOR_ACTIVE
