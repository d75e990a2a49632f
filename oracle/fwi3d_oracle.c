/*
 * fwi3d_oracle.c -- TEST INFRASTRUCTURE ONLY (see fwi3d_oracle.h).
 *
 * CPU restatement of the reference's C v1.4 forward path, in the reference's
 * structure: 12 per-cell coefficient volumes, full-volume psi arrays, one
 * OpenMP loop nest per field component / CPML face, receiver gather per step.
 * This is also the "port" CPU baseline timed beside the GPU path.
 *
 * Each function names the reference lines it restates.  Compile with
 * -ffp-contract=off: the evaluation order below is the reference's, and
 * fused multiply-adds would change the last bits.
 */
#define _GNU_SOURCE
#include "fwi3d_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct or_state {
    or_params p;
    long n;                        /* ix*iy*iz */
    double *vol[OR_NARRAYS];
    double *prof[3][OR_NPROFILES];
};

static int axis_len(const or_params *p, int a)
{
    return a == 0 ? p->ix : (a == 1 ? p->iy : p->iz);
}

static double axis_step(const or_params *p, int a)
{
    return a == 0 ? p->dx : (a == 1 ? p->dy : p->dz);
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* thread count of the following parallel regions (the CPU baseline's 1-thread
 * and all-core runs) */
void or_set_num_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/* STREAM triad a = b + s*c over n doubles per array with the current thread
 * count, first touch by the same threads; best of `reps` passes in GB/s
 * (24 bytes moved per element, as STREAM counts them).  Host-memory bandwidth
 * beside the CPU baseline; no counterpart in the reference. */
double or_stream_triad(long n, int reps)
{
    double *a = (double *)malloc(sizeof(double) * n);
    double *b = (double *)malloc(sizeof(double) * n);
    double *c = (double *)malloc(sizeof(double) * n);
    if (!a || !b || !c) {
        free(a);
        free(b);
        free(c);
        return -1.0;
    }
#pragma omp parallel for schedule(static)
    for (long q = 0; q < n; ++q) {
        a[q] = 0.0;
        b[q] = 1.0;
        c[q] = 2.0;
    }
    double best = 0.0;
    for (int r = 0; r < reps; ++r) {
        const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q)
            a[q] = b[q] + 3.0 * c[q];
        const double t = omp_get_wtime() - t0;
        const double gbs = 24.0 * (double)n / t / 1e9;
        if (gbs > best)
            best = gbs;
    }
    free(a);
    free(b);
    free(c);
    return best;
}

/* satan_i/fwi3d_cpu_lattice_v1.4.c:35-39: cmax from the minimum conductivity,
 * Courant number of the 3-D grid, dt = courant*6/7*factor (factor 0.999f in
 * the reference; 6/7 = 1/(|c1|+|c2|) is the 4th-order stability margin). */
double or_lattice_time_3d(double minval, double omega0, double mu0,
                          double dx, double dy, double dz, double factor,
                          double *cmax_out)
{
    double cmax = sqrt(2.0 * omega0 / mu0 / minval);
    double inv = 1.0 / pow(dx, 2) + 1.0 / pow(dy, 2) + 1.0 / pow(dz, 2);
    double courant = 1.0 / cmax / sqrt(inv);
    if (cmax_out)
        *cmax_out = cmax;
    return courant * 6.0 / 7.0 * factor;
}

or_state *or_create(const or_params *p)
{
    or_state *s = (or_state *)calloc(1, sizeof(or_state));
    if (!s)
        return NULL;
    s->p = *p;
    s->n = (long)p->ix * p->iy * p->iz;
    for (int a = 0; a < OR_NARRAYS; ++a) {
        s->vol[a] = (double *)calloc((size_t)s->n, sizeof(double));
        if (!s->vol[a]) {
            or_destroy(s);
            return NULL;
        }
    }
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < OR_NPROFILES; ++q) {
            s->prof[a][q] = (double *)calloc((size_t)axis_len(p, a), sizeof(double));
            if (!s->prof[a][q]) {
                or_destroy(s);
                return NULL;
            }
        }
    return s;
}

void or_destroy(or_state *s)
{
    if (!s)
        return;
    for (int a = 0; a < OR_NARRAYS; ++a)
        free(s->vol[a]);
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < OR_NPROFILES; ++q)
            free(s->prof[a][q]);
    free(s);
}

double *or_array(or_state *s, int id)
{
    return (id >= 0 && id < OR_NARRAYS) ? s->vol[id] : NULL;
}

double *or_profile(or_state *s, int axis, int id)
{
    if (axis < 0 || axis > 2 || id < 0 || id >= OR_NPROFILES)
        return NULL;
    return s->prof[axis][id];
}

/* satan_i/fwi3d_cpu_cpmlabc_v1.4.c:7-175.  Cubic grading of sigma, linear
 * alpha, kappa = 1 + (kappa_max-1)*ramp^nn, recursive-convolution constants
 * b = exp(-(sigma/kappa+alpha)dt), c = sigma(b-1)/(sigma+kappa*alpha)/kappa.
 * The ramps are evaluated on integer (E) and half-integer (H) offsets with the
 * reference's float-typed offsets, so the edge cell exceeds sigma_max. */
void or_init_cpml(or_state *s)
{
    const or_params *p = &s->p;
    const int nc = p->ncpml;
    const double expo = p->nn + p->order;
    for (int a = 0; a < 3; ++a) {
        const int n = axis_len(p, a);
        const double h = axis_step(p, a);
        const double delta = nc * h;
        const double smax = (p->nn + p->order + 1.0) * p->cmax * log(1.0 / p->rcoef)
                            / (2.0 * delta) * p->opt_to_max[a];
        double **q = s->prof[a];
        for (int i = 0; i < n; ++i) {
            double re, rm, ae, am; /* ramp arguments (float-typed in the reference) */
            int inside = 1;
            if (i < nc) {
                re = (double)(nc - i);
                rm = (double)((float)(nc - i) - 0.5f);
                ae = (double)i;
                am = (double)((float)i + 0.5f);
            } else if (i > n - nc - 1) {
                re = (double)((float)(i - n) + 1.0f + (float)nc);
                rm = (double)((float)(i - n) + 0.5f + (float)nc);
                ae = (double)((float)(-i + n) - 1.0f);
                am = (double)((float)(-i + n) - 0.5f);
            } else {
                inside = 0;
                re = rm = ae = am = 0.0;
            }
            if (inside) {
                const double den = (double)((float)nc - 1.0f);
                q[OR_P_ESIG][i] = smax * pow(re / den, expo);
                q[OR_P_MSIG][i] = smax * pow(rm / den, expo);
                q[OR_P_EKAP][i] = 1.0 + (p->kappa_max - 1.0) * pow(re / den, p->nn);
                q[OR_P_MKAP][i] = 1.0 + (p->kappa_max - 1.0) * pow(rm / den, p->nn);
                q[OR_P_AE][i] = p->alpha_max * pow(ae / den, p->ma);
                q[OR_P_AM][i] = p->alpha_max * pow(am / den, p->ma);
                q[OR_P_BE][i] = exp(-(q[OR_P_ESIG][i] / q[OR_P_EKAP][i] + q[OR_P_AE][i]) * p->dt);
                q[OR_P_BH][i] = exp(-(q[OR_P_MSIG][i] / q[OR_P_MKAP][i] + q[OR_P_AM][i]) * p->dt);
                q[OR_P_CE][i] = q[OR_P_ESIG][i] * (q[OR_P_BE][i] - 1.0)
                                / (q[OR_P_ESIG][i] + q[OR_P_EKAP][i] * q[OR_P_AE][i])
                                / q[OR_P_EKAP][i];
                q[OR_P_CH][i] = q[OR_P_MSIG][i] * (q[OR_P_BH][i] - 1.0)
                                / (q[OR_P_MSIG][i] + q[OR_P_MKAP][i] * q[OR_P_AM][i])
                                / q[OR_P_MKAP][i];
            } else {
                q[OR_P_ESIG][i] = q[OR_P_MSIG][i] = 0.0;
                q[OR_P_EKAP][i] = q[OR_P_MKAP][i] = 1.0;
                q[OR_P_AE][i] = q[OR_P_AM][i] = 0.0;
                q[OR_P_BE][i] = q[OR_P_CE][i] = 0.0;
                q[OR_P_BH][i] = q[OR_P_CH][i] = 0.0;
            }
            q[OR_P_KED][i] = q[OR_P_EKAP][i] * h;
            q[OR_P_KHD][i] = q[OR_P_MKAP][i] * h;
        }
    }
}

/* satan_i/fwi3d_cpu_function_v1.4.c:325-435 (coef1..coef4): fictitious
 * permittivity eps' = sigma/2/omega0 and the lossy-CPML update factors per
 * component, each graded along the component's own axis. */
void or_media_coeff_3d(or_state *s)
{
    const or_params *p = &s->p;
    const int ix = p->ix, iy = p->iy, iz = p->iz;
    const double dt = p->dt;
#pragma omp parallel for
    for (int k = 0; k < iz; ++k)
        for (int j = 0; j < iy; ++j)
            for (int i = 0; i < ix; ++i) {
                const long c = ((long)k * iy + j) * ix + i;
                const double eps2 = s->vol[OR_SIG][c] / 2.0 / p->omega0;
                const int idx[3] = {i, j, k};
                for (int d = 0; d < 3; ++d) {
                    const double se = s->prof[d][OR_P_ESIG][idx[d]] * dt / (2.0 * eps2);
                    const double sm = s->prof[d][OR_P_MSIG][idx[d]] * dt / (2.0 * eps2);
                    s->vol[OR_CA_X + d][c] = (1.0 - se) / (1.0 + se);
                    s->vol[OR_DA_X + d][c] = (1.0 - sm) / (1.0 + sm);
                    s->vol[OR_CB_X + d][c] = dt / eps2 / (1.0 + se);
                    s->vol[OR_DB_X + d][c] = dt / p->mu0 / (1.0 + sm);
                }
            }
}

/* satan_i/fwi3d_cpu_function_v1.4.c:964-986 */
void or_set_zero_eh(or_state *s)
{
    static const int ids[] = {OR_EX, OR_EY, OR_EZ, OR_HX, OR_HY, OR_HZ,
                              OR_PSI_EZX, OR_PSI_EYX, OR_PSI_EXY, OR_PSI_EZY,
                              OR_PSI_EYZ, OR_PSI_EXZ, OR_PSI_HYX, OR_PSI_HZX,
                              OR_PSI_HZY, OR_PSI_HXY, OR_PSI_HXZ, OR_PSI_HYZ};
    for (size_t q = 0; q < sizeof(ids) / sizeof(ids[0]); ++q)
        memset(s->vol[ids[q]], 0, (size_t)s->n * sizeof(double));
}

/* 4th-order staggered differences in the reference's evaluation order:
 * backward (E update, function.c:1203) c1*F0 - c1*F-1 + c2*F+1 - c2*F-2,
 * forward (H update, function.c:1257) c1*F+1 - c1*F0 + c2*F+2 - c2*F-1. */
static inline double dminus(const double *F, long c, long st, double c1, double c2)
{
    return c1 * F[c] - c1 * F[c - st] + c2 * F[c + st] - c2 * F[c - 2 * st];
}

static inline double dplus(const double *F, long c, long st, double c1, double c2)
{
    return c1 * F[c + st] - c1 * F[c] + c2 * F[c + 2 * st] - c2 * F[c - st];
}

/* One component of satan_i/fwi3d_cpu_function_v1.4.c:1184-1236 (sign=+1) or
 * the sign-flipped back-propagation variant :1292-1333 (sign=-1):
 * E_d = ca_d*E_d +/- cb_d*(D-_a H_b / ked_a - D-_b H_a / ked_b), (d,a,b) cyclic.
 * Range: own axis [1,n-1), the two others [2,n-1). */
static void e_component(or_state *s, int d, int sign)
{
    const or_params *p = &s->p;
    const int a = (d + 1) % 3, b = (d + 2) % 3;
    const long st[3] = {1, p->ix, (long)p->ix * p->iy};
    int lo[3], hi[3];
    for (int q = 0; q < 3; ++q) {
        lo[q] = (q == d) ? 1 : 2;
        hi[q] = axis_len(p, q) - 1;
    }
    double *E = s->vol[OR_EX + d];
    const double *Ha = s->vol[OR_HX + a], *Hb = s->vol[OR_HX + b];
    const double *ca = s->vol[OR_CA_X + d], *cb = s->vol[OR_CB_X + d];
    const double *keda = s->prof[a][OR_P_KED], *kedb = s->prof[b][OR_P_KED];
    const double c1 = p->c1, c2 = p->c2;
#pragma omp parallel for
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                const int idx[3] = {i, j, k};
                const long c = k * st[2] + j * st[1] + i;
                const double curl = dminus(Hb, c, st[a], c1, c2) / keda[idx[a]]
                                  - dminus(Ha, c, st[b], c1, c2) / kedb[idx[b]];
                if (sign > 0)
                    E[c] = ca[c] * E[c] + cb[c] * curl;
                else
                    E[c] = ca[c] * E[c] - cb[c] * curl;
            }
}

void or_e_field_cpml42(or_state *s, int sign)
{
    for (int d = 0; d < 3; ++d)
        e_component(s, d, sign);
}

/* One component of satan_i/fwi3d_cpu_function_v1.4.c:1238-1290 (sign=+1) or
 * :1335-1372 (sign=-1): H_d = da_d*H_d -/+ db_d*(D+_a E_b/khd_a - D+_b E_a/khd_b).
 * Range: own axis [1,n-1), the two others [1,n-2). */
static void h_component(or_state *s, int d, int sign)
{
    const or_params *p = &s->p;
    const int a = (d + 1) % 3, b = (d + 2) % 3;
    const long st[3] = {1, p->ix, (long)p->ix * p->iy};
    int lo[3], hi[3];
    for (int q = 0; q < 3; ++q) {
        lo[q] = 1;
        hi[q] = axis_len(p, q) - ((q == d) ? 1 : 2);
    }
    double *H = s->vol[OR_HX + d];
    const double *Ea = s->vol[OR_EX + a], *Eb = s->vol[OR_EX + b];
    const double *da = s->vol[OR_DA_X + d], *db = s->vol[OR_DB_X + d];
    const double *khda = s->prof[a][OR_P_KHD], *khdb = s->prof[b][OR_P_KHD];
    const double c1 = p->c1, c2 = p->c2;
#pragma omp parallel for
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                const int idx[3] = {i, j, k};
                const long c = k * st[2] + j * st[1] + i;
                const double curl = dplus(Eb, c, st[a], c1, c2) / khda[idx[a]]
                                  - dplus(Ea, c, st[b], c1, c2) / khdb[idx[b]];
                if (sign > 0)
                    H[c] = da[c] * H[c] - db[c] * curl;
                else
                    H[c] = da[c] * H[c] + db[c] * curl;
            }
}

void or_h_field_cpml42(or_state *s, int sign)
{
    for (int d = 0; d < 3; ++d)
        h_component(s, d, sign);
}

/* CPML memory update on one face slab of axis a, cells [r0,r1) along a and
 * [0,n-1) along the two other axes (satan_i/fwi3d_cpu_cpmlabc_v1.4.c:402-679).
 * With p=(a+1)%3, q=(a+2)%3:
 *   E pass: psi0 <- D-_a H_p (E_q += cb_q psi0), psi1 <- D-_a H_q (E_p -= cb_p psi1)
 *   H pass: psi0 <- D+_a E_q (H_p += db_p psi0), psi1 <- D+_a E_p (H_q -= db_q psi1)
 * psi = b*psi + c*D/dl, with dl the plain grid step of axis a. */
static void cpml_slab(or_state *s, int a, int r0, int r1, int is_h)
{
    const or_params *p = &s->p;
    if (r1 <= r0)
        return;
    const int pa = (a + 1) % 3, qa = (a + 2) % 3;
    const long st[3] = {1, p->ix, (long)p->ix * p->iy};
    int lo[3] = {0, 0, 0}, hi[3];
    for (int t = 0; t < 3; ++t)
        hi[t] = axis_len(p, t) - 1;
    lo[a] = r0;
    hi[a] = r1;
    const double c1 = p->c1, c2 = p->c2;
    const double dl = axis_step(p, a);
    const double *bb = s->prof[a][is_h ? OR_P_BH : OR_P_BE];
    const double *cc = s->prof[a][is_h ? OR_P_CH : OR_P_CE];
    double *psi0, *psi1, *F0, *F1;
    const double *S0, *S1, *k0, *k1;
    static const int pe[3][2] = {{OR_PSI_EZX, OR_PSI_EYX}, {OR_PSI_EXY, OR_PSI_EZY},
                                 {OR_PSI_EYZ, OR_PSI_EXZ}};
    static const int ph[3][2] = {{OR_PSI_HYX, OR_PSI_HZX}, {OR_PSI_HZY, OR_PSI_HXY},
                                 {OR_PSI_HXZ, OR_PSI_HYZ}};
    if (!is_h) {
        psi0 = s->vol[pe[a][0]];
        psi1 = s->vol[pe[a][1]];
        S0 = s->vol[OR_HX + pa];
        S1 = s->vol[OR_HX + qa];
        F0 = s->vol[OR_EX + qa];
        F1 = s->vol[OR_EX + pa];
        k0 = s->vol[OR_CB_X + qa];
        k1 = s->vol[OR_CB_X + pa];
    } else {
        psi0 = s->vol[ph[a][0]];
        psi1 = s->vol[ph[a][1]];
        S0 = s->vol[OR_EX + qa];
        S1 = s->vol[OR_EX + pa];
        F0 = s->vol[OR_HX + pa];
        F1 = s->vol[OR_HX + qa];
        k0 = s->vol[OR_DB_X + pa];
        k1 = s->vol[OR_DB_X + qa];
    }
#pragma omp parallel for
    for (int k = lo[2]; k < hi[2]; ++k)
        for (int j = lo[1]; j < hi[1]; ++j)
            for (int i = lo[0]; i < hi[0]; ++i) {
                const int idx[3] = {i, j, k};
                const int r = idx[a];
                const long c = k * st[2] + j * st[1] + i;
                const double g0 = is_h ? dplus(S0, c, st[a], c1, c2) : dminus(S0, c, st[a], c1, c2);
                const double g1 = is_h ? dplus(S1, c, st[a], c1, c2) : dminus(S1, c, st[a], c1, c2);
                psi0[c] = bb[r] * psi0[c] + cc[r] * g0 / dl;
                psi1[c] = bb[r] * psi1[c] + cc[r] * g1 / dl;
                F0[c] += k0[c] * psi0[c];
                F1[c] -= k1[c] * psi1[c];
            }
}

/* satan_i/fwi3d_cpu_cpmlabc_v1.4.c:191-196 with xe/ye/ze_pml4 (:402-543):
 * slabs [2,nc-1)|[n-nc+1,n-2) on x, [2,nc)|[n-nc+1,n-2) on y and z. */
void or_epml_abcs4(or_state *s)
{
    const int nc = s->p.ncpml;
    for (int a = 0; a < 3; ++a) {
        const int n = axis_len(&s->p, a);
        cpml_slab(s, a, 2, a == 0 ? nc - 1 : nc, 0);
        cpml_slab(s, a, n - nc + 1, n - 2, 0);
    }
}

/* satan_i/fwi3d_cpu_cpmlabc_v1.4.c:198-203 with xh/yh/zh_pml4 (:545-679):
 * slabs [1,nc-1)|[n-nc,n-2) on every axis. */
void or_hpml_abcs4(or_state *s)
{
    const int nc = s->p.ncpml;
    for (int a = 0; a < 3; ++a) {
        const int n = axis_len(&s->p, a);
        cpml_slab(s, a, 1, nc - 1, 1);
        cpml_slab(s, a, n - nc, n - 2, 1);
    }
}

/* satan_i/fwi3d_cpu_function_v1.4.c:514-529: half-step averaged current
 * injected into E_comp at a 1-based cell, scaled by 1/eps' and the cell volume. */
void or_read_source_3d(or_state *s, int comp, const double *signal,
                       int jx, int jy, int jz, int step)
{
    const or_params *p = &s->p;
    const long c = (long)p->ix * p->iy * (jz - 1) + (long)p->ix * (jy - 1) + jx - 1;
    s->vol[OR_EX + comp][c] -= (signal[step] + signal[step + 1]) / 2.0 * p->dt * 2.0
                               * p->omega0 / s->vol[OR_SIG][c] / p->dx / p->dy / p->dz;
}

/* satan_i/fwi3d_cpu_function_v1.4.c:765-785 (trace part; no file output). */
void or_output(or_state *s, int nrec, const int *rec1, int it, int step,
               double *tx, double *ty, double *tz)
{
    const or_params *p = &s->p;
    for (int r = 0; r < nrec; ++r) {
        const long c = (long)p->ix * p->iy * (rec1[3 * r + 2] - 1)
                     + (long)p->ix * (rec1[3 * r + 1] - 1) + rec1[3 * r] - 1;
        tx[(long)r * it + step] = s->vol[OR_EX][c];
        ty[(long)r * it + step] = s->vol[OR_EY][c];
        tz[(long)r * it + step] = s->vol[OR_EZ][c];
    }
}

/* One forward step in the order of satan_i/fwi3d_cpu_rec2rec_v1.4.c:48-55, with
 * the volume running DFT of E after the step (the reference's copytoEcal +
 * laplaceToFreq_2 pair, function.c:1652-1686 and 2091-2200, accumulated per step
 * instead of storing the history). */
void or_forward(or_state *s, int comp, const double *signal,
                int jx, int jy, int jz, int step0, int nsteps,
                int nrec, const int *rec1, int it,
                double *tx, double *ty, double *tz,
                int nf, const double *w, double *acc)
{
    const long n = s->n;
    for (int step = step0; step < step0 + nsteps; ++step) {
        or_e_field_cpml42(s, 1);
        or_epml_abcs4(s);
        if (signal)
            or_read_source_3d(s, comp, signal, jx, jy, jz, step);
        or_h_field_cpml42(s, 1);
        or_hpml_abcs4(s);
        if (nrec > 0)
            or_output(s, nrec, rec1, it, step, tx, ty, tz);
        if (nf > 0) {
            for (int d = 0; d < 3; ++d) {
                const double *E = s->vol[OR_EX + d];
                for (int f = 0; f < nf; ++f) {
                    const double wr = w[2 * ((long)step * nf + f)];
                    const double wi = w[2 * ((long)step * nf + f) + 1];
                    double *A = acc + 2 * ((long)d * nf + f) * n;
#pragma omp parallel for
                    for (long c = 0; c < n; ++c) {
                        A[2 * c] += E[c] * wr;
                        A[2 * c + 1] += E[c] * wi;
                    }
                }
            }
        }
    }
}

/* satan_i/fwi3d_cpu_waveform_v1.4.c:132-143: first derivative of a Gaussian,
 * beta = pi*fmax^2, delay t0 = pi/fmax. */
void or_firstderiv_gauss(int nt, double dt, double fmax_w, double *out)
{
    const double beta = M_PI * pow(fmax_w, 2);
    const double t0 = M_PI / fmax_w;
    for (int i = 0; i < nt; ++i)
        out[i] = -2.0 * beta * (i * dt - t0) * sqrt(beta / M_PI)
                 * exp(-beta * pow(i * dt - t0, 2.0));
}

/* satan_i/fwi3d_cpu_waveform_v1.4.c:420-438: Ricker wavelet, fp = 1 Hz,
 * phase 1 s, sample offset 200. */
void or_ricker(int nt, double dt, double *out)
{
    const double fp = 1.0, md = 1.0, phase = md / fp;
    for (int i = 0; i < nt; ++i) {
        const double al = pow(M_PI * fp, 2) * pow((double)(i - 200) * dt - phase, 2);
        out[i] = (2.0 * al - 1.0) * exp(-al);
    }
}

/* Fictitious->diffusive running-DFT weights at arbitrary angular frequencies:
 * w[n][f] = dt*exp(-a*n*dt)*exp(i*a*n*dt), a = sqrt(omega0*omega_f)
 * (the kernel of satan_i/fwi3d_cpu_function_v1.4.c:2041-2042 and
 * ver0040/cpu_f_to_d_e.f90:140-142).  Output complex128 [it][nf]. */
void or_dft_weights(int it, double dt, double omega0, int nf,
                    const double *omega, double *w)
{
    for (int f = 0; f < nf; ++f) {
        const double a = sqrt(omega0 * omega[f]);
        for (int n = 0; n < it; ++n) {
            const double x = a * n * dt;
            const double e = dt * exp(-x);
            double sn, cs;
            sincos(x, &sn, &cs);
            w[2 * ((long)n * nf + f)] = e * cs;
            w[2 * ((long)n * nf + f) + 1] = e * sn;
        }
    }
}

/* FFTW_BACKWARD definition (un-normalised, exp(+2*pi*i*j*k/n)) of a real
 * sequence, as the reference applies it to the true waveform
 * (function.c:2006-2007).  Naive O(n^2) with exact index reduction. */
void or_fft_backward_real(int n, const double *in, double *out)
{
    for (int k = 0; k < n; ++k) {
        double re = 0.0, im = 0.0;
        for (int j = 0; j < n; ++j) {
            const long m = ((long)j * k) % n;
            const double ang = 2.0 * M_PI * (double)m / (double)n;
            re += in[j] * cos(ang);
            im += in[j] * sin(ang);
        }
        out[2 * k] = re;
        out[2 * k + 1] = im;
    }
}

/* satan_i/fwi3d_cpu_function_v1.4.c:2038-2056 for one receiver component:
 * bins omega_n = n*2*pi/(it*dt), n < nbins; E_n and the fictitious source
 * spectrum J_n = csqrt(-2*omega0/(i*omega_n)) * sum J[k] * kernel, J_0 := 2*omega0;
 * G = E/J; OBS = G * FFTW_BACKWARD(true waveform)[n].  All outputs complex128.
 * The term is evaluated as ((x*dt)*exp(-a k dt))*cexp(i a k dt) like the
 * reference so that the oracle reproduces it to the last bit. */
void or_laplace_to_freq(int it, double dt, double omega0, int nbins,
                        const double *trace, const double *jsig,
                        const double *strue,
                        double *e_w, double *j_w, double *g_w, double *obs)
{
    double *tru = (double *)malloc(sizeof(double) * 2 * (size_t)it);
    or_fft_backward_real(it, strue, tru);
    const double om = 2.0 * M_PI / it / dt;
#pragma omp parallel for
    for (int n = 0; n < nbins; ++n) {
        const double a = sqrt(omega0 * om * n);
        const double complex jf = csqrt(-2.0 * omega0 / I / om / n);
        double complex ev = 0.0, jv = 0.0;
        for (int k = 0; k < it; ++k) {
            const double decay = exp(-a * k * dt);
            const double complex osc = cexp(I * a * k * dt);
            const double complex jk = jsig[k];
            ev += trace[k] * dt * decay * osc;
            jv += jf * jk * dt * decay * osc;
        }
        if (n == 0)
            jv = 2.0 * omega0;
        const double complex g = ev / jv;
        const double complex o = g * (tru[2 * n] + I * tru[2 * n + 1]);
        e_w[2 * n] = creal(ev);
        e_w[2 * n + 1] = cimag(ev);
        j_w[2 * n] = creal(jv);
        j_w[2 * n + 1] = cimag(jv);
        g_w[2 * n] = creal(g);
        g_w[2 * n + 1] = cimag(g);
        obs[2 * n] = creal(o);
        obs[2 * n + 1] = cimag(o);
    }
    free(tru);
}


/* ------------------------------------------------------------------------
 * The reference's FWI gradient chain (compat mode).
 */

/* laplaceToFreq_2 / laplaceToFreq_back per cell (fwi3d_cpu_function_v1.4.c:
 * 2166-2189, 1883-1894): G = E/J, and E_ff = G * out_True when tw is given.
 * acc_bins holds the volume DFT at bins startF..upperF-1: [3][upperF-startF][N];
 * out spans [3][upperF][N] and stays zero below startF. */
void or_laplace_volume(int n, int upperf, int startf, const double *acc_bins,
                       const double *jw, const double *tw, double *out)
{
    const int nb = upperf - startf;
    const double complex *A = (const double complex *)acc_bins;
    const double complex *J = (const double complex *)jw;
    const double complex *T = (const double complex *)tw;
    double complex *O = (double complex *)out;
    memset(out, 0, sizeof(double complex) * 3 * (size_t)upperf * n);
    for (int d = 0; d < 3; ++d)
        for (int b = 0; b < nb; ++b)
            for (int c = 0; c < n; ++c) {
                const double complex g = A[((size_t)d * nb + b) * n + c] / J[b];
                O[((size_t)d * upperf + startf + b) * n + c] = T ? g * T[b] : g;
            }
}

/* residualE (fwi3d_cpu_function_v1.4.c:2202-2244): dE = Eobs - E_ff at the
 * receivers; err += |dEx| + |dEy| + |dEz| over receivers, then bins */
void or_residual_e(int nrec, int upperf, const double *eobs, const double *effrec,
                   double *de, double *err_sum)
{
    const double complex *ob = (const double complex *)eobs;
    const double complex *ef = (const double complex *)effrec;
    double complex *dd = (double complex *)de;
    for (int irec = 0; irec < nrec; ++irec)
        for (int w = 0; w < upperf; ++w)
            for (int d = 0; d < 3; ++d) {
                const size_t q = ((size_t)d * nrec + irec) * upperf + w;
                dd[q] = ob[q] - ef[q];
                *err_sum += cabs(dd[q]);
            }
}

/* sensitivity (fwi3d_cpu_function_v1.4.c:2246-2290):
 * sense = sum_{irec, idir, w} |Ex_ff Ex_bb| + |Ey_ff Ey_bb| + |Ez_ff Ez_bb| */
void or_sensitivity(int n, int nrec, int upperf, const double *eff, const double *ebb,
                    double *sense)
{
    const double complex *F = (const double complex *)eff;
    const double complex *B = (const double complex *)ebb;
    const size_t vol = (size_t)upperf * n, gvol = (size_t)nrec * 3 * vol;
#pragma omp parallel for
    for (int c = 0; c < n; ++c) {
        double s = 0.0;
        for (int irec = 0; irec < nrec; ++irec)
            for (int idir = 0; idir < 3; ++idir)
                for (int w = 0; w < upperf; ++w) {
                    const size_t ijks = (size_t)w * n + c;
                    const size_t ijkr = ((size_t)irec * 3 + idir) * vol + (size_t)w * n + c;
                    for (int d = 0; d < 3; ++d)
                        s += cabs(F[d * vol + ijks] * B[d * gvol + ijkr]);
                }
        sense[c] = s;
    }
}

/* conv_GdelE (fwi3d_cpu_function_v1.4.c:2292-2349):
 * E_back,d(w) = sum_{irec, idir} G^{irec,idir}_d(w) conj(dE_idir(irec, w)) */
void or_conv_gdele(int n, int nrec, int upperf, const double *ebb, const double *de,
                   double *eback)
{
    const double complex *B = (const double complex *)ebb;
    const double complex *D = (const double complex *)de;
    double complex *E = (double complex *)eback;
    const size_t vol = (size_t)upperf * n, gvol = (size_t)nrec * 3 * vol;
    memset(eback, 0, sizeof(double complex) * 3 * vol);
#pragma omp parallel for
    for (int c = 0; c < n; ++c)
        for (int irec = 0; irec < nrec; ++irec)
            for (int idir = 0; idir < 3; ++idir)
                for (int w = 0; w < upperf; ++w) {
                    const size_t ijkw = (size_t)w * n + c;
                    const size_t ijkr = ((size_t)irec * 3 + idir) * vol + (size_t)w * n + c;
                    const double complex cd = conj(D[((size_t)idir * nrec + irec) * upperf + w]);
                    for (int d = 0; d < 3; ++d)
                        E[d * vol + ijkw] += B[d * gvol + ijkr] * cd;
                }
}

/* calc_gamma (fwi3d_cpu_function_v1.4.c:2351-2384): on ncpml <= i,j,k < n-ncpml,
 * grad += sum_w Re(Ex_ff Ex_back + Ey_ff Ey_back + Ez_ff Ez_back)/sense */
void or_calc_gamma(int ix, int iy, int iz, int nc, int upperf, const double *eff,
                   const double *eback, const double *sense, double *grad)
{
    const double complex *F = (const double complex *)eff;
    const double complex *E = (const double complex *)eback;
    const int n = ix * iy * iz;
    const size_t vol = (size_t)upperf * n;
#pragma omp parallel for
    for (int k = nc; k < iz - nc; ++k)
        for (int j = nc; j < iy - nc; ++j)
            for (int i = nc; i < ix - nc; ++i) {
                const int ijk = k * ix * iy + j * ix + i;
                for (int w = 0; w < upperf; ++w) {
                    const size_t q = (size_t)w * n + ijk;
                    grad[ijk] += creal(F[q] * E[q] + F[vol + q] * E[vol + q] +
                                       F[2 * vol + q] * E[2 * vol + q]) / sense[ijk];
                }
            }
}

/* calc_phi (fwi3d_cpu_function_v1.4.c:2565-2603) */
void or_calc_phi(int ix, int iy, int iz, int iseabed, const double *sig, const double *grad,
                 double *sig2, double *scaler)
{
    double sigmax2 = sig[0], gradmax = grad[0];
    for (int k = 0; k < iz; ++k)
        for (int j = 0; j < iy; ++j)
            for (int i = 0; i < ix; ++i) {
                const int ijk = k * ix * iy + j * ix + i;
                if (k <= iseabed) {
                    if (sig[ijk] > sigmax2)
                        sigmax2 = sig[ijk];
                    if (grad[ijk] > gradmax)
                        gradmax = grad[ijk];
                }
            }
    *scaler = 0.01f * sigmax2 / gradmax;
    for (int c = 0; c < ix * iy * iz; ++c)
        sig2[c] = sig[c] - *scaler * grad[c];
}

/* calc_alpha (fwi3d_cpu_function_v1.4.c:2757-2788), receiver arrays */
void or_calc_alpha(int nrec, int upperf, const double *de, const double *eff2,
                   const double *effl, double *k1, double *k2)
{
    const double complex *D = (const double complex *)de;
    const double complex *F2 = (const double complex *)eff2;
    const double complex *FL = (const double complex *)effl;
    double complex a = k1[0] + I * k1[1], b = k2[0] + I * k2[1];
    for (int irec = 0; irec < nrec; ++irec)
        for (int w = 0; w < upperf; ++w) {
            double complex t[3];
            for (int d = 0; d < 3; ++d) {
                const size_t q = ((size_t)d * nrec + irec) * upperf + w;
                t[d] = F2[q] - FL[q];
            }
            for (int d = 0; d < 3; ++d)
                a += D[((size_t)d * nrec + irec) * upperf + w] * t[d];
            for (int d = 0; d < 3; ++d)
                b += t[d] * t[d];
        }
    k1[0] = creal(a);
    k1[1] = cimag(a);
    k2[0] = creal(b);
    k2[1] = cimag(b);
}

/* update_para2 (fwi3d_cpu_function_v1.4.c:2790-2832) */
void or_update_para2(int ix, int iy, int iz, int nc, int iseabed, const double *k1,
                     const double *k2, double scaler, double minval, double maxval,
                     double *sig, double *grad, double *alpha)
{
    *alpha = 3.f * cabs(k1[0] + I * k1[1]) / cabs(k2[0] + I * k2[1]);
    for (int k = nc; k < iz - nc; ++k)
        for (int j = nc; j < iy - nc; ++j)
            for (int i = nc; i < ix - nc; ++i) {
                const int ijk = k * ix * iy + j * ix + i;
                if (k <= iseabed)
                    sig[ijk] += *alpha * scaler * grad[ijk];
                if (sig[ijk] < minval)
                    sig[ijk] = minval;
                if (sig[ijk] > maxval)
                    sig[ijk] = maxval;
            }
    memset(grad, 0, sizeof(double) * (size_t)ix * iy * iz);
}

/* ===================================================================== */
/* Fortran lineage (ver0040), restated operator by operator over the full */
/* volume in the Fortran loop order; 1-based Fortran ranges written as    */
/* 0-based half-open ones.  Pinned against the reference's own ver0040   */
/* subroutines compiled with amdflang (tests/test_oracle_fortran_ref.py)  */
/* and by mathematics (tests/test_oracle_fortran.py).                     */
/* ===================================================================== */

#define FIDX(i, j, k) (((long)(k) * iy + (j)) * ix + (i))

/* x / c for a complex x and a real c, as the reference's Fortran evaluates it
 * when compiled with amdflang (checked against that build in
 * tests/test_oracle_fortran_ref.py): c becomes the complex (c, 0) and the
 * quotient takes the textbook formula ((xr c + xi 0) / (c c + 0 0),
 * (xi c - xr 0) / (c c + 0 0)), i.e. (x c) / (c c) for each part of a finite
 * x.  That differs from x / c in the last place.  Every complex-by-real
 * division of the ver0040 stencils, psi updates and source takes it; the
 * oracle's real and imaginary runs apply it part by part. */
static inline double orf_cdivr(double x, double c)
{
    return (x * c) / (c * c);
}

/* x**e of the Fortran source with e a whole-number compile-time constant
 * (cpu_cpml4.f90:25-30: nn, order and ma are parameters): the compiler folds
 * it into the integer power, multiplied out by repeated squaring, as the
 * reference compiled with amdflang shows (tests/golden/ref_fortran_step.npz);
 * pow() differs from it in the last place. */
static double orf_pow(double x, double e)
{
    if (e != floor(e) || fabs(e) > 64.0)
        return pow(x, e);
    long k = (long)fabs(e);
    double r = 1.0, sq = x;
    int any = 0;
    while (k) {
        if (k & 1) {
            r = any ? r * sq : sq;
            any = 1;
        }
        k >>= 1;
        if (k)
            sq = sq * sq;
    }
    return e < 0 ? 1.0 / r : r;
}

/* ver0040/cpu_cpml4.f90:8-248: ramps on the 1-based index i = 1..ncpml,
 * sigma_max without a per-axis factor, high side = mirror of the low side */
void orf_init_cpml(or_state *s)
{
    const or_params *p = &s->p;
    const int nc = p->ncpml;
    for (int a = 0; a < 3; ++a) {
        const int n = axis_len(p, a);
        const double h = axis_step(p, a);
        const double delta = nc * h;
        const double smax = (p->nn + p->order + 1.0) * p->cmax * log(1.0 / p->rcoef) / (2.0 * delta);
        double **q = s->prof[a];
        for (int r = 0; r < n; ++r) {
            q[OR_P_ESIG][r] = q[OR_P_MSIG][r] = 0.0;
            q[OR_P_EKAP][r] = q[OR_P_MKAP][r] = 1.0;
            q[OR_P_AE][r] = q[OR_P_AM][r] = 0.0;
            q[OR_P_BE][r] = q[OR_P_BH][r] = q[OR_P_CE][r] = q[OR_P_CH][r] = 0.0;
        }
        for (int i = 1; i <= nc; ++i) {
            const int r = i - 1;
            const double den = (double)nc - 1.0;
            q[OR_P_ESIG][r] = smax * orf_pow(((double)nc - (double)i) / den, p->nn + p->order);
            q[OR_P_MSIG][r] = smax * orf_pow(((double)nc - (double)i - 0.5) / den, p->nn + p->order);
            q[OR_P_EKAP][r] = 1.0 + (p->kappa_max - 1.0) * orf_pow(((double)nc - (double)i) / den, p->nn);
            q[OR_P_MKAP][r] = 1.0 + (p->kappa_max - 1.0) * orf_pow(((double)nc - (double)i - 0.5) / den, p->nn);
            q[OR_P_AE][r] = p->alpha_max * orf_pow(((double)i) / den, p->ma);
            q[OR_P_AM][r] = p->alpha_max * orf_pow(((double)i + 0.5) / den, p->ma);
            q[OR_P_BE][r] = exp(-(q[OR_P_ESIG][r] / q[OR_P_EKAP][r] + q[OR_P_AE][r]) * p->dt);
            q[OR_P_BH][r] = exp(-(q[OR_P_MSIG][r] / q[OR_P_MKAP][r] + q[OR_P_AM][r]) * p->dt);
            q[OR_P_CE][r] = q[OR_P_ESIG][r] * (q[OR_P_BE][r] - 1.0) /
                            (q[OR_P_ESIG][r] + q[OR_P_EKAP][r] * q[OR_P_AE][r]) / q[OR_P_EKAP][r];
            q[OR_P_CH][r] = q[OR_P_MSIG][r] * (q[OR_P_BH][r] - 1.0) /
                            (q[OR_P_MSIG][r] + q[OR_P_MKAP][r] * q[OR_P_AM][r]) / q[OR_P_MKAP][r];
        }
        for (int i = 1; i <= nc; ++i)
            for (int id = 0; id < OR_NPROFILES; ++id)
                q[id][n - i] = q[id][i - 1];
        for (int r = 0; r < n; ++r) {
            q[OR_P_KED][r] = q[OR_P_EKAP][r] * h;
            q[OR_P_KHD][r] = q[OR_P_MKAP][r] * h;
        }
    }
}

/* ver0040/cpu_eh_field_cpml42.f90:8-57: coefficient volumes, per-cell mu in db */
void orf_media_coeff(or_state *s, const double *mu)
{
    const or_params *p = &s->p;
    const int ix = p->ix, iy = p->iy, iz = p->iz;
    const double dt = p->dt;
    for (int k = 0; k < iz; ++k)
        for (int j = 0; j < iy; ++j)
            for (int i = 0; i < ix; ++i) {
                const long c = FIDX(i, j, k);
                const double eps2 = s->vol[OR_SIG][c] / 2.0 / p->omega0;
                const int idx[3] = {i, j, k};
                for (int d = 0; d < 3; ++d) {
                    const double es = s->prof[d][OR_P_ESIG][idx[d]];
                    const double ms = s->prof[d][OR_P_MSIG][idx[d]];
                    s->vol[OR_CA_X + d][c] = (1.0 - ((es * dt) / (2.0 * eps2))) / (1.0 + ((es * dt) / (2.0 * eps2)));
                    s->vol[OR_CB_X + d][c] = dt / eps2 / (1.0 + (es * dt) / (2.0 * eps2));
                    s->vol[OR_DA_X + d][c] = (1.0 - ((ms * dt) / (2.0 * eps2))) / (1.0 + ((ms * dt) / (2.0 * eps2)));
                    s->vol[OR_DB_X + d][c] = dt / mu[c] / (1.0 + (ms * dt) / (2.0 * eps2));
                }
            }
}

/* ver0040/cpu_waveform.f90:20-26: out[n] = s(istep) with istep = n + 1,
 * s = -(2 beta (istep dt - t0) sqrt(beta/pi)) exp(-beta (istep dt - t0)**2),
 * t0 = pi/fmax, beta = pi fmax**2, pi the module's 3.14159265358979d0 */
void orf_waveform(int nstep, double dt, double fmax, double pi, double *out)
{
    const double t0 = pi / fmax;
    const double beta = pi * (fmax * fmax);
    for (int n = 0; n < nstep; ++n) {
        const double u = (double)(n + 1) * dt - t0;
        out[n] = -(2.0 * beta * u * sqrt(beta / pi)) * exp(-beta * (u * u));
    }
}

/* ver0040/cpu_waveform.f90:46-51: Hz(src) -= signal*dt/myu(src)/dx/dy/dz, each point in turn */
void orf_source_hz(or_state *s, int npts, const int *cells, const double *mu, double sval)
{
    const or_params *p = &s->p;
    for (int q = 0; q < npts; ++q) {
        const long c = cells[q];
        s->vol[OR_HZ][c] = s->vol[OR_HZ][c] - orf_cdivr(orf_cdivr(orf_cdivr(orf_cdivr(sval * p->dt, mu[c]), p->dx), p->dy), p->dz);
    }
}

/* ver0040/cpu_eh_field_cpml42.f90:116-156 */
void orf_e_field_cpml4(or_state *s)
{
    const or_params *p = &s->p;
    const int ix = p->ix, iy = p->iy, iz = p->iz;
    const long sy = ix, sz = (long)ix * iy;
    const double c1 = p->c1, c2 = p->c2;
    double *Ex = s->vol[OR_EX], *Ey = s->vol[OR_EY], *Ez = s->vol[OR_EZ];
    const double *Hx = s->vol[OR_HX], *Hy = s->vol[OR_HY], *Hz = s->vol[OR_HZ];
    const double *kedx = s->prof[0][OR_P_KED], *kedy = s->prof[1][OR_P_KED], *kedz = s->prof[2][OR_P_KED];
    for (int k = 2; k < iz - 1; ++k)
        for (int j = 2; j < iy - 1; ++j)
            for (int i = 0; i < ix; ++i) {
                const long c = FIDX(i, j, k);
                Ex[c] = s->vol[OR_CA_X][c] * Ex[c] +
                        s->vol[OR_CB_X][c] * (orf_cdivr((c1 * Hz[c] - c1 * Hz[c - sy] + c2 * Hz[c + sy] - c2 * Hz[c - 2 * sy]), kedy[j]) -
                                              orf_cdivr((c1 * Hy[c] - c1 * Hy[c - sz] + c2 * Hy[c + sz] - c2 * Hy[c - 2 * sz]), kedz[k]));
            }
    for (int k = 2; k < iz - 1; ++k)
        for (int j = 0; j < iy; ++j)
            for (int i = 2; i < ix - 1; ++i) {
                const long c = FIDX(i, j, k);
                Ey[c] = s->vol[OR_CA_Y][c] * Ey[c] +
                        s->vol[OR_CB_Y][c] * (orf_cdivr((c1 * Hx[c] - c1 * Hx[c - sz] + c2 * Hx[c + sz] - c2 * Hx[c - 2 * sz]), kedz[k]) -
                                              orf_cdivr((c1 * Hz[c] - c1 * Hz[c - 1] + c2 * Hz[c + 1] - c2 * Hz[c - 2]), kedx[i]));
            }
    for (int k = 0; k < iz; ++k)
        for (int j = 2; j < iy - 1; ++j)
            for (int i = 2; i < ix - 1; ++i) {
                const long c = FIDX(i, j, k);
                Ez[c] = s->vol[OR_CA_Z][c] * Ez[c] +
                        s->vol[OR_CB_Z][c] * (orf_cdivr((c1 * Hy[c] - c1 * Hy[c - 1] + c2 * Hy[c + 1] - c2 * Hy[c - 2]), kedx[i]) -
                                              orf_cdivr((c1 * Hx[c] - c1 * Hx[c - sy] + c2 * Hx[c + sy] - c2 * Hx[c - 2 * sy]), kedy[j]));
            }
}

/* ver0040/cpu_cpml4.f90:271-358: shells [3, ncpml+1] and [n-ncpml+1, n-1]
 * (1-based) on their axis, every other axis in full */
void orf_cpml_e4(or_state *s)
{
    const or_params *p = &s->p;
    const int ix = p->ix, iy = p->iy, iz = p->iz, nc = p->ncpml;
    const long sy = ix, sz = (long)ix * iy;
    const double c1 = p->c1, c2 = p->c2;
    double *Ex = s->vol[OR_EX], *Ey = s->vol[OR_EY], *Ez = s->vol[OR_EZ];
    const double *Hx = s->vol[OR_HX], *Hy = s->vol[OR_HY], *Hz = s->vol[OR_HZ];
    double *pEzx = s->vol[OR_PSI_EZX], *pEyx = s->vol[OR_PSI_EYX], *pExy = s->vol[OR_PSI_EXY];
    double *pEzy = s->vol[OR_PSI_EZY], *pEyz = s->vol[OR_PSI_EYZ], *pExz = s->vol[OR_PSI_EXZ];
    const double *bex = s->prof[0][OR_P_BE], *cex = s->prof[0][OR_P_CE];
    const double *bey = s->prof[1][OR_P_BE], *cey = s->prof[1][OR_P_CE];
    const double *bez = s->prof[2][OR_P_BE], *cez = s->prof[2][OR_P_CE];
    const int xr[2][2] = {{2, nc + 1}, {ix - nc, ix - 1}};
    for (int side = 0; side < 2; ++side)
        for (int k = 0; k < iz; ++k)
            for (int j = 0; j < iy; ++j)
                for (int i = xr[side][0]; i < xr[side][1]; ++i) {
                    const long c = FIDX(i, j, k);
                    pEzx[c] = bex[i] * pEzx[c] + orf_cdivr(cex[i] * (c1 * Hy[c] - c1 * Hy[c - 1] + c2 * Hy[c + 1] - c2 * Hy[c - 2]), p->dx);
                    pEyx[c] = bex[i] * pEyx[c] + orf_cdivr(cex[i] * (c1 * Hz[c] - c1 * Hz[c - 1] + c2 * Hz[c + 1] - c2 * Hz[c - 2]), p->dx);
                    Ez[c] = Ez[c] + s->vol[OR_CB_Z][c] * pEzx[c];
                    Ey[c] = Ey[c] - s->vol[OR_CB_Y][c] * pEyx[c];
                }
    const int yr[2][2] = {{2, nc + 1}, {iy - nc, iy - 1}};
    for (int side = 0; side < 2; ++side)
        for (int k = 0; k < iz; ++k)
            for (int j = yr[side][0]; j < yr[side][1]; ++j)
                for (int i = 0; i < ix; ++i) {
                    const long c = FIDX(i, j, k);
                    pExy[c] = bey[j] * pExy[c] + orf_cdivr(cey[j] * (c1 * Hz[c] - c1 * Hz[c - sy] + c2 * Hz[c + sy] - c2 * Hz[c - 2 * sy]), p->dy);
                    pEzy[c] = bey[j] * pEzy[c] + orf_cdivr(cey[j] * (c1 * Hx[c] - c1 * Hx[c - sy] + c2 * Hx[c + sy] - c2 * Hx[c - 2 * sy]), p->dy);
                    Ex[c] = Ex[c] + s->vol[OR_CB_X][c] * pExy[c];
                    Ez[c] = Ez[c] - s->vol[OR_CB_Z][c] * pEzy[c];
                }
    const int zr[2][2] = {{2, nc + 1}, {iz - nc, iz - 1}};
    for (int side = 0; side < 2; ++side)
        for (int k = zr[side][0]; k < zr[side][1]; ++k)
            for (int j = 0; j < iy; ++j)
                for (int i = 0; i < ix; ++i) {
                    const long c = FIDX(i, j, k);
                    pEyz[c] = bez[k] * pEyz[c] + orf_cdivr(cez[k] * (c1 * Hx[c] - c1 * Hx[c - sz] + c2 * Hx[c + sz] - c2 * Hx[c - 2 * sz]), p->dz);
                    pExz[c] = bez[k] * pExz[c] + orf_cdivr(cez[k] * (c1 * Hy[c] - c1 * Hy[c - sz] + c2 * Hy[c + sz] - c2 * Hy[c - 2 * sz]), p->dz);
                    Ey[c] = Ey[c] + s->vol[OR_CB_Y][c] * pEyz[c];
                    Ex[c] = Ex[c] - s->vol[OR_CB_X][c] * pExz[c];
                }
}

/* ver0040/cpu_eh_field_cpml42.f90:213-252 */
void orf_h_field_cpml4(or_state *s)
{
    const or_params *p = &s->p;
    const int ix = p->ix, iy = p->iy, iz = p->iz;
    const long sy = ix, sz = (long)ix * iy;
    const double c1 = p->c1, c2 = p->c2;
    const double *Ex = s->vol[OR_EX], *Ey = s->vol[OR_EY], *Ez = s->vol[OR_EZ];
    double *Hx = s->vol[OR_HX], *Hy = s->vol[OR_HY], *Hz = s->vol[OR_HZ];
    const double *khdx = s->prof[0][OR_P_KHD], *khdy = s->prof[1][OR_P_KHD], *khdz = s->prof[2][OR_P_KHD];
    for (int k = 1; k < iz - 2; ++k)
        for (int j = 1; j < iy - 2; ++j)
            for (int i = 0; i < ix; ++i) {
                const long c = FIDX(i, j, k);
                Hx[c] = s->vol[OR_DA_X][c] * Hx[c] -
                        s->vol[OR_DB_X][c] * (orf_cdivr((c1 * Ez[c + sy] - c1 * Ez[c] + c2 * Ez[c + 2 * sy] - c2 * Ez[c - sy]), khdy[j]) -
                                              orf_cdivr((c1 * Ey[c + sz] - c1 * Ey[c] + c2 * Ey[c + 2 * sz] - c2 * Ey[c - sz]), khdz[k]));
            }
    for (int k = 1; k < iz - 2; ++k)
        for (int j = 0; j < iy; ++j)
            for (int i = 1; i < ix - 2; ++i) {
                const long c = FIDX(i, j, k);
                Hy[c] = s->vol[OR_DA_Y][c] * Hy[c] -
                        s->vol[OR_DB_Y][c] * (orf_cdivr((c1 * Ex[c + sz] - c1 * Ex[c] + c2 * Ex[c + 2 * sz] - c2 * Ex[c - sz]), khdz[k]) -
                                              orf_cdivr((c1 * Ez[c + 1] - c1 * Ez[c] + c2 * Ez[c + 2] - c2 * Ez[c - 1]), khdx[i]));
            }
    for (int k = 0; k < iz; ++k)
        for (int j = 1; j < iy - 2; ++j)
            for (int i = 1; i < ix - 2; ++i) {
                const long c = FIDX(i, j, k);
                Hz[c] = s->vol[OR_DA_Z][c] * Hz[c] -
                        s->vol[OR_DB_Z][c] * (orf_cdivr((c1 * Ey[c + 1] - c1 * Ey[c] + c2 * Ey[c + 2] - c2 * Ey[c - 1]), khdx[i]) -
                                              orf_cdivr((c1 * Ex[c + sy] - c1 * Ex[c] + c2 * Ex[c + 2 * sy] - c2 * Ex[c - sy]), khdy[j]));
            }
}

/* ver0040/cpu_cpml4.f90:384-474: shells [2, ncpml] and [n-ncpml, n-2] (1-based);
 * psi_Hyx as written there, with c2 where the difference has c1 (:400, 413) */
void orf_cpml_h4(or_state *s)
{
    const or_params *p = &s->p;
    const int ix = p->ix, iy = p->iy, iz = p->iz, nc = p->ncpml;
    const long sy = ix, sz = (long)ix * iy;
    const double c1 = p->c1, c2 = p->c2;
    const double *Ex = s->vol[OR_EX], *Ey = s->vol[OR_EY], *Ez = s->vol[OR_EZ];
    double *Hx = s->vol[OR_HX], *Hy = s->vol[OR_HY], *Hz = s->vol[OR_HZ];
    double *pHzx = s->vol[OR_PSI_HZX], *pHyx = s->vol[OR_PSI_HYX], *pHxy = s->vol[OR_PSI_HXY];
    double *pHzy = s->vol[OR_PSI_HZY], *pHyz = s->vol[OR_PSI_HYZ], *pHxz = s->vol[OR_PSI_HXZ];
    const double *bhx = s->prof[0][OR_P_BH], *chx = s->prof[0][OR_P_CH];
    const double *bhy = s->prof[1][OR_P_BH], *chy = s->prof[1][OR_P_CH];
    const double *bhz = s->prof[2][OR_P_BH], *chz = s->prof[2][OR_P_CH];
    const int xr[2][2] = {{1, nc}, {ix - nc - 1, ix - 2}};
    for (int side = 0; side < 2; ++side)
        for (int k = 0; k < iz; ++k)
            for (int j = 0; j < iy; ++j)
                for (int i = xr[side][0]; i < xr[side][1]; ++i) {
                    const long c = FIDX(i, j, k);
                    pHzx[c] = bhx[i] * pHzx[c] + orf_cdivr(chx[i] * (c1 * Ey[c + 1] - c1 * Ey[c] + c2 * Ey[c + 2] - c2 * Ey[c - 1]), p->dx);
                    pHyx[c] = bhx[i] * pHyx[c] + orf_cdivr(chx[i] * (c1 * Ez[c + 1] - c2 * Ez[c] + c2 * Ez[c + 2] - c2 * Ez[c - 1]), p->dx);
                    Hz[c] = Hz[c] - s->vol[OR_DB_Z][c] * pHzx[c];
                    Hy[c] = Hy[c] + s->vol[OR_DB_Y][c] * pHyx[c];
                }
    const int yr[2][2] = {{1, nc}, {iy - nc - 1, iy - 2}};
    for (int side = 0; side < 2; ++side)
        for (int k = 0; k < iz; ++k)
            for (int j = yr[side][0]; j < yr[side][1]; ++j)
                for (int i = 0; i < ix; ++i) {
                    const long c = FIDX(i, j, k);
                    pHxy[c] = bhy[j] * pHxy[c] + orf_cdivr(chy[j] * (c1 * Ez[c + sy] - c1 * Ez[c] + c2 * Ez[c + 2 * sy] - c2 * Ez[c - sy]), p->dy);
                    pHzy[c] = bhy[j] * pHzy[c] + orf_cdivr(chy[j] * (c1 * Ex[c + sy] - c1 * Ex[c] + c2 * Ex[c + 2 * sy] - c2 * Ex[c - sy]), p->dy);
                    Hx[c] = Hx[c] - s->vol[OR_DB_X][c] * pHxy[c];
                    Hz[c] = Hz[c] + s->vol[OR_DB_Z][c] * pHzy[c];
                }
    const int zr[2][2] = {{1, nc}, {iz - nc - 1, iz - 2}};
    for (int side = 0; side < 2; ++side)
        for (int k = zr[side][0]; k < zr[side][1]; ++k)
            for (int j = 0; j < iy; ++j)
                for (int i = 0; i < ix; ++i) {
                    const long c = FIDX(i, j, k);
                    pHyz[c] = bhz[k] * pHyz[c] + orf_cdivr(chz[k] * (c1 * Ex[c + sz] - c1 * Ex[c] + c2 * Ex[c + 2 * sz] - c2 * Ex[c - sz]), p->dz);
                    pHxz[c] = bhz[k] * pHxz[c] + orf_cdivr(chz[k] * (c1 * Ey[c + sz] - c1 * Ey[c] + c2 * Ey[c + 2 * sz] - c2 * Ey[c - sz]), p->dz);
                    Hy[c] = Hy[c] - s->vol[OR_DB_Y][c] * pHyz[c];
                    Hx[c] = Hx[c] + s->vol[OR_DB_X][c] * pHxz[c];
                }
}

/* ver0040/cpu_main.f90:104-148: per step source, E, CPML_E, H, CPML_H, output;
 * signal[n] drives step n (the Fortran istep = n + 1); traces [6][nrec][it]
 * hold E and H at the receivers after step n (0-based cells) */
void orf_forward(or_state *s, int npts, const int *cells, const double *mu,
                 const double *signal, int nsteps, int nrec, const int *rec, int it,
                 double *traces)
{
    for (int n = 0; n < nsteps; ++n) {
        if (signal)
            orf_source_hz(s, npts, cells, mu, signal[n]);
        orf_e_field_cpml4(s);
        orf_cpml_e4(s);
        orf_h_field_cpml4(s);
        orf_cpml_h4(s);
        for (int c = 0; c < 6 && traces; ++c)
            for (int r = 0; r < nrec; ++r)
                traces[((long)c * nrec + r) * it + n] = s->vol[OR_EX + c][rec[r]];
    }
}
