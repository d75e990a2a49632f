/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Thin driver that links the reference's own, self-contained CPML and
 * lattice sources (satan_i/fwi3d_cpu_cpmlabc_v1.4.c, fwi3d_cpu_lattice_v1.4.c,
 * compiled where they lie under the reference tree by oracle/Makefile) into
 * oracle/_ref/libref_cpml.so, so that tests can run the reference's init_cpml,
 * lattice_time_3d and xe/ye/ze/xh/yh/zh_pml4 on the same inputs as the oracle.
 *
 * The reference keeps its state in process globals; their definitions come
 * from the reference's own header (fwi3d_cpu_alloli1_v1.4.h).  This file only
 * allocates them and copies data in and out.  The rest of the reference's hot
 * path (fwi3d_cpu_function_v1.4.c) needs fftw3.h, which this image lacks, so
 * it is not built.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "fwi3d_cpu_alloli1_v1.4.h"
#include "fwi3d_cpu_cpmlabc_v1.4.h"
#include "fwi3d_cpu_lattice_v1.4.h"

static double **vol_slot(int id)
{
    /* same ids as the oracle's or_array() for the arrays the CPML code touches */
    static double **tab[] = {
        &EX, &EY, &EZ, &HX, &HY, &HZ, &sig,
        &ca_x, &ca_y, &ca_z, &cb_x, &cb_y, &cb_z,
        &da_x, &da_y, &da_z, &db_x, &db_y, &db_z,
        &psi_Ezx1, &psi_Eyx1, &psi_Exy1, &psi_Ezy1, &psi_Eyz1, &psi_Exz1,
        &psi_Hyx1, &psi_Hzx1, &psi_Hzy1, &psi_Hxy1, &psi_Hxz1, &psi_Hyz1,
    };
    if (id < 0 || id >= (int)(sizeof(tab) / sizeof(tab[0])))
        return NULL;
    return tab[id];
}

static double **prof_slot(int axis, int id)
{
    static double **tab[3][12] = {
        {&esigx, &msigx, &ekappax, &mkappax, &aex, &amx, &be_x, &bh_x, &ce_x, &ch_x, &kedx, &khdx},
        {&esigy, &msigy, &ekappay, &mkappay, &aey, &amy, &be_y, &bh_y, &ce_y, &ch_y, &kedy, &khdy},
        {&esigz, &msigz, &ekappaz, &mkappaz, &aez, &amz, &be_z, &bh_z, &ce_z, &ch_z, &kedz, &khdz},
    };
    if (axis < 0 || axis > 2 || id < 0 || id >= 12)
        return NULL;
    return tab[axis][id];
}

int ref_setup(int nx_, int ny_, int nz_, int nc, double ddx, double ddy, double ddz,
              double omega0, double minv_, double maxv_)
{
    ix = nx_;
    iy = ny_;
    iz = nz_;
    ncpml = nc;
    dx = ddx;
    dy = ddy;
    dz = ddz;
    omega_0 = omega0;
    minval = minv_;
    maxval = maxv_;
    n = ix * iy * iz;
    for (int id = 0; vol_slot(id); ++id) {
        free(*vol_slot(id));
        *vol_slot(id) = (double *)calloc((size_t)n, sizeof(double));
        if (!*vol_slot(id))
            return -1;
    }
    free(vel);
    vel = (double *)calloc((size_t)n, sizeof(double));
    const int len[3] = {ix, iy, iz};
    for (int a = 0; a < 3; ++a)
        for (int id = 0; id < 12; ++id) {
            free(*prof_slot(a, id));
            *prof_slot(a, id) = (double *)calloc((size_t)len[a], sizeof(double));
            if (!*prof_slot(a, id))
                return -1;
        }
    return vel ? 0 : -1;
}

/* Runs the reference's lattice_time_3d (needs sig filled) */
double ref_lattice(double *cmax_out)
{
    lattice_time_3d();
    if (cmax_out)
        *cmax_out = cmax;
    return dt;
}

void ref_set_dt(double v) { dt = v; }
void ref_set_cmax(double v) { cmax = v; }

/* init_cpml writes its profile dumps under ./data; run it inside workdir */
int ref_init_cpml(const char *workdir)
{
    char old[4096];
    if (!getcwd(old, sizeof old))
        return -1;
    if (chdir(workdir) != 0)
        return -2;
    init_cpml();
    return chdir(old) == 0 ? 0 : -3;
}

void ref_epml(void) { epml_abcs4(EX, EY, EZ, HX, HY, HZ); }
void ref_hpml(void) { hpml_abcs4(EX, EY, EZ, HX, HY, HZ); }

int ref_put(int id, const double *src)
{
    double **s = vol_slot(id);
    if (!s)
        return -1;
    memcpy(*s, src, sizeof(double) * (size_t)n);
    return 0;
}

int ref_get(int id, double *dst)
{
    double **s = vol_slot(id);
    if (!s)
        return -1;
    memcpy(dst, *s, sizeof(double) * (size_t)n);
    return 0;
}

int ref_get_profile(int axis, int id, double *dst)
{
    double **s = prof_slot(axis, id);
    if (!s)
        return -1;
    const int len[3] = {ix, iy, iz};
    memcpy(dst, *s, sizeof(double) * (size_t)len[axis]);
    return 0;
}
