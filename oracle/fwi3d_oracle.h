/*
 * fwi3d_oracle.h -- CPU restatement of the reference's fictitious-wave-domain
 * FDTD hot path (satan_i C v1.4), used ONLY as test infrastructure:
 *   - the checker in tests/ and in __graft_entry__.smoke(),
 *   - the `cpu_baseline` leg of bench.py.
 * Nothing in the product (mittet_main_amd/) may include, link or call it.
 *
 * Layout: x fastest, ijk = k*ix*iy + j*ix + i (satan_i/fwi3d_cpu_function_v1.4.c:1200).
 * All arithmetic is FP64 and follows the reference's evaluation order so that a
 * build with -ffp-contract=off reproduces the reference bit for bit.
 *
 * Parity unpinned: the C update ranges along each component's own axis,
 * [1, n-1) for E and H (or_e_field_cpml42 / or_h_field_cpml42,
 * function.c:1201-1231, 1255-1285).  function.c cannot be built here (it
 * includes fftw3.h) and no closed form fixes them; everything else is pinned
 * against the reference's own cpmlabc.c, lattice.c, posiinf.c and ver0040
 * Fortran builds, or by closed forms and invariants (see tests/).
 */
#ifndef FWI3D_ORACLE_H
#define FWI3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_params {
    int ix, iy, iz, ncpml;
    double dx, dy, dz;
    double dt, omega0, mu0;
    double c1, c2;
    double cmax;                 /* sqrt(2*omega0/mu0/minval), lattice.c:36 */
    double rcoef, nn, order, ma; /* cpmlabc.c:13-18 */
    double kappa_max, alpha_max; /* cpmlabc.c:20-21 */
    double opt_to_max[3];        /* cpmlabc.c:22,33 (x only in the reference) */
} or_params;

typedef struct or_state or_state;

/* volume array ids for or_array() */
enum {
    OR_EX = 0, OR_EY, OR_EZ, OR_HX, OR_HY, OR_HZ,
    OR_SIG,
    OR_CA_X, OR_CA_Y, OR_CA_Z, OR_CB_X, OR_CB_Y, OR_CB_Z,
    OR_DA_X, OR_DA_Y, OR_DA_Z, OR_DB_X, OR_DB_Y, OR_DB_Z,
    /* psi: per axis a, [0] feeds the "+" component, [1] the "-" one */
    OR_PSI_EZX, OR_PSI_EYX, OR_PSI_EXY, OR_PSI_EZY, OR_PSI_EYZ, OR_PSI_EXZ,
    OR_PSI_HYX, OR_PSI_HZX, OR_PSI_HZY, OR_PSI_HXY, OR_PSI_HXZ, OR_PSI_HYZ,
    OR_NARRAYS
};
/* 1-D profile ids for or_profile() */
enum {
    OR_P_ESIG = 0, OR_P_MSIG, OR_P_EKAP, OR_P_MKAP, OR_P_AE, OR_P_AM,
    OR_P_BE, OR_P_BH, OR_P_CE, OR_P_CH, OR_P_KED, OR_P_KHD, OR_NPROFILES
};

double or_lattice_time_3d(double minval, double omega0, double mu0,
                          double dx, double dy, double dz, double factor,
                          double *cmax_out);

or_state *or_create(const or_params *p);
void or_destroy(or_state *s);
double *or_array(or_state *s, int id);
double *or_profile(or_state *s, int axis, int id);
int or_num_threads(void);
void or_set_num_threads(int n);
double or_stream_triad(long n, int reps);

void or_init_cpml(or_state *s);
void or_media_coeff_3d(or_state *s);
void or_set_zero_eh(or_state *s);
void or_e_field_cpml42(or_state *s, int sign);
void or_h_field_cpml42(or_state *s, int sign);
void or_epml_abcs4(or_state *s);
void or_hpml_abcs4(or_state *s);
void or_read_source_3d(or_state *s, int comp, const double *signal,
                       int jx, int jy, int jz, int step);
void or_output(or_state *s, int nrec, const int *rec1, int it, int step,
               double *tx, double *ty, double *tz);

/* rec2rec-style driver: steps [step0, step0+nsteps) of E, CPML-E, source, H,
 * CPML-H, receiver gather, plus an optional running volume DFT of E
 * (nf > 0: acc is complex128 [3][nf][N], w is complex128 [it][nf]). */
void or_forward(or_state *s, int comp, const double *signal,
                int jx, int jy, int jz, int step0, int nsteps,
                int nrec, const int *rec1, int it,
                double *tx, double *ty, double *tz,
                int nf, const double *w, double *acc);

void or_firstderiv_gauss(int nt, double dt, double fmax_w, double *out);
void or_ricker(int nt, double dt, double *out);
void or_dft_weights(int it, double dt, double omega0, int nf,
                    const double *omega, double *w);
void or_fft_backward_real(int n, const double *in, double *out);
void or_laplace_to_freq(int it, double dt, double omega0, int nbins,
                        const double *trace, const double *jsig,
                        const double *strue,
                        double *e_w, double *j_w, double *g_w, double *obs);

/* ---- the reference's FWI gradient chain (compat mode), literal loops ----
 * Spectra arrays are complex128 over all bins w in [0, upperF) (zero below
 * startF, as the reference leaves them): volume [3][upperF][N], Green's
 * functions [3 comps][nrec][3 dirs][upperF][N], receiver arrays
 * [3][nrec][upperF]. */
void or_laplace_volume(int n, int upperf, int startf, const double *acc_bins,
                       const double *jw, const double *tw, double *out);
void or_residual_e(int nrec, int upperf, const double *eobs, const double *effrec,
                   double *de, double *err_sum);
void or_sensitivity(int n, int nrec, int upperf, const double *eff, const double *ebb,
                    double *sense);
void or_conv_gdele(int n, int nrec, int upperf, const double *ebb, const double *de,
                   double *eback);
void or_calc_gamma(int ix, int iy, int iz, int nc, int upperf, const double *eff,
                   const double *eback, const double *sense, double *grad);
void or_calc_phi(int ix, int iy, int iz, int iseabed, const double *sig, const double *grad,
                 double *sig2, double *scaler);
void or_calc_alpha(int nrec, int upperf, const double *de, const double *eff2,
                   const double *effl, double *k1, double *k2);
void or_update_para2(int ix, int iy, int iz, int nc, int iseabed, const double *k1,
                     const double *k2, double scaler, double minval, double maxval,
                     double *sig, double *grad, double *alpha);

/* Fortran lineage (ver0040): profiles, coefficients with per-cell mu, the
 * Hz source points, E/H updates and CPML in the Fortran ranges, the step loop.
 * Source and receiver cells are 0-based flat indices. */
void orf_init_cpml(or_state *s);
void orf_media_coeff(or_state *s, const double *mu);
void orf_waveform(int nstep, double dt, double fmax, double pi, double *out);
void orf_source_hz(or_state *s, int npts, const int *cells, const double *mu, double sval);
void orf_e_field_cpml4(or_state *s);
void orf_cpml_e4(or_state *s);
void orf_h_field_cpml4(or_state *s);
void orf_cpml_h4(or_state *s);
void orf_forward(or_state *s, int npts, const int *cells, const double *mu,
                 const double *signal, int nsteps, int nrec, const int *rec, int it,
                 double *traces);

#ifdef __cplusplus
}
#endif
#endif
