/*
 * mittet_fdtd.h -- C ABI of the MI355X fictitious-wave-domain FDTD library
 * (libmittet_fdtd.so, hand-written HIP for gfx950).
 *
 * The reference (satan_i C v1.4) exposes its hot path as void functions over
 * process globals, e.g. `void e_field_cpml42(double *EX, ..., double *HZ)`
 * (satan_i/fwi3d_cpu_function_v1.4.h:1-72, fwi3d_cpu_cpmlabc_v1.4.h:1-17,
 * fwi3d_cpu_lattice_v1.4.h:1).  This ABI keeps the same operator names and
 * argument meaning, made re-entrant: the reference's globals live in an opaque
 * model (grid, conductivity, CPML profiles) and shot (fields, psi shells,
 * receivers, transforms) owned by the library in HBM.
 *
 * Conventions
 *   - every entry point returns an int status (FWI3D_OK == 0, negative on error;
 *     fwi3d_last_error() gives the message).  Nothing calls exit().
 *   - cell coordinates are 0-based (i, j, k), x fastest: c = (k*iy + j)*ix + i.
 *     The reference's 1-based config coordinates map as i = jx-1 (function.c:524).
 *   - host buffers are caller-owned; device buffers are library-owned.
 *   - all arithmetic is FP64.  Complex arrays are interleaved (re, im) doubles.
 *   - work is queued on the stream given at creation (NULL: the library's own
 *     stream); results are valid after fwi3d_*_sync or a get_* call.
 *   - no file I/O.
 */
#ifndef MITTET_FDTD_H
#define MITTET_FDTD_H

#ifdef __cplusplus
extern "C" {
#endif

#define FWI3D_OK 0
#define FWI3D_EINVAL (-1)   /* bad argument or configuration */
#define FWI3D_ENOMEM (-2)   /* device or host allocation failed */
#define FWI3D_EDEVICE (-3)  /* HIP runtime error */
#define FWI3D_ESTATE (-4)   /* call not valid in the object's current state */
#define FWI3D_EBUDGET (-5)  /* request exceeds the admission budget */

/* src_comp of a Fortran-variant shot: magnetic source on Hz */
#define FWI3D_SRC_HZ 5

/* Field / trace component ids */
enum { FWI3D_EX = 0, FWI3D_EY, FWI3D_EZ, FWI3D_HX, FWI3D_HY, FWI3D_HZ };

/* CPML profile ids for fwi3d_model_get_profile (satan_i/fwi3d_cpu_cpmlabc_v1.4.c:40-53) */
enum {
    FWI3D_P_ESIG = 0, FWI3D_P_MSIG, FWI3D_P_EKAP, FWI3D_P_MKAP, FWI3D_P_AE, FWI3D_P_AM,
    FWI3D_P_BE, FWI3D_P_BH, FWI3D_P_CE, FWI3D_P_CH, FWI3D_P_KED, FWI3D_P_KHD,
    FWI3D_NPROFILES
};

/* Every scalar the reference hard-codes or reads from data/model_env.dat. */
typedef struct fwi3d_grid_params {
    int ix, iy, iz;        /* cells per axis, CPML included */
    int ncpml;             /* CPML thickness in cells */
    double dx, dy, dz;     /* m */
    double dt;             /* s (see fwi3d_lattice_time_3d) */
    double omega0;         /* fictitious-domain reference angular frequency, rad/s */
    double mu0;            /* 12.5663706e-7 in the C reference (alloex1.h:3) */
    double c1, c2;         /* 4th-order staggered coefficients (1.125, (double)-0.04167f) */
    double cmax;           /* sqrt(2*omega0/mu0/sigma_min): sets the CPML sigma_max */
    double rcoef;          /* CPML reflection coefficient ((double)0.01f) */
    double nn, order, ma;  /* grading exponents (3, 0, 1) */
    double kappa_max;      /* 1 */
    double alpha_max;      /* (double)0.2f */
    double opt_to_max[3];  /* per-axis sigma_max factor ({10, 1, 1} in the reference) */
} fwi3d_grid_params;

/* One forward-modelling shot. */
typedef struct fwi3d_shot_params {
    int it;                  /* time samples: the loop runs it-1 steps (rec2rec.c:48) */
    int src_comp;            /* 0/1/2: electric dipole along x/y/z */
    int src_i, src_j, src_k; /* 0-based source cell */
    const double *signal;    /* host, length it: source waveform s[n] */
    int nrec;                /* receivers */
    const int *rec_ijk;      /* host, 3*nrec 0-based cells */
    int nf_vol;              /* volume E running-DFT frequencies (0: off) */
    const double *omega_vol; /* host, nf_vol angular frequencies (rad/s) */
    int ring;                /* E history slots between DFT flushes (0: library default) */
    int zchunk;              /* z-cells per thread in the fused kernels (0: library default) */
} fwi3d_shot_params;

typedef struct fwi3d_model fwi3d_model;
typedef struct fwi3d_shot fwi3d_shot;
typedef struct fwi3d_batch fwi3d_batch;

const char *fwi3d_last_error(void);
const char *fwi3d_version(void);

/* ---- setup (L1) ------------------------------------------------------- */

/* satan_i/fwi3d_cpu_lattice_v1.4.c:6-54: cmax = sqrt(2*omega0/mu0/minval),
 * dt = courant*6/7*factor (factor (double)0.999f in the reference). */
double fwi3d_lattice_time_3d(double minval, double omega0, double mu0,
                             double dx, double dy, double dz, double factor,
                             double *cmax_out);

/* satan_i/fwi3d_cpu_function_v1.4.c:1392-1405 (fmax_check): the highest
 * frequency the grid resolves, fmax = cmin/glim/max(dx, dy, dz) with
 * cmin = sqrt(2*omega0/mu0/maxval) (fwi3d_cpu_lattice_v1.4.c:36).  Returns
 * FWI3D_EINVAL when fmax < fmax_w (the reference prints and exits).
 * fmax_out: fmax, may be NULL.  The reference's maxin3 (function.c:1407)
 * returns the largest spacing as an int, which truncates spacings below one
 * metre to 0; the spacing stays a double here.  Its time_check
 * (function.c:1374-1389) multiplies by the global nx, which nothing assigns,
 * so it never fires; it has no counterpart here. */
int fwi3d_fmax_check(double maxval, double omega0, double mu0, double glim,
                     double dx, double dy, double dz, double fmax_w, double *fmax_out);
/* Largest stable dt of the leapfrog on a model: 1/(c_max sqrt(1/dx^2 + 1/dy^2
 * + 1/dz^2)) times 6/7 (variant 0, the C reference's courant*6/7,
 * fwi3d_cpu_lattice_v1.4.c:38-39) or 1/1.16667 (variant 1,
 * ver0040/cpu_parameter_wafe.f90:118), c_max = sqrt(2*omega0/min(mu*sigma)).
 * mu: per-cell permeability or NULL (mu0).  fwi3d_media_coeff_3d and
 * fwi3d_model_set_mu refuse (FWI3D_EINVAL) a model whose limit is below the
 * grid's dt, since such a run diverges.  -1 on a bad argument. */
double fwi3d_stable_dt(const fwi3d_grid_params *p, int variant, const double *sigma,
                       const double *mu);

/* Creates the grid, computes the CPML profiles (replaces init_cpml,
 * satan_i/fwi3d_cpu_cpmlabc_v1.4.c:7-175).  stream: hipStream_t or NULL. */
int fwi3d_model_create(const fwi3d_grid_params *p, void *stream, fwi3d_model **out);
int fwi3d_model_destroy(fwi3d_model *m);
/* Variant 1 is the Fortran lineage (ver0040), beside the C v1.4 semantics of
 * variant 0 (fwi3d_model_create):
 *   profiles of ver0040/cpu_cpml4.f90:8-248 (1-based ramps, high side mirrored;
 *     pass opt_to_max {1,1,1}, rcoef 0.01, alpha_max 0.2, c2 -0.04167 for the
 *     Fortran constants, cpu_parameter_wafe.f90:35-37, 127-131);
 *   per-cell permeability in the H coefficients (cpu_eh_field_cpml42.f90:51-53,
 *     fwi3d_model_set_mu; mu0 everywhere until set);
 *   shots with the magnetic source (src_comp FWI3D_SRC_HZ): Hz -= s[n] dt/mu/dx/dy/dz
 *     at up to 4 points before the E update of step n (cpu_waveform.f90:46-51,
 *     cpu_main.f90:104-148; step n takes s[n], the Fortran istep = n + 1);
 *   E/H updates over the Fortran loop ranges and CPML over the Fortran shells,
 *     psi_Hyx stencil as written there (cpu_eh_field_cpml42.f90:116-252,
 *     cpu_cpml4.f90:271-474).
 * Complex fields (the Fortran arrays are complex(8)) are two real runs: every
 * coefficient is real, so the real and imaginary parts never mix. */
int fwi3d_model_create_variant(const fwi3d_grid_params *p, int variant, void *stream,
                               fwi3d_model **out);
int fwi3d_model_set_mu(fwi3d_model *m, const double *mu);

/* host copy of one 1-D profile (length ix/iy/iz) */
int fwi3d_model_get_profile(const fwi3d_model *m, int axis, int which, double *out);

/* Uploads sigma (ix*iy*iz, S/m, > 0) and derives eps' = sigma/2/omega0 on the
 * device.  Replaces media_coeff_3d (satan_i/fwi3d_cpu_function_v1.4.c:325-435):
 * the 12 coefficient volumes are recomputed in-kernel from eps' and the 1-D
 * profiles instead of being stored.  FWI3D_EINVAL, leaving the model as it
 * was, when the grid's dt exceeds the model's stability limit
 * (fwi3d_stable_dt). */
int fwi3d_media_coeff_3d(fwi3d_model *m, const double *sigma);

/* ---- shots --------------------------------------------------------------- */

int fwi3d_shot_create(fwi3d_model *m, const fwi3d_shot_params *sp, fwi3d_shot **out);
int fwi3d_shot_destroy(fwi3d_shot *s);
/* Fortran variant: one more Hz source point with the shot's waveform (the
 * second transmitter of ver0040/cpu_waveform.f90:49-51); at most 4 points. */
int fwi3d_shot_add_source(fwi3d_shot *s, int i, int j, int k);
/* device bytes this shot would allocate (admission control) */
long long fwi3d_shot_bytes(const fwi3d_grid_params *p, const fwi3d_shot_params *sp);

/* satan_i/fwi3d_cpu_function_v1.4.c:964-986 (+ traces, DFT accumulators, step counter) */
int fwi3d_set_zero_eh(fwi3d_shot *s);

/* ---- per-step operators (L2), same names and order as the reference loop
 * satan_i/fwi3d_cpu_rec2rec_v1.4.c:49-54.  sign = +1 forward, -1 for the
 * sign-flipped *_bp variants (function.c:1292-1372). */
int fwi3d_e_field_cpml42(fwi3d_shot *s, int sign);  /* function.c:1184-1236 */
int fwi3d_epml_abcs4(fwi3d_shot *s);                /* cpmlabc.c:191-196 */
int fwi3d_read_source_3d(fwi3d_shot *s, int step);  /* function.c:514-529 */
int fwi3d_h_field_cpml42(fwi3d_shot *s, int sign);  /* function.c:1238-1290 */
int fwi3d_hpml_abcs4(fwi3d_shot *s);                /* cpmlabc.c:198-203 */
int fwi3d_output(fwi3d_shot *s, int step);          /* function.c:765-785 (traces) */

/* Fused steps: E+CPML-E+source, H+CPML-H, receivers, volume DFT; advances the
 * shot's own step counter by nsteps (bounded by it-1). */
int fwi3d_step_n(fwi3d_shot *s, int nsteps);
/* the whole it-1 step loop of one shot (rec2rec.c:48-57), graph-captured */
int fwi3d_run(fwi3d_shot *s);
int fwi3d_shot_sync(fwi3d_shot *s);
/* Runs nsteps fused steps with HIP events around every launch on the shot's
 * stream and returns the mean duration in ms of each kernel kind (ms holds 9
 * doubles): ms[0] E+CPML-E+source, ms[1] H+CPML-H, ms[2] receiver gather,
 * ms[3] volume-DFT flush (per flush launch), ms[4] whole step (per step, every
 * launch included), ms[5] H ring store and ms[6] H-sum flush (shots with
 * fwi3d_enable_volume_dft_h; else 0), ms[7] the fused E+H launch over the
 * CPML-free interior and ms[8] 1.0 when the shot runs it (then ms[0] and
 * ms[1] cover the CPML shell only), else 0.
 * Advances the shot like fwi3d_step_n.  No counterpart in the reference
 * (its timing is a clock() around main, rec2rec.c:18,65). */
int fwi3d_profile_steps(fwi3d_shot *s, int nsteps, double *ms);
int fwi3d_shot_step(const fwi3d_shot *s); /* steps done so far */
/* Divergence flag, with no counterpart in the reference (which runs on
 * whatever its fields do): 1 once a receiver sample or the source cell's E or
 * H has turned NaN or infinite since fwi3d_set_zero_eh (checked at every
 * receiver gather and at the end of every chunk of steps), else 0; negative
 * on error.  Waits for the shot's (or its batch's) stream. */
int fwi3d_shot_diverged(fwi3d_shot *s);
/* Diagnostic with no counterpart in the reference: copies `bytes` bytes between
 * two device buffers `reps` times with a plain streaming kernel (k_diag_copy),
 * so hardware byte counters can be calibrated against a known transfer.
 * ms_out: mean kernel time. */
int fwi3d_diag_copy(long long bytes, int reps, double *ms_out);

/* ---- readback -------------------------------------------------------------- */

/* which: FWI3D_EX..FWI3D_HZ; out: ix*iy*iz doubles */
int fwi3d_get_field(fwi3d_shot *s, int which, double *out);
int fwi3d_set_field(fwi3d_shot *s, int which, const double *in);
/* traces: [6][nrec][it] doubles, E then H, sample n = after step n */
int fwi3d_get_traces(fwi3d_shot *s, double *out);
/* volume E spectra: complex [3][nf_vol][ix*iy*iz] = sum_n E[n]*w_n(omega) */
int fwi3d_get_volume_dft(fwi3d_shot *s, double *out);
/* Volume H spectra at the shot's nf_vol frequencies, sum_n H[n]*w_n(omega)
 * with H after step n (the sample of the H traces).  The reference transforms
 * H at the receivers only (ver0040/cpu_f_to_d_h.f90:134-167); this carries
 * the volume E sums of laplaceToFreq_2 (satan_i/fwi3d_cpu_function_v1.4.c:
 * 2091-2200) over to H.  The f_to_d_h factor sqrt(-2 omega0/(i omega)) is the
 * caller's (fwi3d_f_to_d_h).  Off by default: H is updated in place, so
 * enabling allocates a second ring of the shot's ring slots (ring*3*8 bytes
 * per cell) that takes a copy of H each step and is summed once per ring like
 * the E slots.  Enable before the first step; not in a batch. */
int fwi3d_enable_volume_dft_h(fwi3d_shot *s);
int fwi3d_get_volume_dft_h(fwi3d_shot *s, double *out);
/* Device bytes fwi3d_enable_volume_dft_h adds to a shot of these parameters,
 * beyond fwi3d_shot_bytes: the H ring (ring*3*N doubles) and the H
 * accumulators (3*nf_vol*N complex).  0 when nf_vol is 0 (H sums need a
 * volume DFT); -1 on a null argument.  Admission control adds the two. */
long long fwi3d_volume_dft_h_bytes(const fwi3d_grid_params *p, const fwi3d_shot_params *sp);

/* ---- fictitious -> diffusive transform (L3) ------------------------------- */

/* Receiver spectra at arbitrary angular frequencies, one wavefront reduction
 * per (component, receiver, frequency):
 *   X(omega) = sum_{n<it} trace[n] * dt*exp(-a n dt)*exp(i a n dt), a = sqrt(omega0*omega).
 * comps: bit mask of FWI3D_EX..FWI3D_HZ; out complex [6][nrec][nfreq] (unused comps 0).
 * Magnetic components get the f_to_d_h factor sqrt(-2*omega0/(i*omega))
 * (ver0040/cpu_f_to_d_h.f90:146-148) when h_factor != 0; at omega = 0 (the
 * reference's bin k = 0, where it divides by zero) they are NaN. */
int fwi3d_receiver_dft(fwi3d_shot *s, int nfreq, const double *omegas, int h_factor,
                       double *out);
/* No counterpart in the reference: fwi3d_receiver_dft for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  The same kernel and the same results,
 * with the frequencies and the spectra staged in a fixed device array of
 * FWI3D_NOALLOC_SCRATCH_DOUBLES doubles that each GPU holds until the process
 * ends, instead of buffers allocated per call.  A request needs
 * 32 * ceil(nfreq / 32) + 12 * nrec * nfreq doubles (nrec: the shot's
 * receivers); a larger one is refused with FWI3D_EINVAL.  Calls that overlap
 * on one GPU share that array and race: the caller serialises them. */
#define FWI3D_NOALLOC_SCRATCH_DOUBLES (1 << 19)
int fwi3d_receiver_dft_noalloc(fwi3d_shot *s, int nfreq, const double *omegas, int h_factor,
                               double *out);
/* f_to_d_e / f_to_d_h on already-accumulated spectra (host arrays, complex):
 * E passes through; H is multiplied by sqrt(-2*omega0/(i*omega)). */
int fwi3d_f_to_d_h(double omega0, int nfreq, const double *omegas, long long count,
                   double *spectra);
/* Source spectrum of the waveform at the same frequencies with the electric-source
 * factor: J(omega) = csqrt(-2*omega0/(i*omega)) * sum J[n] w_n, J(0) := 2*omega0
 * (satan_i/fwi3d_cpu_function_v1.4.c:2044-2050).  out complex [nfreq]. */
int fwi3d_source_spectrum(double dt, double omega0, int it, const double *signal,
                          int nfreq, const double *omegas, double *out);
/* satan_i/fwi3d_cpu_function_v1.4.c:1976-2089 (laplaceToFreq) for every receiver
 * component of the shot at bins omega_n = 2*pi*n/(it*dt), n < nbins:
 * E_n, G_n = E_n/J_n and OBS_n = G_n * FFTW_BACKWARD(strue)[n].
 * e_w, g_w, obs: complex [3][nrec][nbins]; j_w: complex [nbins]. */
int fwi3d_laplace_to_freq(fwi3d_shot *s, int nbins, const double *strue,
                          double *e_w, double *j_w, double *g_w, double *obs);

/* ---- diffusive-time back-transform ----------------------------------------- */

/* out[r][k] = scale * sum_n in[r][n] * exp(sign * 2 pi i k n / nd) for complex
 * rows (interleaved re, im), rows <= 65535, sign +1 or -1, on the GPU (host
 * buffers in and out).  With sign -1 and scale 2/(nd dt) this is the
 * reference's back-transform to diffusive time with FFTW_FORWARD
 * (satan_i/fwi3d_cpu_function_v1.4.c:1507-1527, fwi3d_cpu_convert_v1.4.c:84-93);
 * with sign +1 the one of ver0040/cpu_f_to_d_e.f90:281-320 (FFTW_BACKWARD);
 * with scale 1 a plain un-normalised FFTW transform. */
int fwi3d_dft_rows(const double *in, int rows, int nd, int sign, double scale, double *out);
/* satan_i/fwi3d_cpu_function_v1.4.c:1430-1564 (laplace_fft) for every receiver
 * component of the shot: E_n, J_n (J_0 := 2 omega0) and G_n = E_n/J_n at all
 * bins n < it, then the back-transform x_t[k] = 2/(it dt) sum_n X_n
 * exp(-2 pi i k n / it).  e_t, g_t: complex [3][nrec][it]; j_t: complex [it]. */
int fwi3d_laplace_fft(fwi3d_shot *s, double *e_t, double *j_t, double *g_t);
/* satan_i/fwi3d_cpu_function_v1.4.c:1566-1635 (convolution_GJ_to_E): the
 * spectrum of E as the product FFTW_BACKWARD(G_t)[k] * FFTW_BACKWARD(J_t)[k]
 * for each of `rows` complex rows g_t [rows][nd] and the real source j_t [nd].
 * out: complex [rows][nd]. */
int fwi3d_convolve_gj(const double *g_t, int rows, int nd, const double *j_t, double *out);

/* ---- differentiation ------------------------------------------------------- */

/* Exact discrete adjoint of the forward step for a receiver misfit.  No
 * counterpart in the reference, whose gradient (fwi3d_cpu_inv_fwi_v1.4.c:36-85,
 * function.c:2246-2384) is a reciprocity correlation and not the derivative of
 * its discrete scheme.  The shot must have no volume DFT (nf_vol = 0).
 *   fwi3d_grad_create: checkpoint every `segment` steps (0: library default).
 *   fwi3d_grad_forward: zeroes the shot and runs its whole time loop, storing
 *       checkpoints; traces are then readable with fwi3d_get_traces.
 *   fwi3d_grad_backward: adjoint sweep driven by the receiver-trace sensitivities
 *       g = dJ/dE_trace, layout [3][nrec][it] (sample n = after step n); replays
 *       each segment forward from its checkpoint.  out: dJ/dsigma per cell. */
typedef struct fwi3d_grad fwi3d_grad;
int fwi3d_grad_create(fwi3d_shot *s, int segment, fwi3d_grad **out);
/* Memory-bounded variant, also without a reference counterpart (the reference
 * keeps every time step of every Green's run in host memory, function.c:1652-1724).
 * Recursive recomputation over `levels` depths (1..8): the forward pass stores
 * ~k top checkpoints, each depth below re-runs one segment to store the starts
 * of its k pieces, the last depth replays k steps keeping every state;
 * k = ceil(Nt^(1/(levels+1))), so storage is ~(levels+1) k states for
 * levels + 1 forward sweeps.  levels = 1 is fwi3d_grad_create(s, 0, out); the
 * gradient is bitwise the same for every level count.
 * fwi3d_grad_plan_bytes: device bytes such a gradient would allocate
 * (-1 on bad arguments), for admission control before creating it. */
int fwi3d_grad_create_levels(fwi3d_shot *s, int levels, fwi3d_grad **out);
/* The plan for nt steps (host only): seglen[levels] segment lengths per depth,
 * the top segment count and the number of stored states (incl. the adjoint). */
int fwi3d_grad_plan(int nt, int levels, int *seglen, int *nseg, long long *states);
long long fwi3d_grad_plan_bytes(fwi3d_shot *s, int levels);
int fwi3d_grad_destroy(fwi3d_grad *g);
long long fwi3d_grad_bytes(fwi3d_grad *g);
int fwi3d_grad_forward(fwi3d_grad *g);
int fwi3d_grad_backward(fwi3d_grad *g, const double *dj_dtrace, double *dj_dsigma);

/* ---- the reference's FWI iteration (compat mode) --------------------------- */

/* One iteration of satan_i/fwi3d_cpu_inv_fwi_v1.4.c:27-109 on the GPU.  The
 * frequency bins are the reference's w in [start_f, upper_f), omega_w =
 * w * 2 pi/(it dt) (fwi3d_compat_omegas); every shot passed in must run on the
 * object's model with a volume DFT at exactly those omegas.  Complex arrays are
 * interleaved (re, im); receiver arrays are [3][nrec][upper_f], zero below
 * start_f as in the reference.  Order of a call sequence: reset, source (each
 * shot), green (each receiver x direction), gamma, phi; then forward runs on
 * the returned test model, receivers + alpha (each shot), update.
 *   create: J at the bins from signal_x (csqrt(-2w0/(i w)) factor), T = the
 *           FFTW_BACKWARD of signal_true at the bins   (function.c:2091-2200)
 *   reset:  init_iterate, grad = E_back = 0            (function.c:2386-2413)
 *   source: laplaceToFreq_2 + residualE for source isrc: E_ff = (E/J) T over
 *           the volume, dE = eobs - E_ff at the receivers, err_sum += sum|dE|
 *                                                       (function.c:2091-2244)
 *   green:  laplaceToFreq_back of a run with its source at receiver irec,
 *           direction idir (backpropagation, function.c:2516-2526, 1821-1909),
 *           streamed into sensitivity + conv_GdelE of every source
 *                                                       (function.c:2246-2349)
 *   gamma:  calc_gamma, grad += sum Re(E_ff . E_back)/sense over the cells
 *           ncpml <= i,j,k < n - ncpml                  (function.c:2351-2384)
 *   phi:    calc_phi: scaler = 0.01 max sigma/max grad over k <= iseabed,
 *           sig2 = sigma - scaler grad                  (function.c:2565-2603)
 *   receivers: laplaceToFreq_3 at the receivers (E_ff2) (function.c:2657-2755)
 *   alpha:  calc_alpha: k1 += dE (E_ff2 - E_ff), k2 += (E_ff2 - E_ff)^2
 *                                                       (function.c:2757-2788)
 *   update: update_para2: alpha = 3|k1|/|k2|, sigma += alpha scaler grad where
 *           k <= iseabed, clamp to [minval, maxval] (interior cells); applies
 *           media_coeff_3d to the model, zeroes grad    (function.c:2790-2832) */
typedef struct fwi3d_compat fwi3d_compat;
int fwi3d_compat_create(fwi3d_model *m, int nsrc, int nrec, const int *rec_ijk, int start_f,
                        int upper_f, int it, const double *signal_x, const double *signal_true,
                        fwi3d_compat **out);
int fwi3d_compat_destroy(fwi3d_compat *c);
int fwi3d_compat_omegas(const fwi3d_compat *c, double *out);
int fwi3d_compat_reset(fwi3d_compat *c);
int fwi3d_compat_source(fwi3d_compat *c, int isrc, fwi3d_shot *s, const double *eobs, double *de,
                        double *err_sum);
int fwi3d_compat_green(fwi3d_compat *c, int irec, int idir, fwi3d_shot *g);
int fwi3d_compat_gamma(fwi3d_compat *c, double *grad);
int fwi3d_compat_phi(fwi3d_compat *c, int iseabed, double *scaler, double *sig2);
int fwi3d_compat_receivers(fwi3d_compat *c, fwi3d_shot *s, double *eff);
int fwi3d_compat_alpha(fwi3d_compat *c, int isrc, const double *eff2, double *k1, double *k2);
int fwi3d_compat_update(fwi3d_compat *c, const double *k1, const double *k2, double minval,
                        double maxval, int iseabed, double *alpha, double *sig);

/* ---- multi-shot serving (persistent multi-shot launches) ------------------ */

/* A batch advances all member shots (any models/grid sizes) with one launch per
 * half-step.  Shots join and leave between chunks; each keeps its own step. */
int fwi3d_batch_create(int max_shots, int chunk, void *stream, fwi3d_batch **out);
int fwi3d_batch_destroy(fwi3d_batch *b);
int fwi3d_batch_add(fwi3d_batch *b, fwi3d_shot *s);
int fwi3d_batch_remove(fwi3d_batch *b, fwi3d_shot *s);
/* advances every member by up to nsteps (finished shots idle) */
int fwi3d_batch_step(fwi3d_batch *b, int nsteps);
int fwi3d_batch_sync(fwi3d_batch *b);
int fwi3d_batch_size(const fwi3d_batch *b);

#ifdef __cplusplus
}
#endif
#endif
