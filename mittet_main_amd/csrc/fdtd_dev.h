// fdtd_dev.h -- device-side shot descriptor and kernel launchers shared by the
// HIP kernels (fdtd_kernels.hip) and the C-ABI host layer (fdtd_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace mt {

// Packed per-axis 1-D CPML profile rows on the device: prof[a][row * n_a + idx].
enum { PR_ESIG = 0, PR_MSIG, PR_BE, PR_BH, PR_CE, PR_CH, PR_KED, PR_KHD, PR_N };

// psi arrays per axis: slab arrays of [0,nc) U [n-nc,n) along the axis.
enum { PSI_EP = 0, PSI_EM, PSI_HP, PSI_HM, PSI_N };

constexpr int TX = 64;   // threads along x (one wavefront row, 512 B coalesced)
constexpr int TY = 4;    // rows per block
// E ring slots: the volume DFT reads each slot once per `ring` steps, so it
// costs 24 + 768/ring bytes per cell and step at 8 frequencies
constexpr int RING_MAX = 256;
constexpr int FLUSH_CELLS = 512;   // cells per DFT-flush block (256 lanes x 2)
constexpr int ZX = 32;   // z-marching tile: 32 x 8 columns, one thread per column
constexpr int ZY = 8;
// LDS z-marching kernels of the CPML-free interior: 64 x MT_LY tiles marching
// z-chunks of at most MT_LKB planes (shorter on small grids, so that a launch
// still fills the CUs).  At 288^3, 64 x 8 tiles (half the y-halo re-reads of
// 64 x 4) run the interior kernels 2-2.5 % faster than 64 x 4; 64 x 16 gains
// nothing; the chunk length moves the time there by about 1 %.
#ifndef MT_LKB
#define MT_LKB 16
#endif
#ifndef MT_LY
#define MT_LY 8
#endif
constexpr int LX = 64, LY = MT_LY, LKB = MT_LKB;

// Everything a kernel needs about one shot.  Lives in device memory; blocks
// find their shot through an exclusive prefix of per-shot block counts.
struct ShotDev {
    int ix, iy, iz, nc;
    int n;                    // cells
    int nbx, nby;             // xy tiles
    int it;                   // samples; steps 0 .. it-2 exist
    double dt, omega0, mu0, c1, c2;
    double dl[3];
    const double *prof[3];    // [PR_N][n_a]
    const double *eps;        // eps' = sigma/2/omega0 per cell
    const double *cbi;        // dt/eps' per cell (cb where the CPML sigma is zero)
    int kb;                   // z-chunk length of the z-marching kernels
    int ztx, zty, nkz;        // z-marching tiles along x, y and z-chunks
    int zb, zb_pad;           // z-marching blocks, padded to a multiple of 8
    int fb, fb_pad;           // per-cell fused blocks (one plane each), padded to 8
    int ftx, fty;             // per-cell fused tiles along x and y
    int pair;                 // 1: fused kernels update x-pairs of cells (even ix)
    const int4 *shell;        // tiles of the per-cell kernels when the LDS kernels cover the
                              // interior: {x0, y0, k, yend} (yend 0: a 32 x 8 tile; else a
                              // 16 x 16 tile with rows below yend), fb entries then k = -1
                              // padding; null: every 32 x 8 tile
    int lo[3], hi[3];         // LDS-kernel region [lo, hi) per axis (inside the CPML-free interior)
    int ltx, lty, lkz;        // LDS-kernel tiles along x, y and z-chunks
    int lkb;                  // LDS-kernel z-chunk length (planes)
    int lb, lb_pad;           // LDS-kernel blocks, padded to a multiple of 8
    double *e;                // E ring: slot q, comp d at e + q*sstride + d*n
    long long sstride;        // doubles between ring slots (3n + padding)
    long long astride;        // complex elements between accumulator rows (n + padding)
    int ring, cur;            // ring slots; slot that holds E at chunk start
    double *h;                // H [3][n]: updated in place, except by the fused E+H
                              // launches (see hsel)
    double *h2;               // second H buffer of shots that run the fused E+H interior
                              // kernel (null: none)
    int hpar;                 // 1: the current H lives in h2
    long long hstride;        // nonzero: H in ring slots like E (slot q at h + q*hstride;
                              // a gradient replay's stored states), see hsel
    double *psi[3];           // [PSI_N][slab[a]]
    long long pstride;        // nonzero: psi in ring slots like E (slot q at psi[a] + q*pstride;
                              // a gradient replay's stored states), read and written by the
                              // per-cell code only
    long long slab[3];
    int src_comp;
    int src_cell;             // -1: no source
    const double *amp;        // [it-1] source decrement per step
    int nrec;
    const int *rec_cell;      // [nrec]
    double *traces;           // [6][nrec][it]
    int nf;
    const double *w;          // complex [it][nf]
    double *acc;              // complex [3][nf] rows of astride
    double *acc_h;            // H volume spectra, same layout (null: off)
    double *hring;            // with acc_h: H after each step, ring slots like e
    int *step;                // absolute step of local index j = 0 of the current chunk
    int *flag;                // set to 1 once a gathered sample or the source cell's field
                              // turns NaN or infinite (divergence probe)
    int variant;              // 0: C v1.4 semantics; 1: Fortran ver0040 (fdtd_fortran.hip)
    int nsrc_h;               // variant 1: magnetic (Hz) source points; amp[q * (it - 1) + n]
    int src_h[4];             // their cells
    const double *mu;         // variant 1: per-cell permeability
};

// M_NT: once-touched field values are loaded and stored non-temporally
enum { M_UPD = 1, M_PSI = 2, M_SRC = 4, M_INPLACE = 8, M_NT = 16 };

// H before local step j of a chunk.  In a launch sequence that updates H in
// place (hf = 0) that is always the current buffer.  The fused E+H launches
// (hf = 1) read H from one buffer and write the other, because the halo of a
// fused tile reads old H that neighbouring tiles overwrite: step j then reads
// hsel(S, j, 1) and writes hsel(S, j + 1, 1).  Cells no step updates hold the
// same value in both buffers.  With hstride set, H before step j is ring slot
// (cur + j) % ring whatever hf is, and step j writes slot (cur + j + 1) % ring;
// only the per-cell and LDS kernel families read H through hsel(S, j, .) and
// write it through hsel(S, j + 1, .) (the others update in place).
__host__ __device__ inline double *hsel(const ShotDev &S, int j, int hf)
{
    if (S.hstride)
        return S.h + (size_t)((S.cur + j) % S.ring) * (size_t)S.hstride;
    return ((S.hpar + (hf ? j : 0)) & 1) ? S.h2 : S.h;
}

// Arguments of one adjoint step (fdtd_adjoint.hip), passed by value.
struct AdjArgs {
    int ix, iy, iz, nc, n;
    int ftx, fty, fb, fb_pad;        // 32 x 8 tiles per plane, as the fused forward kernels
    const int4 *shell;               // non-null: fb tiles {x0, y0, k, yend} of the cells
                                     // outside the LDS region (ShotDev::shell's format)
    // LDS region of the adjoint, [lo, hi) on each axis (empty: lo = hi = 0):
    // cells one cell clear of every CPML face, so that neither CPML
    // coefficients nor transposed psi stencils reach them
    int lo[3], hi[3];
    int ltx, lty, lkz, lb, lb_pad;   // 64 x LY tiles, z-chunks of lkb planes
    int lkb;
    int inner_hi[3];                 // cells in [nc, inner_hi[a]) on every axis see no CPML term
    int kappa_one;                   // ked = khd = dl everywhere
    double dt, mu0, c1, c2;
    double dl[3], rdl[3];
    const double *prof[3];
    const double *rp[3];             // per axis: 1/ked (row 0) and 1/khd (row 1)
    const double *eps;
    const double *cbi;               // dt/eps' per cell
    long long slab[3];
    double *lE, *lH, *lpsi[3];       // adjoint state (forward layout)
    double *t3, *tpsi[3];            // per-step weights for the transposed stencils
    const double *uE0, *uH0;         // forward state before the step
    const double *uE1, *upsi1[3];    // forward state after the step
    double *grad;                    // dJ/d eps' per cell (accumulated)
    double *grad_src;                // dJ/d sigma[src] from the source scaling
    int src_cell, src_comp;
    double amp_n, sig_src;
};

hipError_t launch_adjoint_step(const AdjArgs &A, const double *g, int nrec, const int *rec,
                               hipStream_t st);

// Kernel launchers (fdtd_kernels.hip).  `pre` is the exclusive prefix of
// per-shot block counts (device), `nblocks` its total.
hipError_t launch_efield(int mode, const ShotDev *d, int ns, const int *pre, int nblocks,
                         int j, int nover, int sign, hipStream_t st);
hipError_t launch_hfield(int mode, const ShotDev *d, int ns, const int *pre, int nblocks,
                         int j, int sign, hipStream_t st);
hipError_t launch_source(const ShotDev *d, int n, hipStream_t st);
hipError_t launch_efield_z(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           int sign, hipStream_t st);
hipError_t launch_hfield_z(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           int sign, hipStream_t st);
// rec_prev (j > 0): ns more blocks gather the receivers of local step j - 1.
// hf: part of a fused E+H step (H read from hsel(S, j, 1), written to
// hsel(S, j + 1, 1)); 0: H updated in place.  nt (also for the _l launches):
// the once-touched field values go non-temporal (M_NT).
hipError_t launch_efield_f(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           int sign, hipStream_t st, bool rec_prev = false, int hf = 0,
                           bool nt = false);
hipError_t launch_hfield_f(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           int sign, hipStream_t st, int hf = 0, bool nt = false);
hipError_t launch_efield_p(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           int sign, hipStream_t st);
hipError_t launch_hfield_p(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           int sign, hipStream_t st);
hipError_t launch_efield_l(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           int sign, hipStream_t st, bool nt = false);
hipError_t launch_hfield_l(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           int sign, hipStream_t st, bool nt = false);
// Fused E+H step over the LDS region (same blocks as launch_efield_l): after
// launch_efield_f(hf = 1), before launch_hfield_f(hf = 1); shots with h2 and a
// ring of more than one slot
hipError_t launch_ehfield_l(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                            hipStream_t st);
// Fortran-variant step (fdtd_fortran.hip): Hz source points before the E
// update, then per-cell E + CPML-E and H + CPML-H over 32 x 8 tiles
hipError_t launch_source_hz(const ShotDev *d, int ns, int j, hipStream_t st);
hipError_t launch_efield_v(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           hipStream_t st);
hipError_t launch_hfield_v(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           hipStream_t st);
hipError_t launch_receivers(const ShotDev *d, int ns, int j, int nover, int inplace,
                            hipStream_t st, int hf = 0);
// h: the H ring into acc_h (shots with H sums) instead of the E ring into acc
hipError_t launch_flush(const ShotDev *d, int ns, const int *pre_cells, int nblocks, int L,
                        hipStream_t st, bool h = false);
hipError_t launch_advance(const ShotDev *d, int ns, int L, hipStream_t st, int hf = 0);
// H after local step j's H update into its H ring slot (shots with hring)
hipError_t launch_hstore(const ShotDev *d, int ns, const int *pre_cells, int nblocks, int j,
                         hipStream_t st, int hf = 0);
hipError_t launch_receiver_dft(const double *traces, int ncomp_rows, int it, double dt,
                               double omega0, int nfreq, const double *omegas,
                               double *out, hipStream_t st);
hipError_t launch_bin_dft(const double *traces, int nrows, int it, double dt, double omega0,
                          int nbins, double *out, hipStream_t st);
hipError_t launch_fill(double *p, size_t count, double v, hipStream_t st);
hipError_t launch_copy(const double *src, double *dst, size_t count, hipStream_t st);
// one launch that gathers up to 5 segments into a contiguous buffer (save) or
// scatters the buffer back (restore): segment q is `count[q]` doubles at
// `live[q]`, stored from offset sum(count[< q])
struct StateSegs {
    double *live[5];
    size_t count[5];
    int nseg;
    unsigned first[6];   // first block of each segment (set by launch_state_copy)
};
hipError_t launch_state_copy(const StateSegs &segs, double *stored, bool save, hipStream_t st);
// plain 16-byte copy kernel of fwi3d_diag_copy (even count, 16-byte-aligned buffers)
hipError_t launch_diag_copy(const double *src, double *dst, size_t count, hipStream_t st);
hipError_t launch_eps(const double *sig, double *eps, double *cbi, int n, double omega0, double dt,
                      hipStream_t st);

// Complex arithmetic of the reference's C99 `double _Complex` code, usable on
// host and device with identical rounding (no contraction): products are
// (ac - bd) + (ad + bc)i; quotients use Smith's scaling, the algorithm of the
// compiler runtime's complex division for finite operands.
struct cplx {
    double re, im;
};

__host__ __device__ inline cplx cmul(cplx a, cplx b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

__host__ __device__ inline cplx cdiv(cplx a, cplx b)
{
    if (__builtin_fabs(b.re) < __builtin_fabs(b.im)) {
        const double r = b.re / b.im, den = b.re * r + b.im;
        return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
    }
    const double r = b.im / b.re, den = b.im * r + b.re;
    return {(a.im * r + a.re) / den, (a.im - a.re * r) / den};
}

// Reference-style FWI gradient chain (fdtd_compat.hip).  Spectra are complex
// [comp][bin][cell]; `acc` are a shot's running-DFT accumulators with row
// stride `astride` cells.
hipError_t launch_compat_eff(const double *acc, long long astride, int n, int nb,
                             const double *jw, const double *tw, double *eff, hipStream_t st);
hipError_t launch_compat_green(const double *acc, long long astride, int n, int nb,
                               const double *jw, int nsrc, const double *eff,
                               const double *dconj, double *G, double *sense, double *eback,
                               hipStream_t st);
hipError_t launch_compat_gamma(int ix, int iy, int iz, int nc, int nb, int nlow, int nsrc,
                               const double *eff, const double *eback, const double *sense,
                               double *grad, hipStream_t st);
hipError_t launch_compat_gather(const double *acc, long long astride, int nb, int nrec,
                                const int *rec, double *out, hipStream_t st);
hipError_t launch_phi_max(const double *sig, const double *grad, int n, int nxy, int iseabed,
                          double *blockmax, int nblocks, hipStream_t st);
hipError_t launch_phi_apply(const double *sig, const double *grad, double scaler, int n,
                            double *sig2, hipStream_t st);
hipError_t launch_update(double *sig, const double *grad, double step, double minval,
                         double maxval, int ix, int iy, int iz, int nc, int iseabed,
                         hipStream_t st);

}  // namespace mt
