// fdtd_fortran.hip -- the Fortran lineage's step (ver0040), variant 1 of the
// library: magnetic Hz source points applied before the E update
// (ver0040/cpu_waveform.f90:6-58, cpu_main.f90:104-148), E and H updates over
// the Fortran loop ranges (cpu_eh_field_cpml42.f90:116-252), CPML over its
// shells with every other axis in full range (cpu_cpml4.f90:271-474, including
// the c2 in place of c1 of the psi_Hyx stencil, :400, 413), and per-cell
// permeability in the H coefficients (cpu_eh_field_cpml42.f90:51-53).
//
// Every thread owns one cell of a 32 x 8 tile of one plane and applies, in the
// Fortran order, the field update (where the loop range contains the cell) and
// then the x, y and z CPML corrections (where the shell contains it).  The
// per-cell order of operations is the Fortran one, so the result is that of
// the operator-by-operator loops bit for bit.  Cells of a Fortran shell range
// outside the stored slabs ([0, nc) and [n - nc, n)) have zero profiles there
// (their psi stays 0 and adds 0) and are skipped.

#include "fdtd_dev.h"

#define GLOBAL __attribute__((address_space(1)))

namespace mt {

namespace {

__device__ __forceinline__ const GLOBAL double *cg(const double *p) { return (const GLOBAL double *)p; }
__device__ __forceinline__ GLOBAL double *mg(double *p) { return (GLOBAL double *)p; }

__device__ __forceinline__ int shot_of(const int *__restrict__ pre, int ns, int b)
{
    int s = 0;
    while (s + 1 < ns && b >= pre[s + 1])
        ++s;
    return s;
}

__device__ __forceinline__ int slab_at(int r, int n, int nc) { return r < nc ? r : r - (n - 2 * nc); }

// x / c of a complex x by a real c, as the reference's Fortran evaluates it
// compiled with amdflang: c becomes the complex (c, 0) and each part of the
// quotient is (x c) / (c c) (the textbook complex quotient with zero
// imaginary divisor), which differs from x / c in the last place.  Applies to
// every stencil division, psi update and the source scaling; the real and
// imaginary runs of a complex field take it part by part.
__device__ __forceinline__ double cdivr(double x, double c) { return (x * c) / (c * c); }

// c1 F[c] - c1 F[c-s] + c2 F[c+s] - c2 F[c-2s]
__device__ __forceinline__ double dm(const GLOBAL double *F, size_t c, size_t s, double c1, double c2)
{
    return c1 * F[c] - c1 * F[c - s] + c2 * F[c + s] - c2 * F[c - 2 * s];
}

// c1 F[c+s] - c1 F[c] + c2 F[c+2s] - c2 F[c-s]
__device__ __forceinline__ double dp(const GLOBAL double *F, size_t c, size_t s, double c1, double c2)
{
    return c1 * F[c + s] - c1 * F[c] + c2 * F[c + 2 * s] - c2 * F[c - s];
}

__device__ __forceinline__ bool vcell(const ShotDev &S, int b, int &i, int &jj, int &k)
{
    const int q = (b & 7) * (S.fb_pad >> 3) + (b >> 3);
    if (q >= S.fb)
        return false;
    const int tx = q % S.ftx, r = q / S.ftx;
    const int ty = r % S.fty;
    k = r / S.fty;
    i = tx * ZX + (threadIdx.x & (ZX - 1));
    jj = ty * ZY + threadIdx.x / ZX;
    return i < S.ix && jj < S.iy;
}

// E shells: [2, nc + 1) and [n - nc, n - 1), stored part [2, nc) and [n - nc, n - 1)
__device__ __forceinline__ bool e_shell(int r, int n, int nc)
{
    return (r >= 2 && r < nc) || (r >= n - nc && r < n - 1);
}

// H shells: [1, nc) and [n - nc - 1, n - 2), stored part [1, nc) and [n - nc, n - 2)
__device__ __forceinline__ bool h_shell(int r, int n, int nc)
{
    return (r >= 1 && r < nc) || (r >= n - nc && r < n - 2);
}

}  // namespace

// Hz(src) -= s[n] dt / mu(src) / dx / dy / dz for each point, in order
__global__ void k_source_hz(const ShotDev *__restrict__ D, int ns, int j)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns)
        return;
    const ShotDev &S = D[s];
    const int n = *S.step + j;
    if (n >= S.it - 1 || S.nsrc_h <= 0)
        return;
    GLOBAL double *Hz = mg(S.h + 2 * (size_t)S.n);
    for (int q = 0; q < S.nsrc_h && q < 4; ++q)
        Hz[S.src_h[q]] = Hz[S.src_h[q]] - cg(S.amp)[(size_t)q * (S.it - 1) + n];
}

__global__ __launch_bounds__(ZX *ZY) void k_efield_v(const ShotDev *__restrict__ D, int ns,
                                                     const int *__restrict__ pre, int j)
{
    const int s = shot_of(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    const int n = *S.step + j;
    if (n >= S.it - 1)
        return;
    int i, jj, k;
    if (!vcell(S, blockIdx.x - pre[s], i, jj, k))
        return;
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc;
    const size_t N = (size_t)S.n, sy = ix, sz = (size_t)ix * iy;
    const size_t c = ((size_t)k * iy + jj) * ix + i;
    const int si = (S.cur + j) % S.ring, so = (S.cur + j + 1) % S.ring;
    const GLOBAL double *Ein = cg(S.e + (size_t)si * S.sstride);
    GLOBAL double *Eout = mg(S.e + (size_t)so * S.sstride);
    const GLOBAL double *Hx = cg(S.h), *Hy = cg(S.h + N), *Hz = cg(S.h + 2 * N);
    const GLOBAL double *px = cg(S.prof[0]), *py = cg(S.prof[1]), *pz = cg(S.prof[2]);
    const double c1 = S.c1, c2 = S.c2, dt = S.dt;
    const double eps2 = cg(S.eps)[c];
    const double esx = px[PR_ESIG * ix + i], esy = py[PR_ESIG * iy + jj], esz = pz[PR_ESIG * iz + k];
    // media_coeff (cpu_eh_field_cpml42.f90:26-38)
    const double cbx = dt / eps2 / (1.0 + (esx * dt) / (2.0 * eps2));
    const double cby = dt / eps2 / (1.0 + (esy * dt) / (2.0 * eps2));
    const double cbz = dt / eps2 / (1.0 + (esz * dt) / (2.0 * eps2));
    double ex = Ein[c], ey = Ein[N + c], ez = Ein[2 * N + c];
    // e_field_cpml4 (:125-155)
    if (k >= 2 && k < iz - 1 && jj >= 2 && jj < iy - 1) {
        const double ca = (1.0 - ((esx * dt) / (2.0 * eps2))) / (1.0 + ((esx * dt) / (2.0 * eps2)));
        ex = ca * ex + cbx * (cdivr(dm(Hz, c, sy, c1, c2), py[PR_KED * iy + jj]) -
                              cdivr(dm(Hy, c, sz, c1, c2), pz[PR_KED * iz + k]));
    }
    if (k >= 2 && k < iz - 1 && i >= 2 && i < ix - 1) {
        const double ca = (1.0 - ((esy * dt) / (2.0 * eps2))) / (1.0 + ((esy * dt) / (2.0 * eps2)));
        ey = ca * ey + cby * (cdivr(dm(Hx, c, sz, c1, c2), pz[PR_KED * iz + k]) -
                              cdivr(dm(Hz, c, 1, c1, c2), px[PR_KED * ix + i]));
    }
    if (jj >= 2 && jj < iy - 1 && i >= 2 && i < ix - 1) {
        const double ca = (1.0 - ((esz * dt) / (2.0 * eps2))) / (1.0 + ((esz * dt) / (2.0 * eps2)));
        ez = ca * ez + cbz * (cdivr(dm(Hy, c, 1, c1, c2), px[PR_KED * ix + i]) -
                              cdivr(dm(Hx, c, sy, c1, c2), py[PR_KED * iy + jj]));
    }
    // CPML_E4 (cpu_cpml4.f90:281-358): x, then y, then z
    if (e_shell(i, ix, nc)) {
        GLOBAL double *P = mg(S.psi[0]);
        const size_t L = (size_t)S.slab[0];
        const size_t q = ((size_t)k * iy + jj) * (2 * nc) + slab_at(i, ix, nc);
        const double be = px[PR_BE * ix + i], ce = px[PR_CE * ix + i], h = S.dl[0];
        const double p0 = be * P[q] + cdivr(ce * dm(Hy, c, 1, c1, c2), h);      // psi_Ezx
        const double p1 = be * P[L + q] + cdivr(ce * dm(Hz, c, 1, c1, c2), h);  // psi_Eyx
        P[q] = p0;
        P[L + q] = p1;
        ez = ez + cbz * p0;
        ey = ey - cby * p1;
    }
    if (e_shell(jj, iy, nc)) {
        GLOBAL double *P = mg(S.psi[1]);
        const size_t L = (size_t)S.slab[1];
        const size_t q = ((size_t)k * (2 * nc) + slab_at(jj, iy, nc)) * ix + i;
        const double be = py[PR_BE * iy + jj], ce = py[PR_CE * iy + jj], h = S.dl[1];
        const double p0 = be * P[q] + cdivr(ce * dm(Hz, c, sy, c1, c2), h);      // psi_Exy
        const double p1 = be * P[L + q] + cdivr(ce * dm(Hx, c, sy, c1, c2), h);  // psi_Ezy
        P[q] = p0;
        P[L + q] = p1;
        ex = ex + cbx * p0;
        ez = ez - cbz * p1;
    }
    if (e_shell(k, iz, nc)) {
        GLOBAL double *P = mg(S.psi[2]);
        const size_t L = (size_t)S.slab[2];
        const size_t q = ((size_t)slab_at(k, iz, nc) * iy + jj) * ix + i;
        const double be = pz[PR_BE * iz + k], ce = pz[PR_CE * iz + k], h = S.dl[2];
        const double p0 = be * P[q] + cdivr(ce * dm(Hx, c, sz, c1, c2), h);      // psi_Eyz
        const double p1 = be * P[L + q] + cdivr(ce * dm(Hy, c, sz, c1, c2), h);  // psi_Exz
        P[q] = p0;
        P[L + q] = p1;
        ey = ey + cby * p0;
        ex = ex - cbx * p1;
    }
    Eout[c] = ex;
    Eout[N + c] = ey;
    Eout[2 * N + c] = ez;
}

__global__ __launch_bounds__(ZX *ZY) void k_hfield_v(const ShotDev *__restrict__ D, int ns,
                                                     const int *__restrict__ pre, int j)
{
    const int s = shot_of(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    if (*S.step + j >= S.it - 1)
        return;
    int i, jj, k;
    if (!vcell(S, blockIdx.x - pre[s], i, jj, k))
        return;
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc;
    const size_t N = (size_t)S.n, sy = ix, sz = (size_t)ix * iy;
    const size_t c = ((size_t)k * iy + jj) * ix + i;
    const int so = (S.cur + j + 1) % S.ring;
    const GLOBAL double *Ex = cg(S.e + (size_t)so * S.sstride), *Ey = Ex + N, *Ez = Ex + 2 * N;
    GLOBAL double *Hx = mg(S.h), *Hy = mg(S.h + N), *Hz = mg(S.h + 2 * N);
    const GLOBAL double *px = cg(S.prof[0]), *py = cg(S.prof[1]), *pz = cg(S.prof[2]);
    const double c1 = S.c1, c2 = S.c2, dt = S.dt;
    const double eps2 = cg(S.eps)[c], mu = cg(S.mu)[c];
    const double msx = px[PR_MSIG * ix + i], msy = py[PR_MSIG * iy + jj], msz = pz[PR_MSIG * iz + k];
    // media_coeff (cpu_eh_field_cpml42.f90:40-53)
    const double dbx = dt / mu / (1.0 + (msx * dt) / (2.0 * eps2));
    const double dby = dt / mu / (1.0 + (msy * dt) / (2.0 * eps2));
    const double dbz = dt / mu / (1.0 + (msz * dt) / (2.0 * eps2));
    double hx = Hx[c], hy = Hy[c], hz = Hz[c];
    // h_field_cpml4 (:221-251)
    if (k >= 1 && k < iz - 2 && jj >= 1 && jj < iy - 2) {
        const double da = (1.0 - ((msx * dt) / (2.0 * eps2))) / (1.0 + ((msx * dt) / (2.0 * eps2)));
        hx = da * hx - dbx * (cdivr(dp(Ez, c, sy, c1, c2), py[PR_KHD * iy + jj]) -
                              cdivr(dp(Ey, c, sz, c1, c2), pz[PR_KHD * iz + k]));
    }
    if (k >= 1 && k < iz - 2 && i >= 1 && i < ix - 2) {
        const double da = (1.0 - ((msy * dt) / (2.0 * eps2))) / (1.0 + ((msy * dt) / (2.0 * eps2)));
        hy = da * hy - dby * (cdivr(dp(Ex, c, sz, c1, c2), pz[PR_KHD * iz + k]) -
                              cdivr(dp(Ez, c, 1, c1, c2), px[PR_KHD * ix + i]));
    }
    if (jj >= 1 && jj < iy - 2 && i >= 1 && i < ix - 2) {
        const double da = (1.0 - ((msz * dt) / (2.0 * eps2))) / (1.0 + ((msz * dt) / (2.0 * eps2)));
        hz = da * hz - dbz * (cdivr(dp(Ey, c, 1, c1, c2), px[PR_KHD * ix + i]) -
                              cdivr(dp(Ex, c, sy, c1, c2), py[PR_KHD * iy + jj]));
    }
    // CPML_H4 (cpu_cpml4.f90:394-471): x, then y, then z
    if (h_shell(i, ix, nc)) {
        GLOBAL double *P = mg(S.psi[0]) + PSI_HP * (size_t)S.slab[0];
        const size_t L = (size_t)S.slab[0];
        const size_t q = ((size_t)k * iy + jj) * (2 * nc) + slab_at(i, ix, nc);
        const double bh = px[PR_BH * ix + i], ch = px[PR_CH * ix + i], h = S.dl[0];
        // psi_Hyx takes c2 where the stencil has c1 (cpu_cpml4.f90:400, 413)
        const double hyx = c1 * Ez[c + 1] - c2 * Ez[c] + c2 * Ez[c + 2] - c2 * Ez[c - 1];
        const double p0 = bh * P[q] + cdivr(ch * dp(Ey, c, 1, c1, c2), h);  // psi_Hzx
        const double p1 = bh * P[L + q] + cdivr(ch * hyx, h);                // psi_Hyx
        P[q] = p0;
        P[L + q] = p1;
        hz = hz - dbz * p0;
        hy = hy + dby * p1;
    }
    if (h_shell(jj, iy, nc)) {
        GLOBAL double *P = mg(S.psi[1]) + PSI_HP * (size_t)S.slab[1];
        const size_t L = (size_t)S.slab[1];
        const size_t q = ((size_t)k * (2 * nc) + slab_at(jj, iy, nc)) * ix + i;
        const double bh = py[PR_BH * iy + jj], ch = py[PR_CH * iy + jj], h = S.dl[1];
        const double p0 = bh * P[q] + cdivr(ch * dp(Ez, c, sy, c1, c2), h);      // psi_Hxy
        const double p1 = bh * P[L + q] + cdivr(ch * dp(Ex, c, sy, c1, c2), h);  // psi_Hzy
        P[q] = p0;
        P[L + q] = p1;
        hx = hx - dbx * p0;
        hz = hz + dbz * p1;
    }
    if (h_shell(k, iz, nc)) {
        GLOBAL double *P = mg(S.psi[2]) + PSI_HP * (size_t)S.slab[2];
        const size_t L = (size_t)S.slab[2];
        const size_t q = ((size_t)slab_at(k, iz, nc) * iy + jj) * ix + i;
        const double bh = pz[PR_BH * iz + k], ch = pz[PR_CH * iz + k], h = S.dl[2];
        const double p0 = bh * P[q] + cdivr(ch * dp(Ex, c, sz, c1, c2), h);      // psi_Hyz
        const double p1 = bh * P[L + q] + cdivr(ch * dp(Ey, c, sz, c1, c2), h);  // psi_Hxz
        P[q] = p0;
        P[L + q] = p1;
        hy = hy - dby * p0;
        hx = hx + dbx * p1;
    }
    Hx[c] = hx;
    Hy[c] = hy;
    Hz[c] = hz;
}

hipError_t launch_source_hz(const ShotDev *d, int ns, int j, hipStream_t st)
{
    hipLaunchKernelGGL(k_source_hz, dim3((ns + 63) / 64), dim3(64), 0, st, d, ns, j);
    return hipGetLastError();
}

hipError_t launch_efield_v(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           hipStream_t st)
{
    if (nblocks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_efield_v, dim3(nblocks), dim3(ZX * ZY), 0, st, d, ns, pre, j);
    return hipGetLastError();
}

hipError_t launch_hfield_v(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                           hipStream_t st)
{
    if (nblocks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_hfield_v, dim3(nblocks), dim3(ZX * ZY), 0, st, d, ns, pre, j);
    return hipGetLastError();
}

}  // namespace mt
