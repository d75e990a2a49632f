// fdtd_kernels.hip -- gfx950 kernels of the fictitious-wave-domain FDTD step.
//
// One descriptor-driven kernel family serves a single shot and multi-shot
// batches alike: every block looks its shot up in an exclusive prefix of
// per-shot block counts, so shots of different grid sizes share one launch.
//
// E kernel (per cell, x fastest, 64x4 threads per block):
//   E-update of satan_i/fwi3d_cpu_function_v1.4.c:1184-1236, then the CPML-E
//   corrections of satan_i/fwi3d_cpu_cpmlabc_v1.4.c:402-543 in the order
//   x, y, z, then the source (function.c:514-529).  Coefficients
//   (function.c:403-423) are recomputed from eps' and the 1-D profiles, psi is
//   kept only in the boundary slabs.  E is written to the next slot of a ring
//   so the volume DFT can be applied once per chunk (flush kernel).
// H kernel: function.c:1238-1290 then cpmlabc.c:545-679, H in place.
//
// All arithmetic keeps the reference's evaluation order; the library is built
// with -ffp-contract=off so results equal the reference's bit for bit.
#include "fdtd_dev.h"

namespace mt {

__device__ __forceinline__ double dminus(const double *__restrict__ F, int c, int s, double c1,
                                         double c2)
{
    return c1 * F[c] - c1 * F[c - s] + c2 * F[c + s] - c2 * F[c - 2 * s];
}

__device__ __forceinline__ double dplus(const double *__restrict__ F, int c, int s, double c1,
                                        double c2)
{
    return c1 * F[c + s] - c1 * F[c] + c2 * F[c + 2 * s] - c2 * F[c - s];
}

__device__ __forceinline__ int find_shot(const int *__restrict__ pre, int ns, int b)
{
    int s = 0;
    while (s + 1 < ns && b >= pre[s + 1])
        ++s;
    return s;
}

// slab position of index r on an axis of length n with nc-cell shells
__device__ __forceinline__ int slab_pos(int r, int n, int nc)
{
    return r < nc ? r : r - (n - 2 * nc);
}

template <int MODE>
__global__ __launch_bounds__(TX *TY) void k_efield(const ShotDev *__restrict__ D, int ns,
                                                   const int *__restrict__ pre, int j, int nover,
                                                   int sign)
{
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    int n = nover;
    if (n < 0) {
        n = *S.step + j;
        if (n >= S.it - 1)
            return;  // this shot has finished its time loop
    }
    const int b = blockIdx.x - pre[s];
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc, N = S.n;
    const int bx = b % S.nbx, t = b / S.nbx;
    const int by = t % S.nby, k = t / S.nby;
    const int i = bx * TX + threadIdx.x, jj = by * TY + threadIdx.y;
    if (i >= ix || jj >= iy)
        return;
    const int sy = ix, sz = ix * iy;
    const int c = (k * iy + jj) * ix + i;

    int si, so;
    if (MODE & M_INPLACE) {
        si = so = S.cur;
    } else {
        si = (S.cur + j) % S.ring;
        so = (S.cur + j + 1) % S.ring;
    }
    const double *__restrict__ Ein = S.e + (size_t)si * 3 * N;
    double *__restrict__ Eout = S.e + (size_t)so * 3 * N;
    const double *__restrict__ Hx = S.h;
    const double *__restrict__ Hy = S.h + N;
    const double *__restrict__ Hz = S.h + 2 * (size_t)N;
    double ex = Ein[c], ey = Ein[N + c], ez = Ein[2 * (size_t)N + c];

    const double c1 = S.c1, c2 = S.c2, dt = S.dt;
    const double *__restrict__ px = S.prof[0];
    const double *__restrict__ py = S.prof[1];
    const double *__restrict__ pz = S.prof[2];

    const bool rx1 = i >= 1 && i < ix - 1, rx2 = i >= 2 && i < ix - 1;
    const bool ry1 = jj >= 1 && jj < iy - 1, ry2 = jj >= 2 && jj < iy - 1;
    const bool rz1 = k >= 1 && k < iz - 1, rz2 = k >= 2 && k < iz - 1;
    const bool upx = (MODE & M_UPD) && rx1 && ry2 && rz2;
    const bool upy = (MODE & M_UPD) && rx2 && ry1 && rz2;
    const bool upz = (MODE & M_UPD) && rx2 && ry2 && rz1;
    // CPML-E faces: x [2,nc-1)U[ix-nc+1,ix-2); y,z [2,nc)U[n-nc+1,n-2)
    const bool fx = (MODE & M_PSI) && ((i >= 2 && i < nc - 1) || (i >= ix - nc + 1 && i < ix - 2)) &&
                    jj < iy - 1 && k < iz - 1;
    const bool fy = (MODE & M_PSI) && ((jj >= 2 && jj < nc) || (jj >= iy - nc + 1 && jj < iy - 2)) &&
                    i < ix - 1 && k < iz - 1;
    const bool fz = (MODE & M_PSI) && ((k >= 2 && k < nc) || (k >= iz - nc + 1 && k < iz - 2)) &&
                    i < ix - 1 && jj < iy - 1;

    double hzy = 0.0, hyz = 0.0, hxz = 0.0, hzx = 0.0, hyx = 0.0, hxy = 0.0;
    if (upx || fy) hzy = dminus(Hz, c, sy, c1, c2);
    if (upx || fz) hyz = dminus(Hy, c, sz, c1, c2);
    if (upy || fz) hxz = dminus(Hx, c, sz, c1, c2);
    if (upy || fx) hzx = dminus(Hz, c, 1, c1, c2);
    if (upz || fx) hyx = dminus(Hy, c, 1, c1, c2);
    if (upz || fy) hxy = dminus(Hx, c, sy, c1, c2);

    const double eps2 = S.eps[c];
    const double sex = px[PR_ESIG * ix + i] * dt / (2.0 * eps2);
    const double sey = py[PR_ESIG * iy + jj] * dt / (2.0 * eps2);
    const double sez = pz[PR_ESIG * iz + k] * dt / (2.0 * eps2);
    const double cbx = dt / eps2 / (1.0 + sex);
    const double cby = dt / eps2 / (1.0 + sey);
    const double cbz = dt / eps2 / (1.0 + sez);

    if (upx) {
        const double ca = (1.0 - sex) / (1.0 + sex);
        const double cu = hzy / py[PR_KED * iy + jj] - hyz / pz[PR_KED * iz + k];
        ex = sign > 0 ? ca * ex + cbx * cu : ca * ex - cbx * cu;
    }
    if (upy) {
        const double ca = (1.0 - sey) / (1.0 + sey);
        const double cu = hxz / pz[PR_KED * iz + k] - hzx / px[PR_KED * ix + i];
        ey = sign > 0 ? ca * ey + cby * cu : ca * ey - cby * cu;
    }
    if (upz) {
        const double ca = (1.0 - sez) / (1.0 + sez);
        const double cu = hyx / px[PR_KED * ix + i] - hxy / py[PR_KED * iy + jj];
        ez = sign > 0 ? ca * ez + cbz * cu : ca * ez - cbz * cu;
    }
    if (fx) {
        double *__restrict__ P = S.psi[0];
        const size_t L = (size_t)S.slab[0];
        const size_t q = ((size_t)k * iy + jj) * (2 * nc) + slab_pos(i, ix, nc);
        const double be = px[PR_BE * ix + i], ce = px[PR_CE * ix + i], h = S.dl[0];
        const double p0 = be * P[q] + ce * hyx / h;
        const double p1 = be * P[L + q] + ce * hzx / h;
        P[q] = p0;
        P[L + q] = p1;
        ez += cbz * p0;
        ey -= cby * p1;
    }
    if (fy) {
        double *__restrict__ P = S.psi[1];
        const size_t L = (size_t)S.slab[1];
        const size_t q = ((size_t)k * (2 * nc) + slab_pos(jj, iy, nc)) * ix + i;
        const double be = py[PR_BE * iy + jj], ce = py[PR_CE * iy + jj], h = S.dl[1];
        const double p0 = be * P[q] + ce * hzy / h;
        const double p1 = be * P[L + q] + ce * hxy / h;
        P[q] = p0;
        P[L + q] = p1;
        ex += cbx * p0;
        ez -= cbz * p1;
    }
    if (fz) {
        double *__restrict__ P = S.psi[2];
        const size_t L = (size_t)S.slab[2];
        const size_t q = ((size_t)slab_pos(k, iz, nc) * iy + jj) * ix + i;
        const double be = pz[PR_BE * iz + k], ce = pz[PR_CE * iz + k], h = S.dl[2];
        const double p0 = be * P[q] + ce * hxz / h;
        const double p1 = be * P[L + q] + ce * hyz / h;
        P[q] = p0;
        P[L + q] = p1;
        ey += cby * p0;
        ex -= cbx * p1;
    }
    if ((MODE & M_SRC) && c == S.src_cell) {
        const double a = S.amp[n];
        if (S.src_comp == 0)
            ex -= a;
        else if (S.src_comp == 1)
            ey -= a;
        else
            ez -= a;
    }
    Eout[c] = ex;
    Eout[N + c] = ey;
    Eout[2 * (size_t)N + c] = ez;
}

template <int MODE>
__global__ __launch_bounds__(TX *TY) void k_hfield(const ShotDev *__restrict__ D, int ns,
                                                   const int *__restrict__ pre, int j, int sign)
{
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    int so;
    if (MODE & M_INPLACE) {
        so = S.cur;
    } else {
        if (*S.step + j >= S.it - 1)
            return;
        so = (S.cur + j + 1) % S.ring;
    }
    const int b = blockIdx.x - pre[s];
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc, N = S.n;
    const int bx = b % S.nbx, t = b / S.nbx;
    const int by = t % S.nby, k = t / S.nby;
    const int i = bx * TX + threadIdx.x, jj = by * TY + threadIdx.y;
    if (i >= ix || jj >= iy)
        return;
    const int sy = ix, sz = ix * iy;
    const int c = (k * iy + jj) * ix + i;

    const double *__restrict__ Ex = S.e + (size_t)so * 3 * N;
    const double *__restrict__ Ey = Ex + N;
    const double *__restrict__ Ez = Ex + 2 * (size_t)N;
    double *__restrict__ Hx = S.h;
    double *__restrict__ Hy = S.h + N;
    double *__restrict__ Hz = S.h + 2 * (size_t)N;

    const double c1 = S.c1, c2 = S.c2, dt = S.dt;
    const double *__restrict__ px = S.prof[0];
    const double *__restrict__ py = S.prof[1];
    const double *__restrict__ pz = S.prof[2];

    const bool ax1 = i >= 1 && i < ix - 1, ax2 = i >= 1 && i < ix - 2;
    const bool ay1 = jj >= 1 && jj < iy - 1, ay2 = jj >= 1 && jj < iy - 2;
    const bool az1 = k >= 1 && k < iz - 1, az2 = k >= 1 && k < iz - 2;
    const bool upx = (MODE & M_UPD) && ax1 && ay2 && az2;
    const bool upy = (MODE & M_UPD) && ax2 && ay1 && az2;
    const bool upz = (MODE & M_UPD) && ax2 && ay2 && az1;
    // CPML-H faces: [1,nc-1)U[n-nc,n-2) on every axis
    const bool gx = (MODE & M_PSI) && ((i >= 1 && i < nc - 1) || (i >= ix - nc && i < ix - 2)) &&
                    jj < iy - 1 && k < iz - 1;
    const bool gy = (MODE & M_PSI) && ((jj >= 1 && jj < nc - 1) || (jj >= iy - nc && jj < iy - 2)) &&
                    i < ix - 1 && k < iz - 1;
    const bool gz = (MODE & M_PSI) && ((k >= 1 && k < nc - 1) || (k >= iz - nc && k < iz - 2)) &&
                    i < ix - 1 && jj < iy - 1;
    if (!(upx || upy || upz || gx || gy || gz))
        return;

    double ezy = 0.0, eyz = 0.0, exz = 0.0, ezx = 0.0, eyx = 0.0, exy = 0.0;
    if (upx || gy) ezy = dplus(Ez, c, sy, c1, c2);
    if (upx || gz) eyz = dplus(Ey, c, sz, c1, c2);
    if (upy || gz) exz = dplus(Ex, c, sz, c1, c2);
    if (upy || gx) ezx = dplus(Ez, c, 1, c1, c2);
    if (upz || gx) eyx = dplus(Ey, c, 1, c1, c2);
    if (upz || gy) exy = dplus(Ex, c, sy, c1, c2);

    // da = 1 and db = dt/mu0 wherever the magnetic CPML conductivity is zero;
    // eps' is only needed inside the shells.
    const double msx = px[PR_MSIG * ix + i], msy = py[PR_MSIG * iy + jj], msz = pz[PR_MSIG * iz + k];
    double smx = 0.0, smy = 0.0, smz = 0.0;
    if (msx != 0.0 || msy != 0.0 || msz != 0.0) {
        const double eps2 = S.eps[c];
        smx = msx * dt / (2.0 * eps2);
        smy = msy * dt / (2.0 * eps2);
        smz = msz * dt / (2.0 * eps2);
    }
    const double dbx = dt / S.mu0 / (1.0 + smx);
    const double dby = dt / S.mu0 / (1.0 + smy);
    const double dbz = dt / S.mu0 / (1.0 + smz);
    double hx = Hx[c], hy = Hy[c], hz = Hz[c];

    if (upx) {
        const double da = (1.0 - smx) / (1.0 + smx);
        const double cu = ezy / py[PR_KHD * iy + jj] - eyz / pz[PR_KHD * iz + k];
        hx = sign > 0 ? da * hx - dbx * cu : da * hx + dbx * cu;
    }
    if (upy) {
        const double da = (1.0 - smy) / (1.0 + smy);
        const double cu = exz / pz[PR_KHD * iz + k] - ezx / px[PR_KHD * ix + i];
        hy = sign > 0 ? da * hy - dby * cu : da * hy + dby * cu;
    }
    if (upz) {
        const double da = (1.0 - smz) / (1.0 + smz);
        const double cu = eyx / px[PR_KHD * ix + i] - exy / py[PR_KHD * iy + jj];
        hz = sign > 0 ? da * hz - dbz * cu : da * hz + dbz * cu;
    }
    if (gx) {
        double *__restrict__ P = S.psi[0];
        const size_t L = (size_t)S.slab[0];
        const size_t q = ((size_t)k * iy + jj) * (2 * nc) + slab_pos(i, ix, nc);
        const double bh = px[PR_BH * ix + i], ch = px[PR_CH * ix + i], h = S.dl[0];
        const double p0 = bh * P[2 * L + q] + ch * ezx / h;
        const double p1 = bh * P[3 * L + q] + ch * eyx / h;
        P[2 * L + q] = p0;
        P[3 * L + q] = p1;
        hy += dby * p0;
        hz -= dbz * p1;
    }
    if (gy) {
        double *__restrict__ P = S.psi[1];
        const size_t L = (size_t)S.slab[1];
        const size_t q = ((size_t)k * (2 * nc) + slab_pos(jj, iy, nc)) * ix + i;
        const double bh = py[PR_BH * iy + jj], ch = py[PR_CH * iy + jj], h = S.dl[1];
        const double p0 = bh * P[2 * L + q] + ch * exy / h;
        const double p1 = bh * P[3 * L + q] + ch * ezy / h;
        P[2 * L + q] = p0;
        P[3 * L + q] = p1;
        hz += dbz * p0;
        hx -= dbx * p1;
    }
    if (gz) {
        double *__restrict__ P = S.psi[2];
        const size_t L = (size_t)S.slab[2];
        const size_t q = ((size_t)slab_pos(k, iz, nc) * iy + jj) * ix + i;
        const double bh = pz[PR_BH * iz + k], ch = pz[PR_CH * iz + k], h = S.dl[2];
        const double p0 = bh * P[2 * L + q] + ch * eyz / h;
        const double p1 = bh * P[3 * L + q] + ch * exz / h;
        P[2 * L + q] = p0;
        P[3 * L + q] = p1;
        hx += dbx * p0;
        hy -= dby * p1;
    }
    Hx[c] = hx;
    Hy[c] = hy;
    Hz[c] = hz;
}

// read_source_3d as a stand-alone operator (in place on the current slot)
__global__ void k_source(const ShotDev *__restrict__ D, int n)
{
    const ShotDev &S = D[0];
    if (S.src_cell < 0 || n < 0 || n >= S.it - 1)
        return;
    double *E = S.e + ((size_t)S.cur * 3 + S.src_comp) * S.n;
    E[S.src_cell] -= S.amp[n];
}

// Receiver gather: E (from the slot written this step) and H after the step.
__global__ void k_receivers(const ShotDev *__restrict__ D, int j, int nover, int inplace)
{
    const ShotDev &S = D[blockIdx.x];
    int n = nover, so = S.cur;
    if (n < 0) {
        n = *S.step + j;
        if (n >= S.it - 1)
            return;
    }
    if (!inplace)
        so = (S.cur + j + 1) % S.ring;
    const int tot = 6 * S.nrec;
    for (int q = threadIdx.x; q < tot; q += blockDim.x) {
        const int comp = q / S.nrec, r = q - comp * S.nrec;
        const int cell = S.rec_cell[r];
        const double v = comp < 3 ? S.e[((size_t)so * 3 + comp) * S.n + cell]
                                  : S.h[(size_t)(comp - 3) * S.n + cell];
        S.traces[((size_t)comp * S.nrec + r) * S.it + n] = v;
    }
}

// Volume running DFT over the last L ring slots: acc += E[n] * w[n] in step order.
__global__ __launch_bounds__(256) void k_flush(const ShotDev *__restrict__ D, int ns,
                                               const int *__restrict__ pre, int L)
{
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    const int c = (blockIdx.x - pre[s]) * blockDim.x + threadIdx.x;
    if (c >= S.n || S.nf == 0)
        return;
    const int n0 = *S.step;
    int Lv = L;
    if (n0 + Lv > S.it - 1)
        Lv = S.it - 1 - n0;
    if (Lv <= 0)
        return;
    const int N = S.n, nf = S.nf, ring = S.ring, cur = S.cur;
    const double2 *__restrict__ W = reinterpret_cast<const double2 *>(S.w) + (size_t)n0 * nf;
    double2 *__restrict__ A = reinterpret_cast<double2 *>(S.acc);
    for (int d = 0; d < 3; ++d) {
        double ev[RING_MAX];
#pragma unroll
        for (int t = 0; t < RING_MAX; ++t)
            ev[t] = t < Lv ? S.e[((size_t)((cur + 1 + t) % ring) * 3 + d) * N + c] : 0.0;
        for (int f = 0; f < nf; ++f) {
            double2 a = A[((size_t)d * nf + f) * N + c];
#pragma unroll
            for (int t = 0; t < RING_MAX; ++t) {
                if (t < Lv) {
                    const double2 w = W[(size_t)t * nf + f];
                    a.x += ev[t] * w.x;
                    a.y += ev[t] * w.y;
                }
            }
            A[((size_t)d * nf + f) * N + c] = a;
        }
    }
}

__global__ void k_advance(const ShotDev *__restrict__ D, int ns, int L)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns)
        *D[s].step += L;
}

// One wavefront per (row, frequency): sum_n x[n] * dt*exp(-a n dt)*exp(i a n dt),
// a = sqrt(omega0*omega).  Fixed per-lane order then a fixed butterfly: the
// result does not depend on scheduling.
__global__ __launch_bounds__(64) void k_receiver_dft(const double *__restrict__ X, int it,
                                                     double dt, double omega0, int nfreq,
                                                     const double *__restrict__ om,
                                                     double *__restrict__ out)
{
    const int row = blockIdx.y, f = blockIdx.x, lane = threadIdx.x;
    const double a = sqrt(omega0 * om[f]);
    const double *x = X + (size_t)row * it;
    double re = 0.0, im = 0.0;
    for (int q = lane; q < it; q += 64) {
        const double arg = a * q * dt;
        const double e = dt * exp(-arg);
        double sn, cs;
        sincos(arg, &sn, &cs);
        re += x[q] * (e * cs);
        im += x[q] * (e * sn);
    }
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
    }
    if (lane == 0) {
        out[2 * ((size_t)row * nfreq + f)] = re;
        out[2 * ((size_t)row * nfreq + f) + 1] = im;
    }
}

__global__ void k_fill(double *p, size_t count, double v)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < count;
         q += (size_t)gridDim.x * blockDim.x)
        p[q] = v;
}

// eps' = sigma/2/omega0 (satan_i/fwi3d_cpu_function_v1.4.c:335)
__global__ void k_eps(const double *__restrict__ sig, double *__restrict__ eps, int n, double omega0)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n)
        eps[c] = sig[c] / 2.0 / omega0;
}

#define LAUNCH_E(M)                                                                          \
    hipLaunchKernelGGL(k_efield<M>, dim3(nblocks), dim3(TX, TY), 0, st, d, ns, pre, j, nover, \
                       sign)
hipError_t launch_efield(int mode, const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                         int nover, int sign, hipStream_t st)
{
    if (nblocks <= 0)
        return hipSuccess;
    switch (mode) {
    case M_UPD | M_PSI | M_SRC: LAUNCH_E(M_UPD | M_PSI | M_SRC); break;
    case M_UPD | M_PSI: LAUNCH_E(M_UPD | M_PSI); break;
    case M_UPD | M_INPLACE: LAUNCH_E(M_UPD | M_INPLACE); break;
    case M_PSI | M_INPLACE: LAUNCH_E(M_PSI | M_INPLACE); break;
    case M_UPD | M_PSI | M_INPLACE: LAUNCH_E(M_UPD | M_PSI | M_INPLACE); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

#define LAUNCH_H(M) \
    hipLaunchKernelGGL(k_hfield<M>, dim3(nblocks), dim3(TX, TY), 0, st, d, ns, pre, j, sign)
hipError_t launch_hfield(int mode, const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                         int sign, hipStream_t st)
{
    if (nblocks <= 0)
        return hipSuccess;
    switch (mode) {
    case M_UPD | M_PSI: LAUNCH_H(M_UPD | M_PSI); break;
    case M_UPD | M_INPLACE: LAUNCH_H(M_UPD | M_INPLACE); break;
    case M_PSI | M_INPLACE: LAUNCH_H(M_PSI | M_INPLACE); break;
    case M_UPD | M_PSI | M_INPLACE: LAUNCH_H(M_UPD | M_PSI | M_INPLACE); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_source(const ShotDev *d, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_source, dim3(1), dim3(1), 0, st, d, n);
    return hipGetLastError();
}

hipError_t launch_receivers(const ShotDev *d, int ns, int j, int nover, int inplace,
                            hipStream_t st)
{
    if (ns <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_receivers, dim3(ns), dim3(256), 0, st, d, j, nover, inplace);
    return hipGetLastError();
}

hipError_t launch_flush(const ShotDev *d, int ns, const int *pre_cells, int nblocks, int L,
                        hipStream_t st)
{
    if (nblocks <= 0 || L <= 0)
        return hipSuccess;
    if (L > RING_MAX)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_flush, dim3(nblocks), dim3(256), 0, st, d, ns, pre_cells, L);
    return hipGetLastError();
}

hipError_t launch_advance(const ShotDev *d, int ns, int L, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance, dim3((ns + 63) / 64), dim3(64), 0, st, d, ns, L);
    return hipGetLastError();
}

hipError_t launch_receiver_dft(const double *traces, int nrows, int it, double dt, double omega0,
                               int nfreq, const double *omegas, double *out, hipStream_t st)
{
    if (nrows <= 0 || nfreq <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_receiver_dft, dim3(nfreq, nrows), dim3(64), 0, st, traces, it, dt,
                       omega0, nfreq, omegas, out);
    return hipGetLastError();
}

hipError_t launch_fill(double *p, size_t count, double v, hipStream_t st)
{
    if (count == 0)
        return hipSuccess;
    size_t blocks = (count + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, count, v);
    return hipGetLastError();
}

hipError_t launch_eps(const double *sig, double *eps, int n, double omega0, hipStream_t st)
{
    hipLaunchKernelGGL(k_eps, dim3((n + 255) / 256), dim3(256), 0, st, sig, eps, n, omega0);
    return hipGetLastError();
}

}  // namespace mt
