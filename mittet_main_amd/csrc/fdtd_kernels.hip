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

// Minimum resident waves per SIMD requested for the streaming kernels; it caps
// their VGPR budget (512 / waves) so enough waves are in flight to cover HBM
// latency.
#ifndef MT_Z_WAVES
#define MT_Z_WAVES 6
#endif
#ifndef MT_FLUSH_WAVES
#define MT_FLUSH_WAVES 4
#endif
#ifndef MT_FLUSH_G
#define MT_FLUSH_G 8
#endif


// Pointers read out of the shot descriptor are generic; casting them to the
// global address space lets the compiler emit global_load/store instead of
// flat accesses (which also count against the LDS wait counter).
#define GLOBAL __attribute__((address_space(1)))

namespace mt {

__device__ __forceinline__ const GLOBAL double *gc(const double *p) { return (const GLOBAL double *)p; }
__device__ __forceinline__ GLOBAL double *gm(double *p) { return (GLOBAL double *)p; }
__device__ __forceinline__ const GLOBAL int *gc(const int *p) { return (const GLOBAL int *)p; }
__device__ __forceinline__ const GLOBAL double *gc(const GLOBAL double *p) { return p; }
__device__ __forceinline__ GLOBAL double *gm(GLOBAL double *p) { return p; }

// Field values each half-step reads or writes once (own E or H, coefficient,
// the updated field, psi).  With NT they are non-temporal, so they do not
// evict the stencil neighbours that other tiles still read from L2 and the
// Infinity Cache; the host sets NT only for launches whose fields do not fit
// that cache anyway.
template <bool NT> __device__ __forceinline__ double ld_s(const GLOBAL double *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}
template <bool NT> __device__ __forceinline__ void st_s(GLOBAL double *p, double v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

__device__ __forceinline__ double dminus(const GLOBAL double *__restrict__ F, int c, int s,
                                         double c1, double c2)
{
    return c1 * F[c] - c1 * F[c - s] + c2 * F[c + s] - c2 * F[c - 2 * s];
}

__device__ __forceinline__ double dplus(const GLOBAL double *__restrict__ F, int c, int s,
                                        double c1, double c2)
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

// Absolute step of local index 0 of the current chunk.  Every lane of a launch
// reads the same word, which the previous chunk's k_advance wrote: a scalar
// load through the constant address space serves it from the scalar cache
// instead of a vector round trip to L2 per wave.
__device__ __forceinline__ int chunk_step(const ShotDev &S)
{
    return *(const __attribute__((address_space(4))) int *)S.step;
}

// A wave-uniform double (a z-profile value of the current plane) through the
// scalar cache.  A vector load issued after a plane's prefetch would make the
// wait for it also wait for every prefetch load issued before it.
__device__ __forceinline__ double uload(const double *p)
{
    return *(const __attribute__((address_space(4))) double *)p;
}

// One cell of the E update + CPML-E corrections + source, general path.
// Every load a cell needs is issued before the arithmetic that uses any of
// them (the loads a flag does not ask for are skipped and their values left
// zero), so a cell pays one memory round trip instead of one per branch.  The
// arithmetic is the same expressions in the same order as one branch after the
// other, so the result is unchanged to the bit.
__device__ __forceinline__ double dm(double f0, double fm1, double fp1, double fm2, double c1,
                                     double c2)
{
    return c1 * f0 - c1 * fm1 + c2 * fp1 - c2 * fm2;
}

template <int MODE>
__device__ __forceinline__ void efield_cell(const ShotDev &S, const double *H, int i, int jj, int k,
                                            int n, int si, int so, int sign)
{
    constexpr bool NT = (MODE & M_NT) != 0;
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc, N = S.n;
    const int sy = ix, sz = ix * iy;
    const int c = (k * iy + jj) * ix + i;

    const GLOBAL double *__restrict__ Ein = gc(S.e + (size_t)si * S.sstride);
    GLOBAL double *__restrict__ Eout = gm(S.e + (size_t)so * S.sstride);
    const GLOBAL double *__restrict__ Hx = gc(H);
    const GLOBAL double *__restrict__ Hy = gc(H + N);
    const GLOBAL double *__restrict__ Hz = gc(H + 2 * (size_t)N);

    const double c1 = S.c1, c2 = S.c2, dt = S.dt;
    const GLOBAL double *__restrict__ px = gc(S.prof[0]);
    const GLOBAL double *__restrict__ py = gc(S.prof[1]);
    const GLOBAL double *__restrict__ pz = gc(S.prof[2]);

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

    // ---- loads
    double ex = ld_s<NT>(Ein + c), ey = ld_s<NT>(Ein + N + c), ez = ld_s<NT>(Ein + 2 * (size_t)N + c);
    // every backward stencil in bounds: all six derivatives are loaded
    const bool all = rx2 && ry2 && rz2;
    const bool nzy = all || upx || fy, nyz = all || upx || fz, nxz = all || upy || fz;
    const bool nzx = all || upy || fx, nyx = all || upz || fx, nxy = all || upz || fy;
    double z0 = 0.0, zym1 = 0.0, zyp1 = 0.0, zym2 = 0.0, zxm1 = 0.0, zxp1 = 0.0, zxm2 = 0.0;
    double y0 = 0.0, yzm1 = 0.0, yzp1 = 0.0, yzm2 = 0.0, yxm1 = 0.0, yxp1 = 0.0, yxm2 = 0.0;
    double x0 = 0.0, xzm1 = 0.0, xzp1 = 0.0, xzm2 = 0.0, xym1 = 0.0, xyp1 = 0.0, xym2 = 0.0;
    if (nzy || nzx)
        z0 = Hz[c];
    if (nzy) {
        zym1 = Hz[c - sy];
        zyp1 = Hz[c + sy];
        zym2 = Hz[c - 2 * sy];
    }
    if (nzx) {
        zxm1 = Hz[c - 1];
        zxp1 = Hz[c + 1];
        zxm2 = Hz[c - 2];
    }
    if (nyz || nyx)
        y0 = Hy[c];
    if (nyz) {
        yzm1 = Hy[c - sz];
        yzp1 = Hy[c + sz];
        yzm2 = Hy[c - 2 * sz];
    }
    if (nyx) {
        yxm1 = Hy[c - 1];
        yxp1 = Hy[c + 1];
        yxm2 = Hy[c - 2];
    }
    if (nxz || nxy)
        x0 = Hx[c];
    if (nxz) {
        xzm1 = Hx[c - sz];
        xzp1 = Hx[c + sz];
        xzm2 = Hx[c - 2 * sz];
    }
    if (nxy) {
        xym1 = Hx[c - sy];
        xyp1 = Hx[c + sy];
        xym2 = Hx[c - 2 * sy];
    }
    const double eps2 = gc(S.eps)[c];
    const double sgx = px[PR_ESIG * ix + i], sgy = py[PR_ESIG * iy + jj], sgz = pz[PR_ESIG * iz + k];
    double kx = 1.0, ky = 1.0, kz = 1.0;
    if (MODE & M_UPD) {
        kx = px[PR_KED * ix + i];
        ky = py[PR_KED * iy + jj];
        kz = pz[PR_KED * iz + k];
    }
    // psi in place, or (pstride set) read from ring slot si and written to
    // slot so like E
    const size_t pr = S.pstride ? (size_t)si * S.pstride : 0;
    const size_t pw = S.pstride ? (size_t)so * S.pstride : 0;
    const GLOBAL double *Q0 = gc(S.psi[0] + pr), *Q1 = gc(S.psi[1] + pr), *Q2 = gc(S.psi[2] + pr);
    GLOBAL double *P0 = gm(S.psi[0] + pw), *P1 = gm(S.psi[1] + pw), *P2 = gm(S.psi[2] + pw);
    const size_t L0 = (size_t)S.slab[0], L1 = (size_t)S.slab[1], L2 = (size_t)S.slab[2];
    size_t q0 = 0, q1 = 0, q2 = 0;
    double bx = 0.0, cx = 0.0, ax0 = 0.0, ax1 = 0.0;
    double by = 0.0, cy = 0.0, ay0 = 0.0, ay1 = 0.0;
    double bz = 0.0, cz = 0.0, az0 = 0.0, az1 = 0.0;
    if (fx) {
        q0 = ((size_t)k * iy + jj) * (2 * nc) + slab_pos(i, ix, nc);
        bx = px[PR_BE * ix + i];
        cx = px[PR_CE * ix + i];
        ax0 = Q0[q0];
        ax1 = Q0[L0 + q0];
    }
    if (fy) {
        q1 = ((size_t)k * (2 * nc) + slab_pos(jj, iy, nc)) * ix + i;
        by = py[PR_BE * iy + jj];
        cy = py[PR_CE * iy + jj];
        ay0 = Q1[q1];
        ay1 = Q1[L1 + q1];
    }
    if (fz) {
        q2 = ((size_t)slab_pos(k, iz, nc) * iy + jj) * ix + i;
        bz = pz[PR_BE * iz + k];
        cz = pz[PR_CE * iz + k];
        az0 = Q2[q2];
        az1 = Q2[L2 + q2];
    }
    const bool src = (MODE & M_SRC) && c == S.src_cell;
    double amp = 0.0;
    if (src)
        amp = gc(S.amp)[n];

    // ---- arithmetic
    const double hzy = dm(z0, zym1, zyp1, zym2, c1, c2);
    const double hyz = dm(y0, yzm1, yzp1, yzm2, c1, c2);
    const double hxz = dm(x0, xzm1, xzp1, xzm2, c1, c2);
    const double hzx = dm(z0, zxm1, zxp1, zxm2, c1, c2);
    const double hyx = dm(y0, yxm1, yxp1, yxm2, c1, c2);
    const double hxy = dm(x0, xym1, xyp1, xym2, c1, c2);

    // cb = dt/eps'/(1 + s) and ca = (1 - s)/(1 + s) with s = sig dt/(2 eps');
    // where an axis' CPML sigma is zero they are exactly dt/eps' and 1, so the
    // divisions are skipped there (most shell cells lie in one axis' CPML)
    const double dte = dt / eps2;
    double cax = 1.0, cay = 1.0, caz = 1.0, cbx = dte, cby = dte, cbz = dte;
    if (sgx != 0.0) {
        const double se = sgx * dt / (2.0 * eps2);
        cbx = dte / (1.0 + se);
        cax = (1.0 - se) / (1.0 + se);
    }
    if (sgy != 0.0) {
        const double se = sgy * dt / (2.0 * eps2);
        cby = dte / (1.0 + se);
        cay = (1.0 - se) / (1.0 + se);
    }
    if (sgz != 0.0) {
        const double se = sgz * dt / (2.0 * eps2);
        cbz = dte / (1.0 + se);
        caz = (1.0 - se) / (1.0 + se);
    }

    if (upx) {
        const double cu = hzy / ky - hyz / kz;
        ex = sign > 0 ? cax * ex + cbx * cu : cax * ex - cbx * cu;
    }
    if (upy) {
        const double cu = hxz / kz - hzx / kx;
        ey = sign > 0 ? cay * ey + cby * cu : cay * ey - cby * cu;
    }
    if (upz) {
        const double cu = hyx / kx - hxy / ky;
        ez = sign > 0 ? caz * ez + cbz * cu : caz * ez - cbz * cu;
    }
    // The psi values are stored after all arithmetic: a store issued before a
    // later face's branch would make that branch's wait for its loads also
    // wait for the store to complete.
    double px0 = 0.0, px1 = 0.0, py0 = 0.0, py1 = 0.0, pz0 = 0.0, pz1 = 0.0;
    if (fx) {
        const double h = S.dl[0];
        px0 = bx * ax0 + cx * hyx / h;
        px1 = bx * ax1 + cx * hzx / h;
        ez += cbz * px0;
        ey -= cby * px1;
    }
    if (fy) {
        const double h = S.dl[1];
        py0 = by * ay0 + cy * hzy / h;
        py1 = by * ay1 + cy * hxy / h;
        ex += cbx * py0;
        ez -= cbz * py1;
    }
    if (fz) {
        const double h = S.dl[2];
        pz0 = bz * az0 + cz * hxz / h;
        pz1 = bz * az1 + cz * hyz / h;
        ey += cby * pz0;
        ex -= cbx * pz1;
    }
    if (src) {
        if (S.src_comp == 0)
            ex -= amp;
        else if (S.src_comp == 1)
            ey -= amp;
        else
            ez -= amp;
    }
    if (fx) {
        st_s<NT>(P0 + q0, px0);
        st_s<NT>(P0 + L0 + q0, px1);
    }
    if (fy) {
        st_s<NT>(P1 + q1, py0);
        st_s<NT>(P1 + L1 + q1, py1);
    }
    if (fz) {
        st_s<NT>(P2 + q2, pz0);
        st_s<NT>(P2 + L2 + q2, pz1);
    }
    st_s<NT>(Eout + c, ex);
    st_s<NT>(Eout + N + c, ey);
    st_s<NT>(Eout + 2 * (size_t)N + c, ez);
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
        n = chunk_step(S) + j;
        if (n >= S.it - 1)
            return;  // this shot has finished its time loop
    }
    const int b = blockIdx.x - pre[s];
    const int bx = b % S.nbx, t = b / S.nbx;
    const int by = t % S.nby, k = t / S.nby;
    const int i = bx * TX + threadIdx.x, jj = by * TY + threadIdx.y;
    if (i >= S.ix || jj >= S.iy)
        return;
    int si, so;
    if (MODE & M_INPLACE) {
        si = so = S.cur;
    } else {
        si = (S.cur + j) % S.ring;
        so = (S.cur + j + 1) % S.ring;
    }
    efield_cell<MODE>(S, hsel(S, 0, 0), i, jj, k, n, si, so, sign);
}

// One cell of the H update + CPML-H corrections, general path, with the loads
// issued first as in efield_cell.  H is read from Hin and written to Hout (the
// same buffer when updated in place).
__device__ __forceinline__ double dp(double fp1, double f0, double fp2, double fm1, double c1,
                                     double c2)
{
    return c1 * fp1 - c1 * f0 + c2 * fp2 - c2 * fm1;
}

template <int MODE>
__device__ __forceinline__ void hfield_cell(const ShotDev &S, const double *Hin, double *Hout, int i,
                                            int jj, int k, int so, int sign)
{
    constexpr bool NT = (MODE & M_NT) != 0;
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc, N = S.n;
    const int sy = ix, sz = ix * iy;
    const int c = (k * iy + jj) * ix + i;

    const GLOBAL double *__restrict__ Ex = gc(S.e + (size_t)so * S.sstride);
    const GLOBAL double *__restrict__ Ey = gc(Ex + N);
    const GLOBAL double *__restrict__ Ez = gc(Ex + 2 * (size_t)N);

    const double c1 = S.c1, c2 = S.c2, dt = S.dt;
    const GLOBAL double *__restrict__ px = gc(S.prof[0]);
    const GLOBAL double *__restrict__ py = gc(S.prof[1]);
    const GLOBAL double *__restrict__ pz = gc(S.prof[2]);

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

    // ---- loads
    // every forward stencil in bounds: all six derivatives are loaded
    const bool all = ax2 && ay2 && az2;
    const bool nzy = all || upx || gy, nyz = all || upx || gz, nxz = all || upy || gz;
    const bool nzx = all || upy || gx, nyx = all || upz || gx, nxy = all || upz || gy;
    double z0 = 0.0, zyp1 = 0.0, zyp2 = 0.0, zym1 = 0.0, zxp1 = 0.0, zxp2 = 0.0, zxm1 = 0.0;
    double y0 = 0.0, yzp1 = 0.0, yzp2 = 0.0, yzm1 = 0.0, yxp1 = 0.0, yxp2 = 0.0, yxm1 = 0.0;
    double x0 = 0.0, xzp1 = 0.0, xzp2 = 0.0, xzm1 = 0.0, xyp1 = 0.0, xyp2 = 0.0, xym1 = 0.0;
    if (nzy || nzx)
        z0 = Ez[c];
    if (nzy) {
        zyp1 = Ez[c + sy];
        zyp2 = Ez[c + 2 * sy];
        zym1 = Ez[c - sy];
    }
    if (nzx) {
        zxp1 = Ez[c + 1];
        zxp2 = Ez[c + 2];
        zxm1 = Ez[c - 1];
    }
    if (nyz || nyx)
        y0 = Ey[c];
    if (nyz) {
        yzp1 = Ey[c + sz];
        yzp2 = Ey[c + 2 * sz];
        yzm1 = Ey[c - sz];
    }
    if (nyx) {
        yxp1 = Ey[c + 1];
        yxp2 = Ey[c + 2];
        yxm1 = Ey[c - 1];
    }
    if (nxz || nxy)
        x0 = Ex[c];
    if (nxz) {
        xzp1 = Ex[c + sz];
        xzp2 = Ex[c + 2 * sz];
        xzm1 = Ex[c - sz];
    }
    if (nxy) {
        xyp1 = Ex[c + sy];
        xyp2 = Ex[c + 2 * sy];
        xym1 = Ex[c - sy];
    }
    const double msx = px[PR_MSIG * ix + i], msy = py[PR_MSIG * iy + jj], msz = pz[PR_MSIG * iz + k];
    // eps' is only used where a magnetic CPML conductivity is nonzero; loading
    // it beside the conductivities rather than after them saves a round trip
    const double eps2 = gc(S.eps)[c];
    const GLOBAL double *__restrict__ Hi = gc(Hin);
    double hx = ld_s<NT>(Hi + c), hy = ld_s<NT>(Hi + N + c), hz = ld_s<NT>(Hi + 2 * (size_t)N + c);
    double kx = 1.0, ky = 1.0, kz = 1.0;
    if (MODE & M_UPD) {
        kx = px[PR_KHD * ix + i];
        ky = py[PR_KHD * iy + jj];
        kz = pz[PR_KHD * iz + k];
    }
    // psi in place, or (pstride set) read from the ring slot before so and
    // written to slot so like H
    const size_t pr = S.pstride ? (size_t)((so + S.ring - 1) % S.ring) * S.pstride : 0;
    const size_t pw = S.pstride ? (size_t)so * S.pstride : 0;
    const GLOBAL double *Q0 = gc(S.psi[0] + pr), *Q1 = gc(S.psi[1] + pr), *Q2 = gc(S.psi[2] + pr);
    GLOBAL double *P0 = gm(S.psi[0] + pw), *P1 = gm(S.psi[1] + pw), *P2 = gm(S.psi[2] + pw);
    const size_t L0 = (size_t)S.slab[0], L1 = (size_t)S.slab[1], L2 = (size_t)S.slab[2];
    size_t q0 = 0, q1 = 0, q2 = 0;
    double bx = 0.0, cx = 0.0, ax0 = 0.0, ax1v = 0.0;
    double by = 0.0, cy = 0.0, ay0 = 0.0, ay1v = 0.0;
    double bz = 0.0, cz = 0.0, az0 = 0.0, az1v = 0.0;
    if (gx) {
        q0 = ((size_t)k * iy + jj) * (2 * nc) + slab_pos(i, ix, nc);
        bx = px[PR_BH * ix + i];
        cx = px[PR_CH * ix + i];
        ax0 = Q0[2 * L0 + q0];
        ax1v = Q0[3 * L0 + q0];
    }
    if (gy) {
        q1 = ((size_t)k * (2 * nc) + slab_pos(jj, iy, nc)) * ix + i;
        by = py[PR_BH * iy + jj];
        cy = py[PR_CH * iy + jj];
        ay0 = Q1[2 * L1 + q1];
        ay1v = Q1[3 * L1 + q1];
    }
    if (gz) {
        q2 = ((size_t)slab_pos(k, iz, nc) * iy + jj) * ix + i;
        bz = pz[PR_BH * iz + k];
        cz = pz[PR_CH * iz + k];
        az0 = Q2[2 * L2 + q2];
        az1v = Q2[3 * L2 + q2];
    }

    // ---- arithmetic
    const double ezy = dp(zyp1, z0, zyp2, zym1, c1, c2);
    const double eyz = dp(yzp1, y0, yzp2, yzm1, c1, c2);
    const double exz = dp(xzp1, x0, xzp2, xzm1, c1, c2);
    const double ezx = dp(zxp1, z0, zxp2, zxm1, c1, c2);
    const double eyx = dp(yxp1, y0, yxp2, yxm1, c1, c2);
    const double exy = dp(xyp1, x0, xyp2, xym1, c1, c2);

    // db = dt/mu0/(1 + s) and da = (1 - s)/(1 + s) with s = msig dt/(2 eps');
    // exactly dt/mu0 and 1 where an axis' magnetic CPML sigma is zero
    const double dtm = dt / S.mu0;
    double dax = 1.0, day = 1.0, daz = 1.0, dbx = dtm, dby = dtm, dbz = dtm;
    if (msx != 0.0) {
        const double sm = msx * dt / (2.0 * eps2);
        dbx = dtm / (1.0 + sm);
        dax = (1.0 - sm) / (1.0 + sm);
    }
    if (msy != 0.0) {
        const double sm = msy * dt / (2.0 * eps2);
        dby = dtm / (1.0 + sm);
        day = (1.0 - sm) / (1.0 + sm);
    }
    if (msz != 0.0) {
        const double sm = msz * dt / (2.0 * eps2);
        dbz = dtm / (1.0 + sm);
        daz = (1.0 - sm) / (1.0 + sm);
    }

    if (upx) {
        const double cu = ezy / ky - eyz / kz;
        hx = sign > 0 ? dax * hx - dbx * cu : dax * hx + dbx * cu;
    }
    if (upy) {
        const double cu = exz / kz - ezx / kx;
        hy = sign > 0 ? day * hy - dby * cu : day * hy + dby * cu;
    }
    if (upz) {
        const double cu = eyx / kx - exy / ky;
        hz = sign > 0 ? daz * hz - dbz * cu : daz * hz + dbz * cu;
    }
    // psi stored after all arithmetic, as in efield_cell
    double px0 = 0.0, px1 = 0.0, py0 = 0.0, py1 = 0.0, pz0 = 0.0, pz1 = 0.0;
    if (gx) {
        const double h = S.dl[0];
        px0 = bx * ax0 + cx * ezx / h;
        px1 = bx * ax1v + cx * eyx / h;
        hy += dby * px0;
        hz -= dbz * px1;
    }
    if (gy) {
        const double h = S.dl[1];
        py0 = by * ay0 + cy * exy / h;
        py1 = by * ay1v + cy * ezy / h;
        hz += dbz * py0;
        hx -= dbx * py1;
    }
    if (gz) {
        const double h = S.dl[2];
        pz0 = bz * az0 + cz * eyz / h;
        pz1 = bz * az1v + cz * exz / h;
        hx += dbx * pz0;
        hy -= dby * pz1;
    }
    if (gx) {
        st_s<NT>(P0 + 2 * L0 + q0, px0);
        st_s<NT>(P0 + 3 * L0 + q0, px1);
    }
    if (gy) {
        st_s<NT>(P1 + 2 * L1 + q1, py0);
        st_s<NT>(P1 + 3 * L1 + q1, py1);
    }
    if (gz) {
        st_s<NT>(P2 + 2 * L2 + q2, pz0);
        st_s<NT>(P2 + 3 * L2 + q2, pz1);
    }
    GLOBAL double *__restrict__ Ho = gm(Hout);
    st_s<NT>(Ho + c, hx);
    st_s<NT>(Ho + N + c, hy);
    st_s<NT>(Ho + 2 * (size_t)N + c, hz);
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
        if (chunk_step(S) + j >= S.it - 1)
            return;
        so = (S.cur + j + 1) % S.ring;
    }
    const int b = blockIdx.x - pre[s];
    const int bx = b % S.nbx, t = b / S.nbx;
    const int by = t % S.nby, k = t / S.nby;
    const int i = bx * TX + threadIdx.x, jj = by * TY + threadIdx.y;
    if (i >= S.ix || jj >= S.iy)
        return;
    double *H = hsel(S, 0, 0);
    hfield_cell<MODE>(S, H, H, i, jj, k, so, sign);
}

// read_source_3d as a stand-alone operator (in place on the current slot)
__global__ void k_source(const ShotDev *__restrict__ D, int n)
{
    const ShotDev &S = D[0];
    if (S.src_cell < 0 || n < 0 || n >= S.it - 1)
        return;
    double *E = S.e + (size_t)S.cur * S.sstride + (size_t)S.src_comp * S.n;
    E[S.src_cell] -= gc(S.amp)[n];
}

// Trace samples of step n: E from ring slot so and H from buffer H, one block
// per shot.
__device__ __forceinline__ void rec_store(const ShotDev &S, int n, int so, const double *H)
{
    const int tot = 6 * S.nrec;
    for (int q = threadIdx.x; q < tot; q += blockDim.x) {
        const int comp = q / S.nrec, r = q - comp * S.nrec;
        const int cell = S.rec_cell[r];
        const double v = comp < 3 ? S.e[(size_t)so * S.sstride + (size_t)comp * S.n + cell]
                                  : H[(size_t)(comp - 3) * S.n + cell];
        S.traces[((size_t)comp * S.nrec + r) * S.it + n] = v;
        if (!__builtin_isfinite(v))
            *S.flag = 1;
    }
}

// Receiver gather: E (from the slot written this step) and H after the step
// (hf: the step was a fused E+H launch sequence, see hsel).
__global__ void k_receivers(const ShotDev *__restrict__ D, int j, int nover, int inplace, int hf)
{
    const ShotDev &S = D[blockIdx.x];
    int n = nover, so = S.cur;
    if (n < 0) {
        n = chunk_step(S) + j;
        if (n >= S.it - 1)
            return;
    }
    if (!inplace)
        so = (S.cur + j + 1) % S.ring;
    rec_store(S, n, so, inplace ? hsel(S, 0, 0) : hsel(S, j + 1, hf));
}

// Volume running DFT over the last L ring slots: acc += E[n] * w[n] in step
// order.  All NF accumulators of a component stay in registers while the slots
// stream past, so every frequency is an independent dependency chain; each
// accumulator still receives its terms in ascending step order (bit-identical
// to the per-step running sum).
// The chunk's weights (<= RING_MAX steps x FLUSH_NF_MAX frequencies) are staged
// in LDS once per block and read back as wave-uniform LDS broadcasts.  A block
// covers FLUSH_CELLS cells: two adjacent cells per lane with 16-byte loads
// where the strides are even (one weight read serves both; at 288^3 with 8
// frequencies the flush takes 17.4-18.1 ms per 128 steps against 19.0-19.6 ms
// for one cell per lane), else two lane-contiguous halves one cell at a time.
constexpr int FLUSH_NF_MAX = 8;

template <int NF>
__device__ __forceinline__ void flush_cell(const ShotDev &S, const double *src, double *acc, int c,
                                           int Lv, const double2 *__restrict__ wsh)
{
    constexpr int G = MT_FLUSH_G;  // samples loaded together: G loads in flight per lane
    const int N = S.n, ring = S.ring, cur = S.cur;
    GLOBAL double *__restrict__ A = gm(acc);
    for (int d = 0; d < 3; ++d) {
        double2 a[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            a[f].x = A[2 * (((size_t)d * NF + f) * S.astride + c)];
            a[f].y = A[2 * (((size_t)d * NF + f) * S.astride + c) + 1];
        }
        for (int t0 = 0; t0 < Lv; t0 += G) {
            double e[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int slot = (cur + 1 + t0 + g) % ring;
                e[g] = t0 + g < Lv ? gc(src)[(size_t)slot * S.sstride + (size_t)d * N + c] : 0.0;
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (t0 + g < Lv) {
#pragma unroll
                    for (int f = 0; f < NF; ++f) {
                        const double2 w = wsh[(t0 + g) * NF + f];
                        a[f].x += e[g] * w.x;
                        a[f].y += e[g] * w.y;
                    }
                }
            }
        }
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            A[2 * (((size_t)d * NF + f) * S.astride + c)] = a[f].x;
            A[2 * (((size_t)d * NF + f) * S.astride + c) + 1] = a[f].y;
        }
    }
}

// Two adjacent cells per lane (c even; n, slot and accumulator strides even,
// so every access is a 16-byte one): the same sums as flush_cell for each
// cell, with one weight read from LDS per sample for both.
template <int NF>
__device__ __forceinline__ void flush_pair(const ShotDev &S, const double *src, double *acc, int c,
                                           int Lv, const double2 *__restrict__ wsh)
{
    constexpr int G = MT_FLUSH_G / 2;
    const int N = S.n, ring = S.ring, cur = S.cur;
    GLOBAL double *__restrict__ A = gm(acc);
    for (int d = 0; d < 3; ++d) {
        double2 a[NF], b[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const GLOBAL double2 *p =
                (const GLOBAL double2 *)(A + 2 * (((size_t)d * NF + f) * S.astride + c));
            a[f].x = p[0].x;
            a[f].y = p[0].y;
            b[f].x = p[1].x;
            b[f].y = p[1].y;
        }
        for (int t0 = 0; t0 < Lv; t0 += G) {
            double2 e[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int slot = (cur + 1 + t0 + g) % ring;
                double2 v{0.0, 0.0};
                if (t0 + g < Lv) {
                    const GLOBAL double2 *p = (const GLOBAL double2 *)(gc(src) + (size_t)slot * S.sstride +
                                                                       (size_t)d * N + c);
                    v.x = (*p).x;
                    v.y = (*p).y;
                }
                e[g] = v;
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (t0 + g < Lv) {
#pragma unroll
                    for (int f = 0; f < NF; ++f) {
                        const double2 w = wsh[(t0 + g) * NF + f];
                        a[f].x += e[g].x * w.x;
                        a[f].y += e[g].x * w.y;
                        b[f].x += e[g].y * w.x;
                        b[f].y += e[g].y * w.y;
                    }
                }
            }
        }
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            GLOBAL double *p = A + 2 * (((size_t)d * NF + f) * S.astride + c);
            p[0] = a[f].x;
            p[1] = a[f].y;
            p[2] = b[f].x;
            p[3] = b[f].y;
        }
    }
}

// any frequency count: groups of 4 accumulators, weights from global memory
__device__ __forceinline__ void flush_cell_any(const ShotDev &S, const double *src, double *acc,
                                               int c, int n0, int Lv)
{
    const int N = S.n, ring = S.ring, cur = S.cur, nf = S.nf;
    const GLOBAL double *__restrict__ W = gc(S.w) + 2 * (size_t)n0 * nf;
    GLOBAL double *__restrict__ A = gm(acc);
    for (int d = 0; d < 3; ++d)
        for (int f0 = 0; f0 < nf; f0 += 4) {
            const int nq = nf - f0 < 4 ? nf - f0 : 4;
            double2 a[4];
            for (int q = 0; q < nq; ++q) {
                a[q].x = A[2 * (((size_t)d * nf + f0 + q) * S.astride + c)];
                a[q].y = A[2 * (((size_t)d * nf + f0 + q) * S.astride + c) + 1];
            }
            int slot = cur;
            for (int t = 0; t < Lv; ++t) {
                slot = slot + 1 == ring ? 0 : slot + 1;
                const double e = gc(src)[(size_t)slot * S.sstride + (size_t)d * N + c];
                for (int q = 0; q < nq; ++q) {
                    const double wx = W[2 * ((size_t)t * nf + f0 + q)];
                    const double wy = W[2 * ((size_t)t * nf + f0 + q) + 1];
                    a[q].x += e * wx;
                    a[q].y += e * wy;
                }
            }
            for (int q = 0; q < nq; ++q) {
                A[2 * (((size_t)d * nf + f0 + q) * S.astride + c)] = a[q].x;
                A[2 * (((size_t)d * nf + f0 + q) * S.astride + c) + 1] = a[q].y;
            }
        }
}

// HS = 0: the E ring into the E accumulators; HS = 1: the H ring into the H
// accumulators (shots without H sums return), one instance each, so the E
// flush compiles as it does without the H sums.
template <int HS>
__global__ __launch_bounds__(256, MT_FLUSH_WAVES) void k_flush(const ShotDev *__restrict__ D, int ns,
                                                               const int *__restrict__ pre, int L)
{
    __shared__ double2 wsh[RING_MAX * FLUSH_NF_MAX];
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    const int nf = S.nf;
    const int n0 = chunk_step(S);
    const double *src = HS ? S.hring : S.e;
    double *acc = HS ? S.acc_h : S.acc;
    int Lv = L;
    if (n0 + Lv > S.it - 1)
        Lv = S.it - 1 - n0;
    if (nf == 0 || Lv <= 0 || !src || !acc)
        return;  // uniform over the block
    // a block covers FLUSH_CELLS = 2 * blockDim.x cells
    const int base = (blockIdx.x - pre[s]) * FLUSH_CELLS;
    const int ca = base + threadIdx.x, cb = ca + blockDim.x;   // lane-contiguous halves
    if (nf > FLUSH_NF_MAX || (nf != 1 && nf != 2 && nf != 3 && nf != 4 && nf != 8)) {
        if (ca < S.n)
            flush_cell_any(S, src, acc, ca, n0, Lv);
        if (cb < S.n)
            flush_cell_any(S, src, acc, cb, n0, Lv);
        return;
    }
    const GLOBAL double *__restrict__ W = gc(S.w) + 2 * (size_t)n0 * nf;
    for (int q = threadIdx.x; q < Lv * nf; q += blockDim.x) {
        wsh[q].x = W[2 * q];
        wsh[q].y = W[2 * q + 1];
    }
    __syncthreads();
    if ((S.n & 1) == 0 && (S.sstride & 1) == 0 && (S.astride & 1) == 0) {
        const int c = base + 2 * threadIdx.x;
        if (c >= S.n)
            return;
        switch (nf) {
        case 1: flush_pair<1>(S, src, acc, c, Lv, wsh); break;
        case 2: flush_pair<2>(S, src, acc, c, Lv, wsh); break;
        case 3: flush_pair<3>(S, src, acc, c, Lv, wsh); break;
        case 4: flush_pair<4>(S, src, acc, c, Lv, wsh); break;
        default: flush_pair<8>(S, src, acc, c, Lv, wsh); break;
        }
        return;
    }
    for (int c = ca; c <= cb; c += blockDim.x) {
        if (c >= S.n)
            return;
        switch (nf) {
        case 1: flush_cell<1>(S, src, acc, c, Lv, wsh); break;
        case 2: flush_cell<2>(S, src, acc, c, Lv, wsh); break;
        case 3: flush_cell<3>(S, src, acc, c, Lv, wsh); break;
        case 4: flush_cell<4>(S, src, acc, c, Lv, wsh); break;
        default: flush_cell<8>(S, src, acc, c, Lv, wsh); break;
        }
    }
}

// H volume running DFT, staging: H after step j's H update (the sample the H
// receiver traces take) into slot (cur + j + 1) % ring of the H ring, the slot
// the E half-step of that step writes in the E ring.  k_flush (hs = 1) then
// sums the chunk's slots with the E sums' weights, each accumulator in
// ascending step order, so the sums equal per-step running sums bit for bit.
__global__ __launch_bounds__(256) void k_hstore(const ShotDev *__restrict__ D, int ns,
                                                const int *__restrict__ pre, int j, int hf)
{
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    const int n = chunk_step(S) + j;
    if (!S.hring || n >= S.it - 1)
        return;
    const size_t N = (size_t)S.n;
    GLOBAL double *__restrict__ dst = gm(S.hring + (size_t)((S.cur + j + 1) % S.ring) * S.sstride);
    const GLOBAL double *__restrict__ H = gc(hsel(S, j + 1, hf));
    const int base = (blockIdx.x - pre[s]) * FLUSH_CELLS;
    for (int c = base + threadIdx.x; c < base + FLUSH_CELLS && c < S.n; c += blockDim.x) {
        dst[c] = H[c];
        dst[N + c] = H[N + c];
        dst[2 * N + c] = H[2 * N + c];
    }
}

hipError_t launch_hstore(const ShotDev *d, int ns, const int *pre_cells, int nblocks, int j,
                         hipStream_t st, int hf)
{
    if (nblocks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_hstore, dim3(nblocks), dim3(256), 0, st, d, ns, pre_cells, j, hf);
    return hipGetLastError();
}

// Step counters forward by L; then the divergence probe: E (the slot the
// chunk ended in) and H at the source cell (the first Hz point of a
// Fortran-variant shot) must be finite.
__global__ void k_advance(const ShotDev *__restrict__ D, int ns, int L, int hf)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns)
        return;
    const ShotDev &S = D[s];
    *S.step += L;
    const int c = S.src_cell >= 0 ? S.src_cell : (S.nsrc_h > 0 ? S.src_h[0] : -1);
    if (c < 0)
        return;
    const double *E = S.e + (size_t)((S.cur + L) % S.ring) * S.sstride;
    const double *H = hsel(S, L, hf);
    double v = 0.0;
    for (int d = 0; d < 3; ++d)
        v += __builtin_fabs(E[(size_t)d * S.n + c]) + __builtin_fabs(H[(size_t)d * S.n + c]);
    if (!__builtin_isfinite(v))
        *S.flag = 1;
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

__global__ void k_diag_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n2)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n2;
         q += (size_t)gridDim.x * blockDim.x)
        dst[q] = src[q];
}

// element copy: any count, any 8-byte-aligned addresses
__global__ void k_copy(const double *__restrict__ src, double *__restrict__ dst, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x)
        dst[q] = src[q];
}

// eps' = sigma/2/omega0 (satan_i/fwi3d_cpu_function_v1.4.c:335) and dt/eps', the
// E coefficient cb wherever the CPML conductivity is zero (function.c:415-417
// with esig = 0 reduces to dt/eps'/1.0).
__global__ void k_eps(const double *__restrict__ sig, double *__restrict__ eps,
                      double *__restrict__ cbi, int n, double omega0, double dt)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) {
        const double e = sig[c] / 2.0 / omega0;
        eps[c] = e;
        cbi[c] = dt / e;
    }
}


// ------------------------------------------------------------------------
// z-marching kernels of the fused step.  One thread per (i, j) column of a
// 32 x 8 tile walks a z-chunk; the z-stencil operands stay in a 4-deep register
// window, x/y neighbours come through L1/L2.  Blocks are numbered so that each
// XCD (blockIdx % 8) owns a contiguous run of tiles: neighbouring tiles share
// an L2.  Arithmetic is the reference's, operation for operation; where the
// CPML conductivity of a component is zero its coefficients reduce exactly to
// ca = 1, cb = dt/eps' (E) and da = 1, db = dt/mu0 (H), taken without divisions.

struct ZCol {
    int i, jj, k0, k1;
};

__device__ __forceinline__ bool zcol(const ShotDev &S, int b, ZCol &z)
{
    const int L = (b & 7) * (S.zb_pad >> 3) + (b >> 3);
    if (L >= S.zb)
        return false;
    const int tx = L % S.ztx, r = L / S.ztx;
    const int ty = r % S.zty, kz = r / S.zty;
    z.i = tx * ZX + (threadIdx.x & (ZX - 1));
    z.jj = ty * ZY + (threadIdx.x / ZX);
    z.k0 = kz * S.kb;
    z.k1 = min(z.k0 + S.kb, S.iz);
    return z.i < S.ix && z.jj < S.iy;
}

__global__ __launch_bounds__(ZX *ZY, MT_Z_WAVES) void k_efield_z(const ShotDev *__restrict__ D, int ns,
                                                     const int *__restrict__ pre, int j, int sign)
{
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    const int n = chunk_step(S) + j;
    if (n >= S.it - 1)
        return;
    ZCol z;
    if (!zcol(S, blockIdx.x - pre[s], z))
        return;
    const int i = z.i, jj = z.jj;
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc;
    const size_t N = (size_t)S.n;
    const int sy = ix, sz = ix * iy, cxy = jj * ix + i;
    const int si = (S.cur + j) % S.ring, so = (S.cur + j + 1) % S.ring;
    const GLOBAL double *__restrict__ Ein = gc(S.e + (size_t)si * S.sstride);
    GLOBAL double *__restrict__ Eout = gm(S.e + (size_t)so * S.sstride);
    const double *Hc = hsel(S, 0, 0);
    const GLOBAL double *__restrict__ Hx = gc(Hc);
    const GLOBAL double *__restrict__ Hy = gc(Hc + N);
    const GLOBAL double *__restrict__ Hz = gc(Hc + 2 * N);
    const double c1 = S.c1, c2 = S.c2, dt = S.dt;
    const GLOBAL double *__restrict__ px = gc(S.prof[0]);
    const GLOBAL double *__restrict__ py = gc(S.prof[1]);
    const GLOBAL double *__restrict__ pz = gc(S.prof[2]);

    const bool rx1 = i >= 1 && i < ix - 1, rx2 = i >= 2 && i < ix - 1;
    const bool ry1 = jj >= 1 && jj < iy - 1, ry2 = jj >= 2 && jj < iy - 1;
    const bool fxc = ((i >= 2 && i < nc - 1) || (i >= ix - nc + 1 && i < ix - 2)) && jj < iy - 1;
    const bool fyc = ((jj >= 2 && jj < nc) || (jj >= iy - nc + 1 && jj < iy - 2)) && i < ix - 1;
    const double esx = px[PR_ESIG * ix + i], esy = py[PR_ESIG * iy + jj];
    const double kedx = px[PR_KED * ix + i], kedy = py[PR_KED * iy + jj];

    // z window of Hx and Hy at k-2 .. k+1
    double xm2 = 0.0, xm1 = 0.0, x0, ym2 = 0.0, ym1 = 0.0, y0;
    {
        const int c = z.k0 * sz + cxy;
        if (z.k0 >= 2) {
            xm2 = Hx[c - 2 * sz];
            ym2 = Hy[c - 2 * sz];
        }
        if (z.k0 >= 1) {
            xm1 = Hx[c - sz];
            ym1 = Hy[c - sz];
        }
        x0 = Hx[c];
        y0 = Hy[c];
    }
    for (int k = z.k0; k < z.k1; ++k) {
        const int c = k * sz + cxy;
        double xp1 = 0.0, yp1 = 0.0;
        if (k + 1 < iz) {
            xp1 = Hx[c + sz];
            yp1 = Hy[c + sz];
        }
        const bool rz1 = k >= 1 && k < iz - 1, rz2 = k >= 2 && k < iz - 1;
        const bool upx = rx1 && ry2 && rz2, upy = rx2 && ry1 && rz2, upz = rx2 && ry2 && rz1;
        const bool fx = fxc && k < iz - 1, fy = fyc && k < iz - 1;
        const bool fz = ((k >= 2 && k < nc) || (k >= iz - nc + 1 && k < iz - 2)) && i < ix - 1 &&
                        jj < iy - 1;
        double ex = Ein[c], ey = Ein[N + c], ez = Ein[2 * N + c];

        double hzy = 0.0, hyz = 0.0, hxz = 0.0, hzx = 0.0, hyx = 0.0, hxy = 0.0;
        if (upx || fy) hzy = dminus(Hz, c, sy, c1, c2);
        if (upx || fz) hyz = c1 * y0 - c1 * ym1 + c2 * yp1 - c2 * ym2;
        if (upy || fz) hxz = c1 * x0 - c1 * xm1 + c2 * xp1 - c2 * xm2;
        if (upy || fx) hzx = dminus(Hz, c, 1, c1, c2);
        if (upz || fx) hyx = c1 * y0 - c1 * Hy[c - 1] + c2 * Hy[c + 1] - c2 * Hy[c - 2];
        if (upz || fy) hxy = c1 * x0 - c1 * Hx[c - sy] + c2 * Hx[c + sy] - c2 * Hx[c - 2 * sy];

        const double esz = pz[PR_ESIG * iz + k];
        const double cbi = gc(S.cbi)[c];
        double cbx = cbi, cby = cbi, cbz = cbi, sex = 0.0, sey = 0.0, sez = 0.0;
        if (esx != 0.0 || esy != 0.0 || esz != 0.0) {
            const double eps2 = gc(S.eps)[c];
            sex = esx * dt / (2.0 * eps2);
            sey = esy * dt / (2.0 * eps2);
            sez = esz * dt / (2.0 * eps2);
            cbx = dt / eps2 / (1.0 + sex);
            cby = dt / eps2 / (1.0 + sey);
            cbz = dt / eps2 / (1.0 + sez);
        }
        const double kedz = uload(S.prof[2] + PR_KED * iz + k);
        if (upx) {
            const double ca = esx != 0.0 ? (1.0 - sex) / (1.0 + sex) : 1.0;
            const double cu = hzy / kedy - hyz / kedz;
            ex = sign > 0 ? ca * ex + cbx * cu : ca * ex - cbx * cu;
        }
        if (upy) {
            const double ca = esy != 0.0 ? (1.0 - sey) / (1.0 + sey) : 1.0;
            const double cu = hxz / kedz - hzx / kedx;
            ey = sign > 0 ? ca * ey + cby * cu : ca * ey - cby * cu;
        }
        if (upz) {
            const double ca = esz != 0.0 ? (1.0 - sez) / (1.0 + sez) : 1.0;
            const double cu = hyx / kedx - hxy / kedy;
            ez = sign > 0 ? ca * ez + cbz * cu : ca * ez - cbz * cu;
        }
        if (fx) {
            GLOBAL double *__restrict__ P = gm(S.psi[0]);
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
            GLOBAL double *__restrict__ P = gm(S.psi[1]);
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
            GLOBAL double *__restrict__ P = gm(S.psi[2]);
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
        if (c == S.src_cell) {
            const double a = gc(S.amp)[n];
            if (S.src_comp == 0)
                ex -= a;
            else if (S.src_comp == 1)
                ey -= a;
            else
                ez -= a;
        }
        Eout[c] = ex;
        Eout[N + c] = ey;
        Eout[2 * N + c] = ez;
        xm2 = xm1;
        xm1 = x0;
        x0 = xp1;
        ym2 = ym1;
        ym1 = y0;
        y0 = yp1;
    }
}

__global__ __launch_bounds__(ZX *ZY, MT_Z_WAVES) void k_hfield_z(const ShotDev *__restrict__ D, int ns,
                                                     const int *__restrict__ pre, int j, int sign)
{
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    if (chunk_step(S) + j >= S.it - 1)
        return;
    ZCol z;
    if (!zcol(S, blockIdx.x - pre[s], z))
        return;
    const int i = z.i, jj = z.jj;
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc;
    const size_t N = (size_t)S.n;
    const int sy = ix, sz = ix * iy, cxy = jj * ix + i;
    const int so = (S.cur + j + 1) % S.ring;
    const GLOBAL double *__restrict__ Ex = gc(S.e + (size_t)so * S.sstride);
    const GLOBAL double *__restrict__ Ey = gc(Ex + N);
    const GLOBAL double *__restrict__ Ez = gc(Ex + 2 * N);
    double *Hc = hsel(S, 0, 0);
    GLOBAL double *__restrict__ Hx = gm(Hc);
    GLOBAL double *__restrict__ Hy = gm(Hc + N);
    GLOBAL double *__restrict__ Hz = gm(Hc + 2 * N);
    const double c1 = S.c1, c2 = S.c2, dt = S.dt;
    const double dbi = dt / S.mu0 / 1.0;   // db where the magnetic CPML sigma is zero
    const GLOBAL double *__restrict__ px = gc(S.prof[0]);
    const GLOBAL double *__restrict__ py = gc(S.prof[1]);
    const GLOBAL double *__restrict__ pz = gc(S.prof[2]);

    const bool ax1 = i >= 1 && i < ix - 1, ax2 = i >= 1 && i < ix - 2;
    const bool ay1 = jj >= 1 && jj < iy - 1, ay2 = jj >= 1 && jj < iy - 2;
    const bool gxc = ((i >= 1 && i < nc - 1) || (i >= ix - nc && i < ix - 2)) && jj < iy - 1;
    const bool gyc = ((jj >= 1 && jj < nc - 1) || (jj >= iy - nc && jj < iy - 2)) && i < ix - 1;
    const double msx = px[PR_MSIG * ix + i], msy = py[PR_MSIG * iy + jj];
    const double khdx = px[PR_KHD * ix + i], khdy = py[PR_KHD * iy + jj];

    // z window of Ex and Ey at k-1 .. k+2
    double xm1 = 0.0, x0, xp1 = 0.0, ym1 = 0.0, y0, yp1 = 0.0;
    {
        const int c = z.k0 * sz + cxy;
        if (z.k0 >= 1) {
            xm1 = Ex[c - sz];
            ym1 = Ey[c - sz];
        }
        x0 = Ex[c];
        y0 = Ey[c];
        if (z.k0 + 1 < iz) {
            xp1 = Ex[c + sz];
            yp1 = Ey[c + sz];
        }
    }
    for (int k = z.k0; k < z.k1; ++k) {
        const int c = k * sz + cxy;
        double xp2 = 0.0, yp2 = 0.0;
        if (k + 2 < iz) {
            xp2 = Ex[c + 2 * sz];
            yp2 = Ey[c + 2 * sz];
        }
        const bool az1 = k >= 1 && k < iz - 1, az2 = k >= 1 && k < iz - 2;
        const bool upx = ax1 && ay2 && az2, upy = ax2 && ay1 && az2, upz = ax2 && ay2 && az1;
        const bool gx = gxc && k < iz - 1, gy = gyc && k < iz - 1;
        const bool gz = ((k >= 1 && k < nc - 1) || (k >= iz - nc && k < iz - 2)) && i < ix - 1 &&
                        jj < iy - 1;
        if (upx || upy || upz || gx || gy || gz) {
            double ezy = 0.0, eyz = 0.0, exz = 0.0, ezx = 0.0, eyx = 0.0, exy = 0.0;
            if (upx || gy) ezy = dplus(Ez, c, sy, c1, c2);
            if (upx || gz) eyz = c1 * yp1 - c1 * y0 + c2 * yp2 - c2 * ym1;
            if (upy || gz) exz = c1 * xp1 - c1 * x0 + c2 * xp2 - c2 * xm1;
            if (upy || gx) ezx = dplus(Ez, c, 1, c1, c2);
            if (upz || gx) eyx = c1 * Ey[c + 1] - c1 * y0 + c2 * Ey[c + 2] - c2 * Ey[c - 1];
            if (upz || gy) exy = c1 * Ex[c + sy] - c1 * x0 + c2 * Ex[c + 2 * sy] - c2 * Ex[c - sy];

            const double msz = pz[PR_MSIG * iz + k];
            double smx = 0.0, smy = 0.0, smz = 0.0, dbx = dbi, dby = dbi, dbz = dbi;
            if (msx != 0.0 || msy != 0.0 || msz != 0.0) {
                const double eps2 = gc(S.eps)[c];
                smx = msx * dt / (2.0 * eps2);
                smy = msy * dt / (2.0 * eps2);
                smz = msz * dt / (2.0 * eps2);
                dbx = dt / S.mu0 / (1.0 + smx);
                dby = dt / S.mu0 / (1.0 + smy);
                dbz = dt / S.mu0 / (1.0 + smz);
            }
            const double khdz = uload(S.prof[2] + PR_KHD * iz + k);
            double hx = Hx[c], hy = Hy[c], hz = Hz[c];
            if (upx) {
                const double da = msx != 0.0 ? (1.0 - smx) / (1.0 + smx) : 1.0;
                const double cu = ezy / khdy - eyz / khdz;
                hx = sign > 0 ? da * hx - dbx * cu : da * hx + dbx * cu;
            }
            if (upy) {
                const double da = msy != 0.0 ? (1.0 - smy) / (1.0 + smy) : 1.0;
                const double cu = exz / khdz - ezx / khdx;
                hy = sign > 0 ? da * hy - dby * cu : da * hy + dby * cu;
            }
            if (upz) {
                const double da = msz != 0.0 ? (1.0 - smz) / (1.0 + smz) : 1.0;
                const double cu = eyx / khdx - exy / khdy;
                hz = sign > 0 ? da * hz - dbz * cu : da * hz + dbz * cu;
            }
            if (gx) {
                GLOBAL double *__restrict__ P = gm(S.psi[0]);
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
                GLOBAL double *__restrict__ P = gm(S.psi[1]);
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
                GLOBAL double *__restrict__ P = gm(S.psi[2]);
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
        xm1 = x0;
        x0 = xp1;
        xp1 = xp2;
        ym1 = y0;
        y0 = yp1;
        yp1 = yp2;
    }
}

hipError_t launch_efield_z(const ShotDev *d, int ns, const int *pre, int nblocks, int j, int sign,
                           hipStream_t st)
{
    if (nblocks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_efield_z, dim3(nblocks), dim3(ZX * ZY), 0, st, d, ns, pre, j, sign);
    return hipGetLastError();
}

hipError_t launch_hfield_z(const ShotDev *d, int ns, const int *pre, int nblocks, int j, int sign,
                           hipStream_t st)
{
    if (nblocks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_hfield_z, dim3(nblocks), dim3(ZX * ZY), 0, st, d, ns, pre, j, sign);
    return hipGetLastError();
}


// ------------------------------------------------------------------------
// Per-cell kernels of the fused step (32 x 8 tiles, one plane per block,
// blocks numbered so each XCD owns a contiguous run of tiles).  Waves whose
// cells all lie outside every CPML shell take a branch-free path with all 25
// loads issued up front; there the reference's coefficients reduce exactly to
// ca = 1, cb = dt/eps' (E) and da = 1, db = dt/mu0 (H), so no division by the
// coefficient denominators is needed and the result is bit-identical.  Any
// other wave runs the general per-cell code.

__device__ __forceinline__ bool fcell(const ShotDev &S, int b, int &i, int &jj, int &k)
{
    if (!S.shell) {
        // every tile of the grid, each XCD owning a contiguous run of planes
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
    // the shell tile list in launch order: at 288^3 (16-cell CPML) this runs
    // each half-step's shell launch 12 us faster than giving each XCD a
    // contiguous run of the list, while on grids that the per-cell family
    // covers whole (84^3) the contiguous runs are 1.2x faster
    if (b >= S.fb)
        return false;
    const GLOBAL int *t = (const GLOBAL int *)(S.shell + b);   // {x0, y0, k, yend}
    const int x0 = t[0], y0 = t[1], yend = t[3];
    k = t[2];
    if (yend) {   // 16 x 16 tile of an x-edge segment, rows below yend
        i = x0 + (threadIdx.x & 15);
        jj = y0 + (threadIdx.x >> 4);
        if (jj >= yend)
            return false;
    } else {
        i = x0 + (threadIdx.x & (ZX - 1));
        jj = y0 + threadIdx.x / ZX;
    }
    if (k < 0 || i >= S.ix || jj >= S.iy)
        return false;
    // cells of the LDS kernels' region are theirs
    return !(i >= S.lo[0] && i < S.hi[0] && jj >= S.lo[1] && jj < S.hi[1] && k >= S.lo[2] &&
             k < S.hi[2]);
}

// The first recn blocks (recn: 0, or ns rounded up to a multiple of 8 so that
// the tile blocks keep their XCD) gather the receivers of the previous local
// step j - 1, block s for shot s: its E slot is the one this step reads and H
// is not written until the H half-step, so the samples are the ones a receiver
// launch after step j - 1 takes, with one graph node less per step.  Leading,
// they overlap the tile blocks instead of trailing the launch.  hf: the step
// is a fused E+H launch sequence (H before step j is hsel(S, j, 1)).
template <bool NT>
__global__ __launch_bounds__(ZX *ZY) void k_efield_f(const ShotDev *__restrict__ D, int ns,
                                                     const int *__restrict__ pre, int j, int sign,
                                                     int recn, int hf)
{
    if ((int)blockIdx.x < recn) {
        if ((int)blockIdx.x < ns) {
            const ShotDev &R = D[blockIdx.x];
            const int n = chunk_step(R) + j - 1;
            if (n < R.it - 1)
                rec_store(R, n, (R.cur + j) % R.ring, hsel(R, j, hf));
        }
        return;
    }
    const int b = blockIdx.x - recn;
    const int s = find_shot(pre, ns, b);
    const ShotDev &S = D[s];
    const int n = chunk_step(S) + j;
    if (n >= S.it - 1)
        return;
    int i, jj, k;
    if (!fcell(S, b - pre[s], i, jj, k))
        return;
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc;
    const int si = (S.cur + j) % S.ring, so = (S.cur + j + 1) % S.ring;
    const bool inner = i >= nc && i < ix - nc && jj >= nc && jj < iy - nc && k >= nc && k < iz - nc;
    const double *Hc = hsel(S, j, hf);
    if (!__all(inner)) {
        efield_cell<M_UPD | M_PSI | M_SRC | (NT ? M_NT : 0)>(S, Hc, i, jj, k, n, si, so, sign);
        return;
    }
    const size_t N = (size_t)S.n;
    const int sy = ix, sz = ix * iy;
    const int c = (k * iy + jj) * ix + i;
    const GLOBAL double *__restrict__ Ein = gc(S.e + (size_t)si * S.sstride);
    GLOBAL double *__restrict__ Eout = gm(S.e + (size_t)so * S.sstride);
    const GLOBAL double *__restrict__ Hx = gc(Hc);
    const GLOBAL double *__restrict__ Hy = gc(Hc + N);
    const GLOBAL double *__restrict__ Hz = gc(Hc + 2 * N);
    const double ex0 = ld_s<NT>(Ein + c), ey0 = ld_s<NT>(Ein + N + c), ez0 = ld_s<NT>(Ein + 2 * N + c);
    const double cb = ld_s<NT>(gc(S.cbi) + c);
    const double z0 = Hz[c], zym1 = Hz[c - sy], zyp1 = Hz[c + sy], zym2 = Hz[c - 2 * sy];
    const double zxm1 = Hz[c - 1], zxp1 = Hz[c + 1], zxm2 = Hz[c - 2];
    const double y0 = Hy[c], yzm1 = Hy[c - sz], yzp1 = Hy[c + sz], yzm2 = Hy[c - 2 * sz];
    const double yxm1 = Hy[c - 1], yxp1 = Hy[c + 1], yxm2 = Hy[c - 2];
    const double x0 = Hx[c], xzm1 = Hx[c - sz], xzp1 = Hx[c + sz], xzm2 = Hx[c - 2 * sz];
    const double xym1 = Hx[c - sy], xyp1 = Hx[c + sy], xym2 = Hx[c - 2 * sy];
    const double kedx = gc(S.prof[0])[PR_KED * ix + i];
    const double kedy = gc(S.prof[1])[PR_KED * iy + jj];
    const double kedz = gc(S.prof[2])[PR_KED * iz + k];
    const double c1 = S.c1, c2 = S.c2;
    const double hzy = c1 * z0 - c1 * zym1 + c2 * zyp1 - c2 * zym2;
    const double hyz = c1 * y0 - c1 * yzm1 + c2 * yzp1 - c2 * yzm2;
    const double hxz = c1 * x0 - c1 * xzm1 + c2 * xzp1 - c2 * xzm2;
    const double hzx = c1 * z0 - c1 * zxm1 + c2 * zxp1 - c2 * zxm2;
    const double hyx = c1 * y0 - c1 * yxm1 + c2 * yxp1 - c2 * yxm2;
    const double hxy = c1 * x0 - c1 * xym1 + c2 * xyp1 - c2 * xym2;
    const double cux = hzy / kedy - hyz / kedz;
    const double cuy = hxz / kedz - hzx / kedx;
    const double cuz = hyx / kedx - hxy / kedy;
    double ex = sign > 0 ? ex0 + cb * cux : ex0 - cb * cux;
    double ey = sign > 0 ? ey0 + cb * cuy : ey0 - cb * cuy;
    double ez = sign > 0 ? ez0 + cb * cuz : ez0 - cb * cuz;
    if (c == S.src_cell) {
        const double a = gc(S.amp)[n];
        if (S.src_comp == 0)
            ex -= a;
        else if (S.src_comp == 1)
            ey -= a;
        else
            ez -= a;
    }
    st_s<NT>(Eout + c, ex);
    st_s<NT>(Eout + N + c, ey);
    st_s<NT>(Eout + 2 * N + c, ez);
}

template <bool NT>
__global__ __launch_bounds__(ZX *ZY) void k_hfield_f(const ShotDev *__restrict__ D, int ns,
                                                     const int *__restrict__ pre, int j, int sign,
                                                     int hf)
{
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    if (chunk_step(S) + j >= S.it - 1)
        return;
    int i, jj, k;
    if (!fcell(S, blockIdx.x - pre[s], i, jj, k))
        return;
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc;
    const int so = (S.cur + j + 1) % S.ring;
    const bool inner = i >= nc && i < ix - nc && jj >= nc && jj < iy - nc && k >= nc && k < iz - nc;
    double *Hi = hsel(S, j, hf), *Ho = hsel(S, j + 1, hf);
    if (!__all(inner)) {
        hfield_cell<M_UPD | M_PSI | (NT ? M_NT : 0)>(S, Hi, Ho, i, jj, k, so, sign);
        return;
    }
    const size_t N = (size_t)S.n;
    const int sy = ix, sz = ix * iy;
    const int c = (k * iy + jj) * ix + i;
    const GLOBAL double *__restrict__ Ex = gc(S.e + (size_t)so * S.sstride);
    const GLOBAL double *__restrict__ Ey = gc(Ex + N);
    const GLOBAL double *__restrict__ Ez = gc(Ex + 2 * N);
    const GLOBAL double *__restrict__ Hxi = gc(Hi);
    const GLOBAL double *__restrict__ Hyi = gc(Hi + N);
    const GLOBAL double *__restrict__ Hzi = gc(Hi + 2 * N);
    GLOBAL double *__restrict__ Hx = gm(Ho);
    GLOBAL double *__restrict__ Hy = gm(Ho + N);
    GLOBAL double *__restrict__ Hz = gm(Ho + 2 * N);
    const double hx0 = ld_s<NT>(Hxi + c), hy0 = ld_s<NT>(Hyi + c), hz0 = ld_s<NT>(Hzi + c);
    const double z0 = Ez[c], zyp1 = Ez[c + sy], zyp2 = Ez[c + 2 * sy], zym1 = Ez[c - sy];
    const double zxp1 = Ez[c + 1], zxp2 = Ez[c + 2], zxm1 = Ez[c - 1];
    const double y0 = Ey[c], yzp1 = Ey[c + sz], yzp2 = Ey[c + 2 * sz], yzm1 = Ey[c - sz];
    const double yxp1 = Ey[c + 1], yxp2 = Ey[c + 2], yxm1 = Ey[c - 1];
    const double x0 = Ex[c], xzp1 = Ex[c + sz], xzp2 = Ex[c + 2 * sz], xzm1 = Ex[c - sz];
    const double xyp1 = Ex[c + sy], xyp2 = Ex[c + 2 * sy], xym1 = Ex[c - sy];
    const double khdx = gc(S.prof[0])[PR_KHD * ix + i];
    const double khdy = gc(S.prof[1])[PR_KHD * iy + jj];
    const double khdz = gc(S.prof[2])[PR_KHD * iz + k];
    const double c1 = S.c1, c2 = S.c2;
    const double db = S.dt / S.mu0 / 1.0;
    const double ezy = c1 * zyp1 - c1 * z0 + c2 * zyp2 - c2 * zym1;
    const double eyz = c1 * yzp1 - c1 * y0 + c2 * yzp2 - c2 * yzm1;
    const double exz = c1 * xzp1 - c1 * x0 + c2 * xzp2 - c2 * xzm1;
    const double ezx = c1 * zxp1 - c1 * z0 + c2 * zxp2 - c2 * zxm1;
    const double eyx = c1 * yxp1 - c1 * y0 + c2 * yxp2 - c2 * yxm1;
    const double exy = c1 * xyp1 - c1 * x0 + c2 * xyp2 - c2 * xym1;
    const double cux = ezy / khdy - eyz / khdz;
    const double cuy = exz / khdz - ezx / khdx;
    const double cuz = eyx / khdx - exy / khdy;
    st_s<NT>(Hx + c, sign > 0 ? hx0 - db * cux : hx0 + db * cux);
    st_s<NT>(Hy + c, sign > 0 ? hy0 - db * cuy : hy0 + db * cuy);
    st_s<NT>(Hz + c, sign > 0 ? hz0 - db * cuz : hz0 + db * cuz);
}

hipError_t launch_efield_f(const ShotDev *d, int ns, const int *pre, int nblocks, int j, int sign,
                           hipStream_t st, bool rec_prev, int hf, bool nt)
{
    if (nblocks <= 0)
        return rec_prev ? hipErrorInvalidValue : hipSuccess;
    const int recn = rec_prev && j > 0 ? (ns + 7) & ~7 : 0;
    if (nt)
        hipLaunchKernelGGL(k_efield_f<true>, dim3(recn + nblocks), dim3(ZX * ZY), 0, st, d, ns, pre, j,
                           sign, recn, hf);
    else
        hipLaunchKernelGGL(k_efield_f<false>, dim3(recn + nblocks), dim3(ZX * ZY), 0, st, d, ns, pre, j,
                           sign, recn, hf);
    return hipGetLastError();
}

hipError_t launch_hfield_f(const ShotDev *d, int ns, const int *pre, int nblocks, int j, int sign,
                           hipStream_t st, int hf, bool nt)
{
    if (nblocks <= 0)
        return hipSuccess;
    if (nt)
        hipLaunchKernelGGL(k_hfield_f<true>, dim3(nblocks), dim3(ZX * ZY), 0, st, d, ns, pre, j, sign, hf);
    else
        hipLaunchKernelGGL(k_hfield_f<false>, dim3(nblocks), dim3(ZX * ZY), 0, st, d, ns, pre, j, sign, hf);
    return hipGetLastError();
}


// ------------------------------------------------------------------------
// LDS z-marching kernels of the CPML-free interior.  A block owns a 64 x LY
// column of the shot's LDS region and marches a z-chunk: the x and y
// stencil neighbours of a plane come from LDS tiles that the block fills
// once per plane, the z neighbours from a register queue along each lane's
// own column, so every plane is read from HBM once per column (plus halo)
// instead of once per stencil point.  The tiles of the next plane are
// fetched into registers while the current plane is computed; one barrier per
// plane (the LDS ring holds the planes a slow lane may still read).  The
// arithmetic is the per-cell kernels' fast path, term for term.

namespace {

constexpr int LNT = LX * LY;

// tile of LX x LY cells with a halo of XL/XR cells before/after along x and
// YL/YR along y
template <int XL, int XR, int YL, int YR> struct LTile {
    static constexpr int W = LX + XL + XR, H = LY + YL + YR, N = W * H;
    static constexpr int R = (N + LNT - 1) / LNT;   // elements per lane
};

template <class T, int XL, int YL>
__device__ __forceinline__ void tile_load(const GLOBAL double *f, size_t pl, int sx, int x0,
                                          int y0, int xmax, int ymax, double (&r)[T::R])
{
#pragma unroll
    for (int q = 0; q < T::R; ++q) {
        const int e = threadIdx.x + q * LNT;
        double v = 0.0;
        if (e < T::N) {
            const int gy = y0 - YL + e / T::W, gx = x0 - XL + e % T::W;
            if (gx <= xmax && gy <= ymax)
                v = f[pl + (size_t)gy * sx + gx];
        }
        r[q] = v;
    }
}

template <class T>
__device__ __forceinline__ void tile_store(double *lds, const double (&r)[T::R])
{
#pragma unroll
    for (int q = 0; q < T::R; ++q) {
        const int e = threadIdx.x + q * LNT;
        if (e < T::N)
            lds[e] = r[q];
    }
}

// block -> (shot region tile, z-chunk); false for padding blocks
__device__ __forceinline__ bool lcol(const ShotDev &S, int b, int &x0, int &y0, int &k0, int &k1)
{
    const int L = (b & 7) * (S.lb_pad >> 3) + (b >> 3);
    if (L >= S.lb)
        return false;
    const int tx = L % S.ltx, r = L / S.ltx;
    const int ty = r % S.lty, kz = r / S.lty;
    x0 = S.lo[0] + tx * LX;
    y0 = S.lo[1] + ty * LY;
    k0 = S.lo[2] + kz * S.lkb;
    k1 = min(k0 + S.lkb, S.hi[2]);
    return true;
}

}  // namespace

template <bool NT>
__global__ __launch_bounds__(LNT) void k_efield_l(const ShotDev *__restrict__ D, int ns,
                                                  const int *__restrict__ pre, int j, int sign)
{
    // Hz: x -2..+1, y -2..+1; Hy: x -2..+1; Hx: y -2..+1
    using TZ = LTile<2, 1, 2, 1>;
    using TY = LTile<2, 1, 0, 0>;
    using TX = LTile<0, 0, 2, 1>;
    __shared__ double sz[3][TZ::N], sy[3][TY::N], sx[3][TX::N];
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    const int n = chunk_step(S) + j;
    if (n >= S.it - 1)
        return;
    int x0, y0, k0, k1;
    if (!lcol(S, blockIdx.x - pre[s], x0, y0, k0, k1))
        return;
    const int lx = threadIdx.x % LX, ly = threadIdx.x / LX;
    const int i = x0 + lx, jj = y0 + ly;
    const bool valid = i < S.hi[0] && jj < S.hi[1];
    // idle lanes of a partial tile read their row's last cell (stays in bounds)
    const int ci = min(i, S.hi[0] - 1), cj = min(jj, S.hi[1] - 1);
    const int ix = S.ix, iy = S.iy;
    const size_t N = (size_t)S.n, pl = (size_t)ix * iy;
    const int xmax = min(x0 + LX, S.hi[0]), ymax = min(y0 + LY, S.hi[1]);
    const int si = (S.cur + j) % S.ring, so = (S.cur + j + 1) % S.ring;
    const GLOBAL double *Ein = gc(S.e + (size_t)si * S.sstride);
    GLOBAL double *Eout = gm(S.e + (size_t)so * S.sstride);
    const double *Hc = hsel(S, j, 0);
    const GLOBAL double *Hx = gc(Hc), *Hy = gc(Hc + N), *Hz = gc(Hc + 2 * N);
    const GLOBAL double *CB = gc(S.cbi);
    const double kedx = gc(S.prof[0])[PR_KED * ix + ci];
    const double kedy = gc(S.prof[1])[PR_KED * iy + cj];
    const double c1 = S.c1, c2 = S.c2;
    const size_t col = (size_t)cj * ix + ci;
    double yq2 = Hy[(size_t)(k0 - 2) * pl + col], yq1 = Hy[(size_t)(k0 - 1) * pl + col];
    double xq2 = Hx[(size_t)(k0 - 2) * pl + col], xq1 = Hx[(size_t)(k0 - 1) * pl + col];
    double rz[TZ::R], ry[TY::R], rx[TX::R];
    int b0 = 0, b1 = 1, b2 = 2;   // LDS buffers of planes k, k+1 and the one being filled
    tile_load<TZ, 2, 2>(Hz, (size_t)k0 * pl, ix, x0, y0, xmax, ymax, rz);
    tile_load<TY, 2, 0>(Hy, (size_t)k0 * pl, ix, x0, y0, xmax, ymax, ry);
    tile_load<TX, 0, 2>(Hx, (size_t)k0 * pl, ix, x0, y0, xmax, ymax, rx);
    tile_store<TZ>(sz[b0], rz);
    tile_store<TY>(sy[b0], ry);
    tile_store<TX>(sx[b0], rx);
    tile_load<TZ, 2, 2>(Hz, (size_t)(k0 + 1) * pl, ix, x0, y0, xmax, ymax, rz);
    tile_load<TY, 2, 0>(Hy, (size_t)(k0 + 1) * pl, ix, x0, y0, xmax, ymax, ry);
    tile_load<TX, 0, 2>(Hx, (size_t)(k0 + 1) * pl, ix, x0, y0, xmax, ymax, rx);
    size_t c = (size_t)k0 * pl + col;
    double ex0 = ld_s<NT>(Ein + c), ey0 = ld_s<NT>(Ein + N + c), ez0 = ld_s<NT>(Ein + 2 * N + c);
    double cb = ld_s<NT>(CB + c);
    for (int k = k0; k < k1; ++k) {
        tile_store<TZ>(sz[b1], rz);
        tile_store<TY>(sy[b1], ry);
        tile_store<TX>(sx[b1], rx);
        __syncthreads();
        const bool more = k + 1 < k1;
        const size_t cn = c + pl;
        double nex = 0.0, ney = 0.0, nez = 0.0, ncb = 0.0;
        if (more) {
            tile_load<TZ, 2, 2>(Hz, (size_t)(k + 2) * pl, ix, x0, y0, xmax, ymax, rz);
            tile_load<TY, 2, 0>(Hy, (size_t)(k + 2) * pl, ix, x0, y0, xmax, ymax, ry);
            tile_load<TX, 0, 2>(Hx, (size_t)(k + 2) * pl, ix, x0, y0, xmax, ymax, rx);
            nex = ld_s<NT>(Ein + cn);
            ney = ld_s<NT>(Ein + N + cn);
            nez = ld_s<NT>(Ein + 2 * N + cn);
            ncb = ld_s<NT>(CB + cn);
        }
        const double kedz = uload(S.prof[2] + PR_KED * S.iz + k);
        const double *Z = sz[b0], *Y = sy[b0], *X = sx[b0];
        const int zc = (ly + 2) * TZ::W + lx + 2;
        const double z0 = Z[zc], zym1 = Z[zc - TZ::W], zyp1 = Z[zc + TZ::W], zym2 = Z[zc - 2 * TZ::W];
        const double zxm1 = Z[zc - 1], zxp1 = Z[zc + 1], zxm2 = Z[zc - 2];
        const int yc = ly * TY::W + lx + 2;
        const double y0v = Y[yc], yxm1 = Y[yc - 1], yxp1 = Y[yc + 1], yxm2 = Y[yc - 2];
        const double yzp1 = sy[b1][yc];
        const int xc = (ly + 2) * TX::W + lx;
        const double x0v = X[xc], xym1 = X[xc - TX::W], xyp1 = X[xc + TX::W], xym2 = X[xc - 2 * TX::W];
        const double xzp1 = sx[b1][xc];
        const double hzy = c1 * z0 - c1 * zym1 + c2 * zyp1 - c2 * zym2;
        const double hyz = c1 * y0v - c1 * yq1 + c2 * yzp1 - c2 * yq2;
        const double hxz = c1 * x0v - c1 * xq1 + c2 * xzp1 - c2 * xq2;
        const double hzx = c1 * z0 - c1 * zxm1 + c2 * zxp1 - c2 * zxm2;
        const double hyx = c1 * y0v - c1 * yxm1 + c2 * yxp1 - c2 * yxm2;
        const double hxy = c1 * x0v - c1 * xym1 + c2 * xyp1 - c2 * xym2;
        const double cux = hzy / kedy - hyz / kedz;
        const double cuy = hxz / kedz - hzx / kedx;
        const double cuz = hyx / kedx - hxy / kedy;
        if (valid) {
            double ex = sign > 0 ? ex0 + cb * cux : ex0 - cb * cux;
            double ey = sign > 0 ? ey0 + cb * cuy : ey0 - cb * cuy;
            double ez = sign > 0 ? ez0 + cb * cuz : ez0 - cb * cuz;
            if ((long long)c == S.src_cell) {
                const double a = gc(S.amp)[n];
                if (S.src_comp == 0)
                    ex -= a;
                else if (S.src_comp == 1)
                    ey -= a;
                else
                    ez -= a;
            }
            st_s<NT>(Eout + c, ex);
            st_s<NT>(Eout + N + c, ey);
            st_s<NT>(Eout + 2 * N + c, ez);
        }
        yq2 = yq1;
        yq1 = y0v;
        xq2 = xq1;
        xq1 = x0v;
        ex0 = nex;
        ey0 = ney;
        ez0 = nez;
        cb = ncb;
        c = cn;
        const int t = b0;
        b0 = b1;
        b1 = b2;
        b2 = t;
    }
}

template <bool NT>
__global__ __launch_bounds__(LNT) void k_hfield_l(const ShotDev *__restrict__ D, int ns,
                                                  const int *__restrict__ pre, int j, int sign)
{
    // Ez: x -1..+2, y -1..+2; Ey: x -1..+2; Ex: y -1..+2
    using TZ = LTile<1, 2, 1, 2>;
    using TY = LTile<1, 2, 0, 0>;
    using TX = LTile<0, 0, 1, 2>;
    __shared__ double sz[4][TZ::N], sy[4][TY::N], sx[4][TX::N];
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    if (chunk_step(S) + j >= S.it - 1)
        return;
    int x0, y0, k0, k1;
    if (!lcol(S, blockIdx.x - pre[s], x0, y0, k0, k1))
        return;
    const int lx = threadIdx.x % LX, ly = threadIdx.x / LX;
    const int i = x0 + lx, jj = y0 + ly;
    const bool valid = i < S.hi[0] && jj < S.hi[1];
    const int ci = min(i, S.hi[0] - 1), cj = min(jj, S.hi[1] - 1);
    const int ix = S.ix, iy = S.iy;
    const size_t N = (size_t)S.n, pl = (size_t)ix * iy;
    const int xmax = min(x0 + LX, S.hi[0]) + 1, ymax = min(y0 + LY, S.hi[1]) + 1;
    const int so = (S.cur + j + 1) % S.ring;
    const GLOBAL double *Ex = gc(S.e + (size_t)so * S.sstride);
    const GLOBAL double *Ey = Ex + N, *Ez = Ex + 2 * N;
    // H in place, or from ring slot j to slot j + 1 (see hsel)
    const double *Hc = hsel(S, j, 0);
    double *Hn = hsel(S, j + 1, 0);
    const GLOBAL double *Hxi = gc(Hc), *Hyi = gc(Hc + N), *Hzi = gc(Hc + 2 * N);
    GLOBAL double *Hx = gm(Hn), *Hy = gm(Hn + N), *Hz = gm(Hn + 2 * N);
    const double khdx = gc(S.prof[0])[PR_KHD * ix + ci];
    const double khdy = gc(S.prof[1])[PR_KHD * iy + cj];
    const double c1 = S.c1, c2 = S.c2;
    const double db = S.dt / S.mu0 / 1.0;
    const size_t col = (size_t)cj * ix + ci;
    double yq1 = Ey[(size_t)(k0 - 1) * pl + col], xq1 = Ex[(size_t)(k0 - 1) * pl + col];
    double rz[TZ::R], ry[TY::R], rx[TX::R];
    int b0 = 0, b1 = 1, b2 = 2, b3 = 3;   // planes k, k+1, k+2 (being filled), free
    for (int p = 0; p < 2; ++p) {
        const int bp = p == 0 ? b0 : b1;
        tile_load<TZ, 1, 1>(Ez, (size_t)(k0 + p) * pl, ix, x0, y0, xmax, ymax, rz);
        tile_load<TY, 1, 0>(Ey, (size_t)(k0 + p) * pl, ix, x0, y0, xmax, ymax, ry);
        tile_load<TX, 0, 1>(Ex, (size_t)(k0 + p) * pl, ix, x0, y0, xmax, ymax, rx);
        tile_store<TZ>(sz[bp], rz);
        tile_store<TY>(sy[bp], ry);
        tile_store<TX>(sx[bp], rx);
    }
    tile_load<TZ, 1, 1>(Ez, (size_t)(k0 + 2) * pl, ix, x0, y0, xmax, ymax, rz);
    tile_load<TY, 1, 0>(Ey, (size_t)(k0 + 2) * pl, ix, x0, y0, xmax, ymax, ry);
    tile_load<TX, 0, 1>(Ex, (size_t)(k0 + 2) * pl, ix, x0, y0, xmax, ymax, rx);
    size_t c = (size_t)k0 * pl + col;
    double hx0 = ld_s<NT>(Hxi + c), hy0 = ld_s<NT>(Hyi + c), hz0 = ld_s<NT>(Hzi + c);
    for (int k = k0; k < k1; ++k) {
        tile_store<TZ>(sz[b2], rz);
        tile_store<TY>(sy[b2], ry);
        tile_store<TX>(sx[b2], rx);
        __syncthreads();
        const bool more = k + 1 < k1;
        const size_t cn = c + pl;
        double nhx = 0.0, nhy = 0.0, nhz = 0.0;
        if (more) {
            tile_load<TZ, 1, 1>(Ez, (size_t)(k + 3) * pl, ix, x0, y0, xmax, ymax, rz);
            tile_load<TY, 1, 0>(Ey, (size_t)(k + 3) * pl, ix, x0, y0, xmax, ymax, ry);
            tile_load<TX, 0, 1>(Ex, (size_t)(k + 3) * pl, ix, x0, y0, xmax, ymax, rx);
            nhx = ld_s<NT>(Hxi + cn);
            nhy = ld_s<NT>(Hyi + cn);
            nhz = ld_s<NT>(Hzi + cn);
        }
        const double khdz = uload(S.prof[2] + PR_KHD * S.iz + k);
        const double *Z = sz[b0], *Y = sy[b0], *X = sx[b0];
        const int zc = (ly + 1) * TZ::W + lx + 1;
        const double z0 = Z[zc], zyp1 = Z[zc + TZ::W], zyp2 = Z[zc + 2 * TZ::W], zym1 = Z[zc - TZ::W];
        const double zxp1 = Z[zc + 1], zxp2 = Z[zc + 2], zxm1 = Z[zc - 1];
        const int yc = ly * TY::W + lx + 1;
        const double y0v = Y[yc], yxp1 = Y[yc + 1], yxp2 = Y[yc + 2], yxm1 = Y[yc - 1];
        const double yzp1 = sy[b1][yc], yzp2 = sy[b2][yc];
        const int xc = (ly + 1) * TX::W + lx;
        const double x0v = X[xc], xyp1 = X[xc + TX::W], xyp2 = X[xc + 2 * TX::W], xym1 = X[xc - TX::W];
        const double xzp1 = sx[b1][xc], xzp2 = sx[b2][xc];
        const double ezy = c1 * zyp1 - c1 * z0 + c2 * zyp2 - c2 * zym1;
        const double eyz = c1 * yzp1 - c1 * y0v + c2 * yzp2 - c2 * yq1;
        const double exz = c1 * xzp1 - c1 * x0v + c2 * xzp2 - c2 * xq1;
        const double ezx = c1 * zxp1 - c1 * z0 + c2 * zxp2 - c2 * zxm1;
        const double eyx = c1 * yxp1 - c1 * y0v + c2 * yxp2 - c2 * yxm1;
        const double exy = c1 * xyp1 - c1 * x0v + c2 * xyp2 - c2 * xym1;
        const double cux = ezy / khdy - eyz / khdz;
        const double cuy = exz / khdz - ezx / khdx;
        const double cuz = eyx / khdx - exy / khdy;
        if (valid) {
            st_s<NT>(Hx + c, sign > 0 ? hx0 - db * cux : hx0 + db * cux);
            st_s<NT>(Hy + c, sign > 0 ? hy0 - db * cuy : hy0 + db * cuy);
            st_s<NT>(Hz + c, sign > 0 ? hz0 - db * cuz : hz0 + db * cuz);
        }
        yq1 = y0v;
        xq1 = x0v;
        hx0 = nhx;
        hy0 = nhy;
        hz0 = nhz;
        c = cn;
        const int t = b0;
        b0 = b1;
        b1 = b2;
        b2 = b3;
        b3 = t;
    }
}

hipError_t launch_efield_l(const ShotDev *d, int ns, const int *pre, int nblocks, int j, int sign,
                           hipStream_t st, bool nt)
{
    if (nblocks <= 0)
        return hipSuccess;
    if (nt)
        hipLaunchKernelGGL(k_efield_l<true>, dim3(nblocks), dim3(LNT), 0, st, d, ns, pre, j, sign);
    else
        hipLaunchKernelGGL(k_efield_l<false>, dim3(nblocks), dim3(LNT), 0, st, d, ns, pre, j, sign);
    return hipGetLastError();
}

hipError_t launch_hfield_l(const ShotDev *d, int ns, const int *pre, int nblocks, int j, int sign,
                           hipStream_t st, bool nt)
{
    if (nblocks <= 0)
        return hipSuccess;
    if (nt)
        hipLaunchKernelGGL(k_hfield_l<true>, dim3(nblocks), dim3(LNT), 0, st, d, ns, pre, j, sign);
    else
        hipLaunchKernelGGL(k_hfield_l<false>, dim3(nblocks), dim3(LNT), 0, st, d, ns, pre, j, sign);
    return hipGetLastError();
}


// ------------------------------------------------------------------------
// Fused E+H z-march over the LDS region: one launch per step in place of
// k_efield_l + k_hfield_l.  Plane p's E is computed, then H of plane p - 2,
// the plane whose forward z stencil E(p) completes.  The H update of a tile
// needs new E one cell before and two cells after it along x and y, in cells
// that neighbouring blocks own: the block recomputes E there from old H (halo
// items), or reads it from the E slot where the cell lies outside the region
// (the shell launch before this one has written it).  The same holds for the
// planes k0 - 1, k1 and k1 + 1 around a z-chunk.  Old H is read from
// hsel(S, j, 1) and new H written to hsel(S, j + 1, 1), and E goes from ring
// slot si to slot so, so no block overwrites a value another block still
// reads.  Per step this moves E, H and dt/eps' in and E and H out (104 B per
// cell) instead of the 152 B of the two launches.
//
// Work items: lane t owns the tile cell (t % LX, t / LX) and, for t < FB_N,
// one halo cell: first the rows -1, LY, LY + 1 of the tile's columns (their
// Ez and Ex are needed), then the columns -1, LX, LX + 1 of its rows (Ez and
// Ey).  The arithmetic is the per-cell fast path term for term, so every value
// is bitwise the two-launch schedule's.  On the region kappa = 1, so ked and
// khd are one value per axis (the host checks this before it selects the
// kernel).  LDS: old H of plane p, and new E of planes p, p - 1 and p - 2; two
// barriers per plane.

namespace {

constexpr int FB_N = 3 * LX + 3 * LY;
// old-H tiles of a plane, as offsets (u, v) from the tile origin
constexpr int FZW = LX + 6, FZH = LY + 6;   // Hz: u in [-3, LX + 3), v in [-3, LY + 3)
constexpr int FYW = LX + 6, FYH = LY + 3;   // Hy: u in [-3, LX + 3), v in [-1, LY + 2)
constexpr int FXW = LX + 3, FXH = LY + 6;   // Hx: u in [-1, LX + 2), v in [-3, LY + 3)
constexpr int FZN = FZW * FZH, FYN = FYW * FYH, FXN = FXW * FXH;
constexpr int FHN = FZN + FYN + FXN;
constexpr int FHR = (FHN + LNT - 1) / LNT;
// new-E tiles of a plane
constexpr int GZW = LX + 3, GZH = LY + 3;   // Ez: u in [-1, LX + 2), v in [-1, LY + 2)
constexpr int GYW = LX + 3, GYH = LY;       // Ey: u in [-1, LX + 2), v in [0, LY)
constexpr int GXW = LX, GXH = LY + 3;       // Ex: u in [0, LX), v in [-1, LY + 2)
constexpr int GZN = GZW * GZH, GYN = GYW * GYH, GXN = GXW * GXH;
constexpr int GEN = GZN + GYN + GXN;
static_assert(FB_N <= LNT, "halo items exceed the block");

__device__ __forceinline__ int fz_at(int u, int v) { return (v + 3) * FZW + u + 3; }
__device__ __forceinline__ int fy_at(int u, int v) { return FZN + (v + 1) * FYW + u + 3; }
__device__ __forceinline__ int fx_at(int u, int v) { return FZN + FYN + (v + 3) * FXW + u + 1; }
__device__ __forceinline__ int gz_at(int u, int v) { return (v + 1) * GZW + u + 1; }
__device__ __forceinline__ int gy_at(int u, int v) { return GZN + v * GYW + u + 1; }
__device__ __forceinline__ int gx_at(int u, int v) { return GZN + GYN + (v + 1) * GXW + u; }

}  // namespace

__global__ __launch_bounds__(LNT, 4) void k_ehfield_l(const ShotDev *__restrict__ D, int ns,
                                                      const int *__restrict__ pre, int j)
{
    __shared__ double sh[FHN], se[3][GEN];
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    const int n = chunk_step(S) + j;
    if (n >= S.it - 1)
        return;
    int x0, y0, k0, k1;
    if (!lcol(S, blockIdx.x - pre[s], x0, y0, k0, k1))
        return;
    const int ix = S.ix, iy = S.iy, iz = S.iz;
    const size_t N = (size_t)S.n, pl = (size_t)ix * iy;
    const int lo2 = S.lo[2], hi2 = S.hi[2];
    const int si = (S.cur + j) % S.ring, so = (S.cur + j + 1) % S.ring;
    const GLOBAL double *Ein = gc(S.e + (size_t)si * S.sstride);
    GLOBAL double *Eout = gm(S.e + (size_t)so * S.sstride);
    const GLOBAL double *Hin = gc(hsel(S, j, 1));
    GLOBAL double *Hout = gm(hsel(S, j + 1, 1));
    const GLOBAL double *CB = gc(S.cbi);
    const double c1 = S.c1, c2 = S.c2;
    const double db = S.dt / S.mu0 / 1.0;
    const double kedx = gc(S.prof[0])[PR_KED * ix + S.lo[0]];
    const double kedy = gc(S.prof[1])[PR_KED * iy + S.lo[1]];
    const double kedz = gc(S.prof[2])[PR_KED * iz + lo2];
    const double khdx = gc(S.prof[0])[PR_KHD * ix + S.lo[0]];
    const double khdy = gc(S.prof[1])[PR_KHD * iy + S.lo[1]];
    const double khdz = gc(S.prof[2])[PR_KHD * iz + lo2];
    const int t = threadIdx.x;

    // old-H tile elements this lane loads: offset within a plane (component
    // included; the host keeps 3 N below 2^31), -1 outside the grid
    int ho[FHR];
#pragma unroll
    for (int q = 0; q < FHR; ++q) {
        int e = t + q * LNT, comp, u, v;
        ho[q] = -1;
        if (e >= FHN)
            continue;
        if (e < FZN) {
            comp = 2, u = e % FZW - 3, v = e / FZW - 3;
        } else if (e < FZN + FYN) {
            e -= FZN;
            comp = 1, u = e % FYW - 3, v = e / FYW - 1;
        } else {
            e -= FZN + FYN;
            comp = 0, u = e % FXW - 1, v = e / FXW - 3;
        }
        const int gx = x0 + u, gy = y0 + v;
        if (gx >= 0 && gx < ix && gy >= 0 && gy < iy)
            ho[q] = comp * S.n + gy * ix + gx;
    }

    // tile item (A): computes and owns E and H where it lies in the region
    const int uA = t % LX, vA = t / LX;
    const bool ingA = x0 + uA < ix && y0 + vA < iy;
    const bool own = x0 + uA < S.hi[0] && y0 + vA < S.hi[1];
    const int colA = ingA ? (y0 + vA) * ix + x0 + uA : 0;
    // halo item (B): yh on the halo rows (Ez, Ex), else on the halo columns (Ez, Ey)
    const bool hasB = t < FB_N, yh = t < 3 * LX;
    int uB, vB;
    if (yh) {
        const int r = t / LX;
        uB = t % LX;
        vB = r == 0 ? -1 : LY - 1 + r;
    } else {
        const int m = t - 3 * LX, c3 = m % 3;
        uB = c3 == 0 ? -1 : LX - 1 + c3;
        vB = m / 3;
    }
    const int gxB = x0 + uB, gyB = y0 + vB;
    const bool ingB = hasB && gxB >= 0 && gxB < ix && gyB >= 0 && gyB < iy;
    const bool inrB = ingB && gxB >= S.lo[0] && gxB < S.hi[0] && gyB >= S.lo[1] && gyB < S.hi[1];
    const int colB = ingB ? gyB * ix + gxB : 0;
    const int dB = yh ? 0 : 1;                     // B's second component
    const size_t qoB = yh ? N : 0;                 // B's z queue: Hy on the rows, Hx on the columns
    const int e2B = yh ? gx_at(uB, vB) : gy_at(uB, vB);

    const int pa = k0 - 1, pb = k1 + 1;
    // own-column old H at planes p - 2, p - 1 and p + 1 (plane p comes from LDS)
    double qyA[3], qxA[3], qB[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int pz = pa - 2 + (r < 2 ? r : 3);
        const bool okz = pz >= 0 && pz < iz;
        qyA[r] = ingA && okz ? Hin[N + (size_t)pz * pl + colA] : 0.0;
        qxA[r] = ingA && okz ? Hin[(size_t)pz * pl + colA] : 0.0;
        qB[r] = ingB && okz ? Hin[qoB + (size_t)pz * pl + colB] : 0.0;
    }
    // E(p) inputs: old E and dt/eps' where the item computes, else the new E
    // the shell launch wrote
    double eA[3], eBz, eB2, cbA, cbB;
    double rh[FHR];
    {
        const bool z = pa >= lo2 && pa < hi2;
        const bool cA = own && z, cB = inrB && z;
        const size_t cAi = (size_t)pa * pl + colA, cBi = (size_t)pa * pl + colB;
        const GLOBAL double *sA = cA ? Ein : Eout, *sB = cB ? Ein : Eout;
#pragma unroll
        for (int d = 0; d < 3; ++d)
            eA[d] = ingA ? sA[d * N + cAi] : 0.0;
        cbA = cA ? CB[cAi] : 0.0;
        eBz = ingB ? sB[2 * N + cBi] : 0.0;
        eB2 = ingB ? sB[dB * N + cBi] : 0.0;
        cbB = cB ? CB[cBi] : 0.0;
#pragma unroll
        for (int q = 0; q < FHR; ++q)
            rh[q] = ho[q] >= 0 ? Hin[(size_t)pa * pl + ho[q]] : 0.0;
    }
    double d3x = 0.0, d3y = 0.0;              // new Ex, Ey of plane p - 3 at the tile cell

    for (int p = pa; p <= pb; ++p) {
        const bool zc = p >= lo2 && p < hi2;   // E computed (not read) at plane p
#pragma unroll
        for (int q = 0; q < FHR; ++q) {
            const int e = t + q * LNT;
            if (e < FHN)
                sh[e] = rh[q];
        }
        __syncthreads();   // sh holds plane p; se[p % 3] is free

        double *eb = se[p % 3];
        // E(p) at the tile cell
        const double yA = sh[fy_at(uA, vA)], xA = sh[fx_at(uA, vA)];
        double ex = eA[0], ey = eA[1], ez = eA[2];
        if (zc && own) {
            const double *Z = sh + fz_at(uA, vA);
            const double z0 = Z[0], zym1 = Z[-FZW], zyp1 = Z[FZW], zym2 = Z[-2 * FZW];
            const double zxm1 = Z[-1], zxp1 = Z[1], zxm2 = Z[-2];
            const double *Y = sh + fy_at(uA, vA);
            const double yxm1 = Y[-1], yxp1 = Y[1], yxm2 = Y[-2];
            const double *X = sh + fx_at(uA, vA);
            const double xym1 = X[-FXW], xyp1 = X[FXW], xym2 = X[-2 * FXW];
            const double hzy = c1 * z0 - c1 * zym1 + c2 * zyp1 - c2 * zym2;
            const double hyz = c1 * yA - c1 * qyA[1] + c2 * qyA[2] - c2 * qyA[0];
            const double hxz = c1 * xA - c1 * qxA[1] + c2 * qxA[2] - c2 * qxA[0];
            const double hzx = c1 * z0 - c1 * zxm1 + c2 * zxp1 - c2 * zxm2;
            const double hyx = c1 * yA - c1 * yxm1 + c2 * yxp1 - c2 * yxm2;
            const double hxy = c1 * xA - c1 * xym1 + c2 * xyp1 - c2 * xym2;
            const double cux = hzy / kedy - hyz / kedz;
            const double cuy = hxz / kedz - hzx / kedx;
            const double cuz = hyx / kedx - hxy / kedy;
            ex = eA[0] + cbA * cux;
            ey = eA[1] + cbA * cuy;
            ez = eA[2] + cbA * cuz;
            if ((long long)((size_t)p * pl + colA) == S.src_cell) {
                const double a = gc(S.amp)[n];
                if (S.src_comp == 0)
                    ex -= a;
                else if (S.src_comp == 1)
                    ey -= a;
                else
                    ez -= a;
            }
        }
        eb[gz_at(uA, vA)] = ez;
        eb[gy_at(uA, vA)] = ey;
        eb[gx_at(uA, vA)] = ex;
        if (own && p >= k0 && p < k1) {
            const size_t c = (size_t)p * pl + colA;
            Eout[c] = ex;
            Eout[N + c] = ey;
            Eout[2 * N + c] = ez;
        }
        // E(p) at the halo cell
        double ownB = 0.0;
        if (hasB) {
            ownB = sh[yh ? fy_at(uB, vB) : fx_at(uB, vB)];
            double ezb = eBz, e2b = eB2;
            if (zc && inrB) {
                const double *Y = sh + fy_at(uB, vB);
                const double *X = sh + fx_at(uB, vB);
                const double *Z = sh + fz_at(uB, vB);
                const double hyx = c1 * Y[0] - c1 * Y[-1] + c2 * Y[1] - c2 * Y[-2];
                const double hxy = c1 * X[0] - c1 * X[-FXW] + c2 * X[FXW] - c2 * X[-2 * FXW];
                const double cuz = hyx / kedx - hxy / kedy;
                ezb = eBz + cbB * cuz;
                if (yh) {   // Ex
                    const double hzy = c1 * Z[0] - c1 * Z[-FZW] + c2 * Z[FZW] - c2 * Z[-2 * FZW];
                    const double hyz = c1 * ownB - c1 * qB[1] + c2 * qB[2] - c2 * qB[0];
                    const double cux = hzy / kedy - hyz / kedz;
                    e2b = eB2 + cbB * cux;
                } else {    // Ey
                    const double hxz = c1 * ownB - c1 * qB[1] + c2 * qB[2] - c2 * qB[0];
                    const double hzx = c1 * Z[0] - c1 * Z[-1] + c2 * Z[1] - c2 * Z[-2];
                    const double cuy = hxz / kedz - hzx / kedx;
                    e2b = eB2 + cbB * cuy;
                }
                if ((long long)((size_t)p * pl + colB) == S.src_cell) {
                    const double a = gc(S.amp)[n];
                    if (S.src_comp == 2)
                        ezb -= a;
                    else if (S.src_comp == dB)
                        e2b -= a;
                }
            }
            eb[gz_at(uB, vB)] = ezb;
            eb[e2B] = e2b;
        }
        __syncthreads();   // se[p % 3] holds plane p; sh is free

        // queues forward one plane; the old-H tile of plane p + 1 is fetched
        // while H of plane p - 2 is computed, the rest after it
        qyA[0] = qyA[1];
        qyA[1] = yA;
        qxA[0] = qxA[1];
        qxA[1] = xA;
        qB[0] = qB[1];
        qB[1] = ownB;
        const bool more = p < pb;
        if (more) {
#pragma unroll
            for (int q = 0; q < FHR; ++q)
                rh[q] = ho[q] >= 0 ? Hin[(size_t)(p + 1) * pl + ho[q]] : 0.0;
        }

        // H of plane q = p - 2 at the tile cell: E(q) from se[q % 3], E(q + 1)
        // from se[(q + 1) % 3], E(q + 2) = this plane's, E(q - 1) from d3
        if (own && p - 2 >= k0) {
            const size_t c = (size_t)(p - 2) * pl + colA;
            const double hA0 = Hin[c], hA1 = Hin[N + c], hA2 = Hin[2 * N + c];
            const double *E0 = se[(p - 2) % 3], *E1 = se[(p - 1) % 3];
            const double *Z = E0 + gz_at(uA, vA);
            const double z0 = Z[0], zyp1 = Z[GZW], zyp2 = Z[2 * GZW], zym1 = Z[-GZW];
            const double zxp1 = Z[1], zxp2 = Z[2], zxm1 = Z[-1];
            const double *Y = E0 + gy_at(uA, vA);
            const double y0v = Y[0], yxp1 = Y[1], yxp2 = Y[2], yxm1 = Y[-1];
            const double *X = E0 + gx_at(uA, vA);
            const double x0v = X[0], xyp1 = X[GXW], xyp2 = X[2 * GXW], xym1 = X[-GXW];
            const double yzp1 = E1[gy_at(uA, vA)], xzp1 = E1[gx_at(uA, vA)];
            const double ezy = c1 * zyp1 - c1 * z0 + c2 * zyp2 - c2 * zym1;
            const double eyz = c1 * yzp1 - c1 * y0v + c2 * ey - c2 * d3y;
            const double exz = c1 * xzp1 - c1 * x0v + c2 * ex - c2 * d3x;
            const double ezx = c1 * zxp1 - c1 * z0 + c2 * zxp2 - c2 * zxm1;
            const double eyx = c1 * yxp1 - c1 * y0v + c2 * yxp2 - c2 * yxm1;
            const double exy = c1 * xyp1 - c1 * x0v + c2 * xyp2 - c2 * xym1;
            const double cux = ezy / khdy - eyz / khdz;
            const double cuy = exz / khdz - ezx / khdx;
            const double cuz = eyx / khdx - exy / khdy;
            Hout[c] = hA0 - db * cux;
            Hout[N + c] = hA1 - db * cuy;
            Hout[2 * N + c] = hA2 - db * cuz;
        }
        if (more) {
            const int pn = p + 1, pz = p + 2;
            const bool z = pn >= lo2 && pn < hi2;
            const bool cA = own && z, cB = inrB && z;
            const size_t cAi = (size_t)pn * pl + colA, cBi = (size_t)pn * pl + colB;
            const GLOBAL double *sA = cA ? Ein : Eout, *sB = cB ? Ein : Eout;
#pragma unroll
            for (int d = 0; d < 3; ++d)
                eA[d] = ingA ? sA[d * N + cAi] : 0.0;
            cbA = cA ? CB[cAi] : 0.0;
            eBz = ingB ? sB[2 * N + cBi] : 0.0;
            eB2 = ingB ? sB[dB * N + cBi] : 0.0;
            cbB = cB ? CB[cBi] : 0.0;
            const bool okz = pz < iz;
            qyA[2] = ingA && okz ? Hin[N + (size_t)pz * pl + colA] : 0.0;
            qxA[2] = ingA && okz ? Hin[(size_t)pz * pl + colA] : 0.0;
            qB[2] = ingB && okz ? Hin[qoB + (size_t)pz * pl + colB] : 0.0;
        }
        if (p - 2 >= pa) {
            d3x = se[(p - 2) % 3][gx_at(uA, vA)];
            d3y = se[(p - 2) % 3][gy_at(uA, vA)];
        }
    }
}

hipError_t launch_ehfield_l(const ShotDev *d, int ns, const int *pre, int nblocks, int j,
                            hipStream_t st)
{
    if (nblocks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_ehfield_l, dim3(nblocks), dim3(LNT), 0, st, d, ns, pre, j);
    return hipGetLastError();
}


// ------------------------------------------------------------------------
// Pair variant of the fused per-cell kernels (even ix): every lane updates the
// x-pair (i, i+1) with 16-byte loads, halving the load instructions of the
// stencil.  Tiles are 32 x 16 cells (16 x 16 lanes); same arithmetic, same
// fast/general split (a wave with any non-interior pair runs the general code
// for both of its cells).

__device__ __forceinline__ double2 ld2(const GLOBAL double *p, size_t q)
{
    const GLOBAL double2 *v = (const GLOBAL double2 *)(p + q);
    double2 r;
    r.x = (*v).x;
    r.y = (*v).y;
    return r;
}

__device__ __forceinline__ void st2(GLOBAL double *p, size_t q, double a, double b)
{
    p[q] = a;
    p[q + 1] = b;
}

__device__ __forceinline__ bool fpair(const ShotDev &S, int b, int &i, int &jj, int &k)
{
    const int L = (b & 7) * (S.fb_pad >> 3) + (b >> 3);
    if (L >= S.fb)
        return false;
    const int tx = L % S.ftx, r = L / S.ftx;
    const int ty = r % S.fty;
    k = r / S.fty;
    i = tx * 32 + 2 * (threadIdx.x & 15);
    jj = ty * 16 + (threadIdx.x >> 4);
    return i < S.ix && jj < S.iy;
}

__global__ __launch_bounds__(256) void k_efield_p(const ShotDev *__restrict__ D, int ns,
                                                  const int *__restrict__ pre, int j, int sign)
{
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    const int n = chunk_step(S) + j;
    if (n >= S.it - 1)
        return;
    int i, jj, k;
    if (!fpair(S, blockIdx.x - pre[s], i, jj, k))
        return;
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc;
    const int si = (S.cur + j) % S.ring, so = (S.cur + j + 1) % S.ring;
    const bool inner =
        i >= nc && i + 1 < ix - nc && jj >= nc && jj < iy - nc && k >= nc && k < iz - nc;
    if (!__all(inner)) {
        efield_cell<M_UPD | M_PSI | M_SRC>(S, hsel(S, 0, 0), i, jj, k, n, si, so, sign);
        efield_cell<M_UPD | M_PSI | M_SRC>(S, hsel(S, 0, 0), i + 1, jj, k, n, si, so, sign);
        return;
    }
    const size_t N = (size_t)S.n;
    const int sy = ix, sz = ix * iy;
    const size_t c = ((size_t)k * iy + jj) * ix + i;
    const GLOBAL double *__restrict__ Ein = gc(S.e + (size_t)si * S.sstride);
    GLOBAL double *__restrict__ Eout = gm(S.e + (size_t)so * S.sstride);
    const double *Hc = hsel(S, 0, 0);
    const GLOBAL double *__restrict__ Hx = gc(Hc);
    const GLOBAL double *__restrict__ Hy = gc(Hc + N);
    const GLOBAL double *__restrict__ Hz = gc(Hc + 2 * N);
    const double2 ex0 = ld2(Ein, c), ey0 = ld2(Ein, N + c), ez0 = ld2(Ein, 2 * N + c);
    const double2 cb = ld2(gc(S.cbi), c);
    const double2 z0 = ld2(Hz, c), zym1 = ld2(Hz, c - sy), zyp1 = ld2(Hz, c + sy),
                  zym2 = ld2(Hz, c - 2 * sy), zxm = ld2(Hz, c - 2), zxp = ld2(Hz, c + 2);
    const double2 y0 = ld2(Hy, c), yzm1 = ld2(Hy, c - sz), yzp1 = ld2(Hy, c + sz),
                  yzm2 = ld2(Hy, c - 2 * sz), yxm = ld2(Hy, c - 2), yxp = ld2(Hy, c + 2);
    const double2 x0 = ld2(Hx, c), xzm1 = ld2(Hx, c - sz), xzp1 = ld2(Hx, c + sz),
                  xzm2 = ld2(Hx, c - 2 * sz), xym1 = ld2(Hx, c - sy), xyp1 = ld2(Hx, c + sy),
                  xym2 = ld2(Hx, c - 2 * sy);
    const double2 kedx = ld2(gc(S.prof[0]), (size_t)PR_KED * ix + i);
    const double kedy = gc(S.prof[1])[PR_KED * iy + jj];
    const double kedz = gc(S.prof[2])[PR_KED * iz + k];
    const double c1 = S.c1, c2 = S.c2;
    // cell a = i
    const double hzy_a = c1 * z0.x - c1 * zym1.x + c2 * zyp1.x - c2 * zym2.x;
    const double hyz_a = c1 * y0.x - c1 * yzm1.x + c2 * yzp1.x - c2 * yzm2.x;
    const double hxz_a = c1 * x0.x - c1 * xzm1.x + c2 * xzp1.x - c2 * xzm2.x;
    const double hzx_a = c1 * z0.x - c1 * zxm.y + c2 * z0.y - c2 * zxm.x;
    const double hyx_a = c1 * y0.x - c1 * yxm.y + c2 * y0.y - c2 * yxm.x;
    const double hxy_a = c1 * x0.x - c1 * xym1.x + c2 * xyp1.x - c2 * xym2.x;
    // cell b = i + 1
    const double hzy_b = c1 * z0.y - c1 * zym1.y + c2 * zyp1.y - c2 * zym2.y;
    const double hyz_b = c1 * y0.y - c1 * yzm1.y + c2 * yzp1.y - c2 * yzm2.y;
    const double hxz_b = c1 * x0.y - c1 * xzm1.y + c2 * xzp1.y - c2 * xzm2.y;
    const double hzx_b = c1 * z0.y - c1 * z0.x + c2 * zxp.x - c2 * zxm.y;
    const double hyx_b = c1 * y0.y - c1 * y0.x + c2 * yxp.x - c2 * yxm.y;
    const double hxy_b = c1 * x0.y - c1 * xym1.y + c2 * xyp1.y - c2 * xym2.y;
    const double cux_a = hzy_a / kedy - hyz_a / kedz, cux_b = hzy_b / kedy - hyz_b / kedz;
    const double cuy_a = hxz_a / kedz - hzx_a / kedx.x, cuy_b = hxz_b / kedz - hzx_b / kedx.y;
    const double cuz_a = hyx_a / kedx.x - hxy_a / kedy, cuz_b = hyx_b / kedx.y - hxy_b / kedy;
    double exa, exb, eya, eyb, eza, ezb;
    if (sign > 0) {
        exa = ex0.x + cb.x * cux_a; exb = ex0.y + cb.y * cux_b;
        eya = ey0.x + cb.x * cuy_a; eyb = ey0.y + cb.y * cuy_b;
        eza = ez0.x + cb.x * cuz_a; ezb = ez0.y + cb.y * cuz_b;
    } else {
        exa = ex0.x - cb.x * cux_a; exb = ex0.y - cb.y * cux_b;
        eya = ey0.x - cb.x * cuy_a; eyb = ey0.y - cb.y * cuy_b;
        eza = ez0.x - cb.x * cuz_a; ezb = ez0.y - cb.y * cuz_b;
    }
    if ((long long)c == S.src_cell || (long long)c + 1 == S.src_cell) {
        const double a = gc(S.amp)[n];
        const bool second = (long long)c + 1 == S.src_cell;
        if (S.src_comp == 0) {
            if (second) exb -= a; else exa -= a;
        } else if (S.src_comp == 1) {
            if (second) eyb -= a; else eya -= a;
        } else {
            if (second) ezb -= a; else eza -= a;
        }
    }
    st2(Eout, c, exa, exb);
    st2(Eout, N + c, eya, eyb);
    st2(Eout, 2 * N + c, eza, ezb);
}

__global__ __launch_bounds__(256) void k_hfield_p(const ShotDev *__restrict__ D, int ns,
                                                  const int *__restrict__ pre, int j, int sign)
{
    const int s = find_shot(pre, ns, blockIdx.x);
    const ShotDev &S = D[s];
    if (chunk_step(S) + j >= S.it - 1)
        return;
    int i, jj, k;
    if (!fpair(S, blockIdx.x - pre[s], i, jj, k))
        return;
    const int ix = S.ix, iy = S.iy, iz = S.iz, nc = S.nc;
    const int so = (S.cur + j + 1) % S.ring;
    const bool inner =
        i >= nc && i + 1 < ix - nc && jj >= nc && jj < iy - nc && k >= nc && k < iz - nc;
    if (!__all(inner)) {
        double *H = hsel(S, 0, 0);
        hfield_cell<M_UPD | M_PSI>(S, H, H, i, jj, k, so, sign);
        hfield_cell<M_UPD | M_PSI>(S, H, H, i + 1, jj, k, so, sign);
        return;
    }
    const size_t N = (size_t)S.n;
    const int sy = ix, sz = ix * iy;
    const size_t c = ((size_t)k * iy + jj) * ix + i;
    const GLOBAL double *__restrict__ Ex = gc(S.e + (size_t)so * S.sstride);
    const GLOBAL double *__restrict__ Ey = Ex + N;
    const GLOBAL double *__restrict__ Ez = Ex + 2 * N;
    double *Hc = hsel(S, 0, 0);
    GLOBAL double *__restrict__ Hx = gm(Hc);
    GLOBAL double *__restrict__ Hy = gm(Hc + N);
    GLOBAL double *__restrict__ Hz = gm(Hc + 2 * N);
    const double2 hx0 = ld2(Hx, c), hy0 = ld2(Hy, c), hz0 = ld2(Hz, c);
    const double2 z0 = ld2(Ez, c), zyp1 = ld2(Ez, c + sy), zyp2 = ld2(Ez, c + 2 * sy),
                  zym1 = ld2(Ez, c - sy), zxm = ld2(Ez, c - 2), zxp = ld2(Ez, c + 2);
    const double2 y0 = ld2(Ey, c), yzp1 = ld2(Ey, c + sz), yzp2 = ld2(Ey, c + 2 * sz),
                  yzm1 = ld2(Ey, c - sz), yxm = ld2(Ey, c - 2), yxp = ld2(Ey, c + 2);
    const double2 x0 = ld2(Ex, c), xzp1 = ld2(Ex, c + sz), xzp2 = ld2(Ex, c + 2 * sz),
                  xzm1 = ld2(Ex, c - sz), xyp1 = ld2(Ex, c + sy), xyp2 = ld2(Ex, c + 2 * sy),
                  xym1 = ld2(Ex, c - sy);
    const double2 khdx = ld2(gc(S.prof[0]), (size_t)PR_KHD * ix + i);
    const double khdy = gc(S.prof[1])[PR_KHD * iy + jj];
    const double khdz = gc(S.prof[2])[PR_KHD * iz + k];
    const double c1 = S.c1, c2 = S.c2;
    const double db = S.dt / S.mu0 / 1.0;
    // cell a = i: forward x stencil uses i-1 .. i+2
    const double ezy_a = c1 * zyp1.x - c1 * z0.x + c2 * zyp2.x - c2 * zym1.x;
    const double eyz_a = c1 * yzp1.x - c1 * y0.x + c2 * yzp2.x - c2 * yzm1.x;
    const double exz_a = c1 * xzp1.x - c1 * x0.x + c2 * xzp2.x - c2 * xzm1.x;
    const double ezx_a = c1 * z0.y - c1 * z0.x + c2 * zxp.x - c2 * zxm.y;
    const double eyx_a = c1 * y0.y - c1 * y0.x + c2 * yxp.x - c2 * yxm.y;
    const double exy_a = c1 * xyp1.x - c1 * x0.x + c2 * xyp2.x - c2 * xym1.x;
    // cell b = i + 1: i .. i+3
    const double ezy_b = c1 * zyp1.y - c1 * z0.y + c2 * zyp2.y - c2 * zym1.y;
    const double eyz_b = c1 * yzp1.y - c1 * y0.y + c2 * yzp2.y - c2 * yzm1.y;
    const double exz_b = c1 * xzp1.y - c1 * x0.y + c2 * xzp2.y - c2 * xzm1.y;
    const double ezx_b = c1 * zxp.x - c1 * z0.y + c2 * zxp.y - c2 * z0.x;
    const double eyx_b = c1 * yxp.x - c1 * y0.y + c2 * yxp.y - c2 * y0.x;
    const double exy_b = c1 * xyp1.y - c1 * x0.y + c2 * xyp2.y - c2 * xym1.y;
    const double cux_a = ezy_a / khdy - eyz_a / khdz, cux_b = ezy_b / khdy - eyz_b / khdz;
    const double cuy_a = exz_a / khdz - ezx_a / khdx.x, cuy_b = exz_b / khdz - ezx_b / khdx.y;
    const double cuz_a = eyx_a / khdx.x - exy_a / khdy, cuz_b = eyx_b / khdx.y - exy_b / khdy;
    if (sign > 0) {
        st2(Hx, c, hx0.x - db * cux_a, hx0.y - db * cux_b);
        st2(Hy, c, hy0.x - db * cuy_a, hy0.y - db * cuy_b);
        st2(Hz, c, hz0.x - db * cuz_a, hz0.y - db * cuz_b);
    } else {
        st2(Hx, c, hx0.x + db * cux_a, hx0.y + db * cux_b);
        st2(Hy, c, hy0.x + db * cuy_a, hy0.y + db * cuy_b);
        st2(Hz, c, hz0.x + db * cuz_a, hz0.y + db * cuz_b);
    }
}

hipError_t launch_efield_p(const ShotDev *d, int ns, const int *pre, int nblocks, int j, int sign,
                           hipStream_t st)
{
    if (nblocks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_efield_p, dim3(nblocks), dim3(256), 0, st, d, ns, pre, j, sign);
    return hipGetLastError();
}

hipError_t launch_hfield_p(const ShotDev *d, int ns, const int *pre, int nblocks, int j, int sign,
                           hipStream_t st)
{
    if (nblocks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_hfield_p, dim3(nblocks), dim3(256), 0, st, d, ns, pre, j, sign);
    return hipGetLastError();
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
                            hipStream_t st, int hf)
{
    if (ns <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_receivers, dim3(ns), dim3(256), 0, st, d, j, nover, inplace, hf);
    return hipGetLastError();
}

hipError_t launch_flush(const ShotDev *d, int ns, const int *pre_cells, int nblocks, int L,
                        hipStream_t st, bool h)
{
    if (nblocks <= 0 || L <= 0)
        return hipSuccess;
    if (L > RING_MAX)
        return hipErrorInvalidValue;
    if (h)
        hipLaunchKernelGGL(k_flush<1>, dim3(nblocks), dim3(256), 0, st, d, ns, pre_cells, L);
    else
        hipLaunchKernelGGL(k_flush<0>, dim3(nblocks), dim3(256), 0, st, d, ns, pre_cells, L);
    return hipGetLastError();
}

hipError_t launch_advance(const ShotDev *d, int ns, int L, hipStream_t st, int hf)
{
    hipLaunchKernelGGL(k_advance, dim3((ns + 63) / 64), dim3(64), 0, st, d, ns, L, hf);
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

hipError_t launch_copy(const double *src, double *dst, size_t count, hipStream_t st)
{
    if (count == 0)
        return hipSuccess;
    size_t blocks = (count + 255) / 256;
    if (blocks > 262144)
        blocks = 262144;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, st, src, dst, count);
    return hipGetLastError();
}

// Copies the shot's live state to a stored state (save) or back.  Segment k
// is stored right after the segments before it and owns the blocks
// [first[k], first[k + 1]); each block copies one chunk of 2 * kStateChunk
// doubles of its segment, every lane issuing its kStateUnroll loads before any
// store, as 16-byte units when both sides of the segment start 16-byte
// aligned, else as doubles.  Blocks are sized per segment, so a short segment
// (a psi slab) launches no idle blocks.
constexpr int kStateLanes = 256, kStateUnroll = 4;
constexpr size_t kStateChunk = (size_t)kStateLanes * kStateUnroll;

__global__ __launch_bounds__(kStateLanes) void k_state_copy(StateSegs segs, double *__restrict__ stored,
                                                            int save)
{
    const unsigned b = blockIdx.x;
    int k = 0;
    while (k + 1 < segs.nseg && b >= segs.first[k + 1])
        ++k;
    size_t base = 0;
    for (int q = 0; q < k; ++q)
        base += segs.count[q];
    double *live = segs.live[k], *st = stored + base;
    const double *__restrict__ src = save ? live : st;
    double *__restrict__ dst = save ? st : live;
    const size_t n = segs.count[k];
    const size_t blk = b - segs.first[k];
    if ((((uintptr_t)live | (uintptr_t)st) & 15) == 0) {
        const double2 *__restrict__ s2 = reinterpret_cast<const double2 *>(src);
        double2 *__restrict__ d2 = reinterpret_cast<double2 *>(dst);
        const size_t n2 = n / 2, u0 = blk * kStateChunk + threadIdx.x;
        double2 v[kStateUnroll];
#pragma unroll
        for (int q = 0; q < kStateUnroll; ++q) {
            const size_t u = u0 + (size_t)q * kStateLanes;
            if (u < n2)
                v[q] = s2[u];
        }
#pragma unroll
        for (int q = 0; q < kStateUnroll; ++q) {
            const size_t u = u0 + (size_t)q * kStateLanes;
            if (u < n2)
                d2[u] = v[q];
        }
        if ((n & 1) && blk == 0 && threadIdx.x == 0)
            dst[n - 1] = src[n - 1];
    } else {
        const size_t e0 = blk * 2 * kStateChunk + threadIdx.x;
        double v[2 * kStateUnroll];
#pragma unroll
        for (int q = 0; q < 2 * kStateUnroll; ++q) {
            const size_t e = e0 + (size_t)q * kStateLanes;
            if (e < n)
                v[q] = src[e];
        }
#pragma unroll
        for (int q = 0; q < 2 * kStateUnroll; ++q) {
            const size_t e = e0 + (size_t)q * kStateLanes;
            if (e < n)
                dst[e] = v[q];
        }
    }
}

hipError_t launch_state_copy(const StateSegs &segs_in, double *stored, bool save, hipStream_t st)
{
    StateSegs segs = segs_in;
    size_t nb = 0;
    for (int k = 0; k < segs.nseg; ++k) {
        segs.first[k] = (unsigned)nb;
        nb += (segs.count[k] + 2 * kStateChunk - 1) / (2 * kStateChunk);
    }
    segs.first[segs.nseg] = (unsigned)nb;
    if (nb == 0)
        return hipSuccess;
    if (nb > 0xffffffffu)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_state_copy, dim3((unsigned)nb), dim3(kStateLanes), 0, st, segs, stored,
                       save ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_diag_copy(const double *src, double *dst, size_t count, hipStream_t st)
{
    const size_t n2 = count / 2;
    size_t blocks = (n2 + 255) / 256;
    if (blocks > 262144)
        blocks = 262144;
    hipLaunchKernelGGL(k_diag_copy, dim3((unsigned)blocks), dim3(256), 0, st,
                       reinterpret_cast<const double2 *>(src), reinterpret_cast<double2 *>(dst), n2);
    return hipGetLastError();
}

hipError_t launch_eps(const double *sig, double *eps, double *cbi, int n, double omega0, double dt,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_eps, dim3((n + 255) / 256), dim3(256), 0, st, sig, eps, cbi, n, omega0, dt);
    return hipGetLastError();
}

}  // namespace mt
