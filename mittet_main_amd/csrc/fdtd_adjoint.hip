// fdtd_adjoint.hip -- exact discrete adjoint of the fused forward step and the
// gradient of a receiver misfit with respect to eps' (hence sigma).
//
// Forward step n (state u = (E, H, psiE, psiH), all linear for fixed sigma):
//   E-phase : E' = ca.E + cb.curl-(H)  on the reference's update ranges,
//             psiE' = be.psiE + ce.D-(H)/dl on the CPML faces, E += +-cb.psiE'
//   source  : E'[src] -= amp[n],  amp ~ 1/sigma[src]
//   H-phase : H' = da.H - db.curl+(E'), psiH' = bh.psiH + ch.D+(E')/dl, H += +-db.psiH'
// (satan_i/fwi3d_cpu_function_v1.4.c:1184-1290, fwi3d_cpu_cpmlabc_v1.4.c:402-679).
// The adjoint applies the transposed sub-steps in reverse order.  Transposes
// are written as gathers: the transpose of D+ is -D- and of D- is -D+, taken
// over the same range masks and with the weights of the cell the forward
// stencil was centred on.  Gradient terms use
//   d ca/de = 2K/(e+K)^2, d cb/de = -dt/(e+K)^2          (K = esig*dt/2)
//   d da/de = 2Km/(e+Km)^2, d db/de = (dt/mu0) Km/(e+Km)^2 (Km = msig*dt/2)
// with e = eps' = sigma/(2 omega0): ca = (e-K)/(e+K), cb = dt/(e+K) are the
// reference's coefficients rewritten.
#include "fdtd_dev.h"

#define GLOBAL __attribute__((address_space(1)))

namespace mt {

namespace {

__device__ __forceinline__ int slabq(int a, const AdjArgs &A, int i, int j, int k)
{
    const int nc = A.nc, ix = A.ix, iy = A.iy, iz = A.iz;
    if (a == 0) {
        const int s = i < nc ? i : i - (ix - 2 * nc);
        return (k * iy + j) * (2 * nc) + s;
    }
    if (a == 1) {
        const int s = j < nc ? j : j - (iy - 2 * nc);
        return (k * (2 * nc) + s) * ix + i;
    }
    const int s = k < nc ? k : k - (iz - 2 * nc);
    return (s * iy + j) * ix + i;
}

// reference update ranges (0-based, half open)
__device__ __forceinline__ bool in_e(const AdjArgs &A, int d, int i, int j, int k)
{
    const int idx[3] = {i, j, k}, n[3] = {A.ix, A.iy, A.iz};
    for (int q = 0; q < 3; ++q) {
        const int lo = q == d ? 1 : 2;
        if (idx[q] < lo || idx[q] >= n[q] - 1)
            return false;
    }
    return true;
}

__device__ __forceinline__ bool in_h(const AdjArgs &A, int d, int i, int j, int k)
{
    const int idx[3] = {i, j, k}, n[3] = {A.ix, A.iy, A.iz};
    for (int q = 0; q < 3; ++q) {
        const int hi = q == d ? n[q] - 1 : n[q] - 2;
        if (idx[q] < 1 || idx[q] >= hi)
            return false;
    }
    return true;
}

// CPML faces: E on x [2,nc-1)U[n-nc+1,n-2), y,z [2,nc)U[n-nc+1,n-2); H on
// [1,nc-1)U[n-nc,n-2); the two other axes [0,n-1).
__device__ __forceinline__ bool face_e(const AdjArgs &A, int a, int i, int j, int k)
{
    const int idx[3] = {i, j, k}, n[3] = {A.ix, A.iy, A.iz};
    const int r = idx[a], nn = n[a], nc = A.nc;
    const bool in = (r >= 2 && r < (a == 0 ? nc - 1 : nc)) || (r >= nn - nc + 1 && r < nn - 2);
    if (!in)
        return false;
    for (int q = 0; q < 3; ++q)
        if (q != a && idx[q] >= n[q] - 1)
            return false;
    return true;
}

__device__ __forceinline__ bool face_h(const AdjArgs &A, int a, int i, int j, int k)
{
    const int idx[3] = {i, j, k}, n[3] = {A.ix, A.iy, A.iz};
    const int r = idx[a], nn = n[a], nc = A.nc;
    const bool in = (r >= 1 && r < nc - 1) || (r >= nn - nc && r < nn - 2);
    if (!in)
        return false;
    for (int q = 0; q < 3; ++q)
        if (q != a && idx[q] >= n[q] - 1)
            return false;
    return true;
}

__device__ __forceinline__ double prof(const AdjArgs &A, int a, int row, int idx)
{
    const int n = a == 0 ? A.ix : (a == 1 ? A.iy : A.iz);
    return ((const GLOBAL double *)A.prof[a])[row * n + idx];
}

// 1/ked (h = false) or 1/khd (h = true) along axis a at index idx
__device__ __forceinline__ double rprof(const AdjArgs &A, int a, bool h, int idx)
{
    const int n = a == 0 ? A.ix : (a == 1 ? A.iy : A.iz);
    return ((const GLOBAL double *)A.rp[a])[(h ? n : 0) + idx];
}

// 32 x 8 tiles of one plane per block, numbered so that each XCD owns a
// contiguous run of tiles (the numbering of the fused forward kernels)
__device__ __forceinline__ bool cell_of(const AdjArgs &A, int &i, int &j, int &k, int &c)
{
    const int b = blockIdx.x;
    const int L = (b & 7) * (A.fb_pad >> 3) + (b >> 3);
    if (L >= A.fb)
        return false;
    const int tx = L % A.ftx, r = L / A.ftx;
    const int ty = r % A.fty;
    k = r / A.fty;
    i = tx * ZX + (threadIdx.x & (ZX - 1));
    j = ty * ZY + threadIdx.x / ZX;
    if (i >= A.ix || j >= A.iy)
        return false;
    c = (k * A.iy + j) * A.ix + i;
    return true;
}

// Cell-local CPML-free interior [nc, n - nc): ca = da = 1, cb = dt/eps',
// db = dt/mu0, ked = khd = dl, no psi face; only the curl terms and the
// cb-derivative of the E update remain.
__device__ __forceinline__ bool inner(const AdjArgs &A, int i, int j, int k)
{
    return i >= A.nc && i < A.inner_hi[0] && j >= A.nc && j < A.inner_hi[1] && k >= A.nc &&
           k < A.inner_hi[2];
}

// the transposed curl stencils (offsets -2..+2) stay inside the grid
__device__ __forceinline__ bool core(const AdjArgs &A, int i, int j, int k)
{
    return i >= 2 && i < A.ix - 2 && j >= 2 && j < A.iy - 2 && k >= 2 && k < A.iz - 2;
}

// along axis a, the transposed psi stencils of this cell (offsets -2..+1 for
// psi-H, -1..+2 for psi-E) reach a CPML face cell
__device__ __forceinline__ bool near_face(const AdjArgs &A, int a, int r)
{
    const int n = a == 0 ? A.ix : (a == 1 ? A.iy : A.iz);
    return r < A.nc + 1 || r >= n - A.nc - 1;
}

__device__ __forceinline__ double dminus_at(const GLOBAL double *w, int c, int s, double c1, double c2)
{
    return c1 * w[c] - c1 * w[c - s] + c2 * w[c + s] - c2 * w[c - 2 * s];
}

__device__ __forceinline__ double dplus_at(const GLOBAL double *w, int c, int s, double c1, double c2)
{
    return c1 * w[c + s] - c1 * w[c] + c2 * w[c + 2 * s] - c2 * w[c - s];
}

}  // namespace

// adjoint receiver source: lE[comp][rec] += g[comp][rec] (state after step n)
__global__ void k_adj_inject(AdjArgs A, const double *__restrict__ g, int nrec,
                             const int *__restrict__ rec)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3 * nrec)
        return;
    const int comp = q / nrec, r = q - comp * nrec;
    A.lE[(size_t)comp * A.n + rec[r]] += g[q];
}

// H-phase part 1 (cell-local): weights for the transposed curls, psiH adjoint,
// H-phase gradient, lH <- da.lH.
__global__ __launch_bounds__(ZX *ZY) void k_adj_h1(AdjArgs A)
{
    int i, j, k, c;
    if (!cell_of(A, i, j, k, c))
        return;
    const int N = A.n, sy = A.ix, sz = A.ix * A.iy;
    const double dt = A.dt, c1 = A.c1, c2 = A.c2;
    const double dbi = dt / A.mu0;
    if (__all(inner(A, i, j, k))) {
        for (int d = 0; d < 3; ++d)
            A.t3[(size_t)d * N + c] = dbi * A.lH[(size_t)d * N + c];
        return;
    }
    const int idx[3] = {i, j, k};
    const long long st[3] = {1, sy, sz};
    const double e = A.eps[c];
    // one reciprocal per axis: r = 1/(e + Km), Km = msig dt/2; then
    // db = dt/mu0 e r, da = (e - Km) r, dda/de = 2 Km r^2, ddb/de = dt/mu0 Km r^2
    double lh[3], ms[3], km[3], rr[3], db[3];
    for (int d = 0; d < 3; ++d) {
        lh[d] = A.lH[(size_t)d * N + c];
        ms[d] = prof(A, d, PR_MSIG, idx[d]);
        km[d] = ms[d] * dt / 2.0;
        rr[d] = 1.0 / (e + km[d]);
        db[d] = dbi * e * rr[d];
    }
    double g = 0.0;
    const GLOBAL double *E1 = (const GLOBAL double *)A.uE1;
    const GLOBAL double *H0 = (const GLOBAL double *)A.uH0;
    for (int d = 0; d < 3; ++d) {
        const int a = (d + 1) % 3, b = (d + 2) % 3;
        const bool up = in_h(A, d, i, j, k);
        A.t3[(size_t)d * N + c] = up ? db[d] * lh[d] : 0.0;
        if (up && ms[d] != 0.0) {
            // gradient: lH_d (dda/de H_d^n - ddb/de curl+_d(E^{n+1}))
            const GLOBAL double *Eb = E1 + (size_t)b * N, *Ea = E1 + (size_t)a * N;
            const double r2 = rr[d] * rr[d];
            const double da_de = 2.0 * km[d] * r2;
            const double db_de = dbi * km[d] * r2;
            const double dpa = c1 * Eb[c + st[a]] - c1 * Eb[c] + c2 * Eb[c + 2 * st[a]] - c2 * Eb[c - st[a]];
            const double dpb = c1 * Ea[c + st[b]] - c1 * Ea[c] + c2 * Ea[c + 2 * st[b]] - c2 * Ea[c - st[b]];
            const double cu = dpa * rprof(A, a, true, idx[a]) - dpb * rprof(A, b, true, idx[b]);
            g += lh[d] * (da_de * H0[(size_t)d * N + c] - db_de * cu);
        }
    }
    for (int a = 0; a < 3; ++a) {
        if (!face_h(A, a, i, j, k))
            continue;
        const int p = (a + 1) % 3, q = (a + 2) % 3;
        const size_t L = (size_t)A.slab[a], s = slabq(a, A, i, j, k);
        double *P = A.lpsi[a];
        double *T = A.tpsi[a];
        const double bh = prof(A, a, PR_BH, idx[a]), ch = prof(A, a, PR_CH, idx[a]);
        // psiH0 feeds H_p (+db_p), psiH1 feeds H_q (-db_q)
        const double m0 = P[2 * L + s] + db[p] * lh[p];
        const double m1 = P[3 * L + s] - db[q] * lh[q];
        P[2 * L + s] = bh * m0;
        P[3 * L + s] = bh * m1;
        T[2 * L + s] = ch * m0 * A.rdl[a];
        T[3 * L + s] = ch * m1 * A.rdl[a];
        // gradient through db in H_p += db_p psi0', H_q -= db_q psi1'
        const GLOBAL double *Q1 = (const GLOBAL double *)A.upsi1[a];
        const double dbp = dbi * km[p] * rr[p] * rr[p];
        const double dbq = dbi * km[q] * rr[q] * rr[q];
        g += lh[p] * dbp * Q1[2 * L + s] - lh[q] * dbq * Q1[3 * L + s];
    }
    for (int d = 0; d < 3; ++d) {
        if (in_h(A, d, i, j, k) && ms[d] != 0.0)
            A.lH[(size_t)d * N + c] = lh[d] * (e - km[d]) * rr[d];
    }
    if (g != 0.0)
        A.grad[c] += g;
}

// weighted value at a neighbour for a transposed stencil (0 outside the grid)
__device__ __forceinline__ double wnb(const AdjArgs &A, const double *T, int comp, int ni, int nj,
                                      int nk, int axis, bool h)
{
    if (ni < 0 || nj < 0 || nk < 0 || ni >= A.ix || nj >= A.iy || nk >= A.iz)
        return 0.0;
    const int c = (nk * A.iy + nj) * A.ix + ni;
    const int id[3] = {ni, nj, nk};
    return T[(size_t)comp * A.n + c] * rprof(A, axis, h, id[axis]);
}

__device__ __forceinline__ double wpsi(const AdjArgs &A, int a, int arr, int ni, int nj, int nk,
                                       bool h)
{
    if (ni < 0 || nj < 0 || nk < 0 || ni >= A.ix || nj >= A.iy || nk >= A.iz)
        return 0.0;
    if (h ? !face_h(A, a, ni, nj, nk) : !face_e(A, a, ni, nj, nk))
        return 0.0;
    return A.tpsi[a][(size_t)arr * A.slab[a] + slabq(a, A, ni, nj, nk)];
}

// H-phase part 2 (gathers): lE += transposed curl+ of the weights and the
// transposed CPML-H stencils; then the source gradient.
__global__ __launch_bounds__(ZX *ZY) void k_adj_h2(AdjArgs A)
{
    int i, j, k, c;
    if (!cell_of(A, i, j, k, c))
        return;
    const double c1 = A.c1, c2 = A.c2;
    const int N = A.n;
    double acc[3] = {0.0, 0.0, 0.0};
    // D-_a w (c) = c1 w(c) - c1 w(c-s) + c2 w(c+s) - c2 w(c-2s)
    auto dminus_psi = [&](int a, int arr) {
        int o[3] = {0, 0, 0};
        auto at = [&](int off) {
            o[a] = off;
            return wpsi(A, a, arr, i + o[0], j + o[1], k + o[2], true);
        };
        return c1 * at(0) - c1 * at(-1) + c2 * at(1) - c2 * at(-2);
    };
    if (A.kappa_one && __all(core(A, i, j, k))) {
        const int st[3] = {1, A.ix, A.ix * A.iy};
        const GLOBAL double *T = (const GLOBAL double *)A.t3;
        for (int d = 0; d < 3; ++d) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            const GLOBAL double *Td = T + (size_t)d * N;
            acc[b] += dminus_at(Td, c, st[a], c1, c2) * A.rdl[a];
            acc[a] -= dminus_at(Td, c, st[b], c1, c2) * A.rdl[b];
        }
    } else {
        auto dminus = [&](int comp, int axis) {
            int o[3] = {0, 0, 0};
            auto at = [&](int off) {
                o[axis] = off;
                return wnb(A, A.t3, comp, i + o[0], j + o[1], k + o[2], axis, true);
            };
            return c1 * at(0) - c1 * at(-1) + c2 * at(1) - c2 * at(-2);
        };
        for (int d = 0; d < 3; ++d) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            // H_d -= db (D+_a E_b/khd_a - D+_b E_a/khd_b):  lE_b += D-_a(k/khd_a), lE_a -= D-_b(k/khd_b)
            acc[b] += dminus(d, a);
            acc[a] -= dminus(d, b);
        }
    }
    const int idx[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
        if (!__any(near_face(A, a, idx[a])))
            continue;
        const int p = (a + 1) % 3, q = (a + 2) % 3;
        // psiH0 <- D+_a E_q, psiH1 <- D+_a E_p; transpose of D+ is -D-
        acc[q] -= dminus_psi(a, 2);
        acc[p] -= dminus_psi(a, 3);
    }
    for (int d = 0; d < 3; ++d)
        A.lE[(size_t)d * N + c] += acc[d];
    if (c == A.src_cell) {
        // E[src] -= amp (amp ~ 1/sigma): d/dsigma = +amp/sigma  (kept in sigma units)
        A.grad_src[0] += A.lE[(size_t)A.src_comp * N + c] * A.amp_n / A.sig_src;
    }
}

// E-phase part 1 (cell-local): weights, psiE adjoint, E-phase gradient, lE <- ca.lE.
__global__ __launch_bounds__(ZX *ZY) void k_adj_e1(AdjArgs A)
{
    int i, j, k, c;
    if (!cell_of(A, i, j, k, c))
        return;
    const int N = A.n, sy = A.ix, sz = A.ix * A.iy;
    const long long st[3] = {1, sy, sz};
    const double e = A.eps[c], dt = A.dt, c1 = A.c1, c2 = A.c2;
    const GLOBAL double *H0 = (const GLOBAL double *)A.uH0;
    if (__all(inner(A, i, j, k))) {
        const double r = 1.0 / e, cb = dt * r, cbd = -dt * r * r;
        double g = 0.0;
        for (int d = 0; d < 3; ++d) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            const double le = A.lE[(size_t)d * N + c];
            A.t3[(size_t)d * N + c] = cb * le;
            const double cu = dminus_at(H0 + (size_t)b * N, c, (int)st[a], c1, c2) * A.rdl[a] -
                              dminus_at(H0 + (size_t)a * N, c, (int)st[b], c1, c2) * A.rdl[b];
            g += le * cbd * cu;
        }
        A.grad[c] += g;
        return;
    }
    const int idx[3] = {i, j, k};
    // one reciprocal per axis: r = 1/(e + K), K = esig dt/2; then cb = dt r,
    // ca = (e - K) r, dca/de = 2 K r^2, dcb/de = -dt r^2
    double le[3], ke[3], rr[3], cb[3], cbd[3];
    for (int d = 0; d < 3; ++d) {
        le[d] = A.lE[(size_t)d * N + c];
        ke[d] = prof(A, d, PR_ESIG, idx[d]) * dt / 2.0;
        rr[d] = 1.0 / (e + ke[d]);
        cb[d] = dt * rr[d];
        cbd[d] = -dt * rr[d] * rr[d];
    }
    double g = 0.0;
    const GLOBAL double *E0 = (const GLOBAL double *)A.uE0;
    for (int d = 0; d < 3; ++d) {
        const int a = (d + 1) % 3, b = (d + 2) % 3;
        const bool up = in_e(A, d, i, j, k);
        A.t3[(size_t)d * N + c] = up ? cb[d] * le[d] : 0.0;
        if (up) {
            const GLOBAL double *Hb = H0 + (size_t)b * N, *Ha = H0 + (size_t)a * N;
            const double dma = c1 * Hb[c] - c1 * Hb[c - st[a]] + c2 * Hb[c + st[a]] - c2 * Hb[c - 2 * st[a]];
            const double dmb = c1 * Ha[c] - c1 * Ha[c - st[b]] + c2 * Ha[c + st[b]] - c2 * Ha[c - 2 * st[b]];
            const double cu = dma * rprof(A, a, false, idx[a]) - dmb * rprof(A, b, false, idx[b]);
            double gd = cbd[d] * cu;
            if (ke[d] != 0.0)
                gd += 2.0 * ke[d] * rr[d] * rr[d] * E0[(size_t)d * N + c];
            g += le[d] * gd;
        }
    }
    for (int a = 0; a < 3; ++a) {
        if (!face_e(A, a, i, j, k))
            continue;
        const int p = (a + 1) % 3, q = (a + 2) % 3;
        const size_t L = (size_t)A.slab[a], s = slabq(a, A, i, j, k);
        double *P = A.lpsi[a];
        double *T = A.tpsi[a];
        const double be = prof(A, a, PR_BE, idx[a]), ce = prof(A, a, PR_CE, idx[a]);
        // psiE0 feeds E_q (+cb_q), psiE1 feeds E_p (-cb_p)
        const double m0 = P[s] + cb[q] * le[q];
        const double m1 = P[L + s] - cb[p] * le[p];
        P[s] = be * m0;
        P[L + s] = be * m1;
        T[s] = ce * m0 * A.rdl[a];
        T[L + s] = ce * m1 * A.rdl[a];
        const GLOBAL double *Q1 = (const GLOBAL double *)A.upsi1[a];
        g += le[q] * cbd[q] * Q1[s] - le[p] * cbd[p] * Q1[L + s];
    }
    for (int d = 0; d < 3; ++d)
        if (in_e(A, d, i, j, k) && ke[d] != 0.0)
            A.lE[(size_t)d * N + c] = le[d] * (e - ke[d]) * rr[d];
    A.grad[c] += g;
}

// E-phase part 2 (gathers): lH += transposed curl- of the weights and the
// transposed CPML-E stencils.
__global__ __launch_bounds__(ZX *ZY) void k_adj_e2(AdjArgs A)
{
    int i, j, k, c;
    if (!cell_of(A, i, j, k, c))
        return;
    const double c1 = A.c1, c2 = A.c2;
    const int N = A.n;
    double acc[3] = {0.0, 0.0, 0.0};
    // D+_a w (c) = c1 w(c+s) - c1 w(c) + c2 w(c+2s) - c2 w(c-s)
    auto dplus_psi = [&](int a, int arr) {
        int o[3] = {0, 0, 0};
        auto at = [&](int off) {
            o[a] = off;
            return wpsi(A, a, arr, i + o[0], j + o[1], k + o[2], false);
        };
        return c1 * at(1) - c1 * at(0) + c2 * at(2) - c2 * at(-1);
    };
    if (A.kappa_one && __all(core(A, i, j, k))) {
        const int st[3] = {1, A.ix, A.ix * A.iy};
        const GLOBAL double *T = (const GLOBAL double *)A.t3;
        for (int d = 0; d < 3; ++d) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            const GLOBAL double *Td = T + (size_t)d * N;
            acc[b] -= dplus_at(Td, c, st[a], c1, c2) * A.rdl[a];
            acc[a] += dplus_at(Td, c, st[b], c1, c2) * A.rdl[b];
        }
    } else {
        auto dplus = [&](int comp, int axis) {
            int o[3] = {0, 0, 0};
            auto at = [&](int off) {
                o[axis] = off;
                return wnb(A, A.t3, comp, i + o[0], j + o[1], k + o[2], axis, false);
            };
            return c1 * at(1) - c1 * at(0) + c2 * at(2) - c2 * at(-1);
        };
        for (int d = 0; d < 3; ++d) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            // E_d += cb (D-_a H_b/ked_a - D-_b H_a/ked_b): lH_b -= D+_a(nu/ked_a), lH_a += D+_b(nu/ked_b)
            acc[b] -= dplus(d, a);
            acc[a] += dplus(d, b);
        }
    }
    const int idx[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
        if (!__any(near_face(A, a, idx[a])))
            continue;
        const int p = (a + 1) % 3, q = (a + 2) % 3;
        // psiE0 <- D-_a H_p, psiE1 <- D-_a H_q; transpose of D- is -D+
        acc[p] -= dplus_psi(a, 0);
        acc[q] -= dplus_psi(a, 1);
    }
    for (int d = 0; d < 3; ++d)
        A.lH[(size_t)d * N + c] += acc[d];
}

hipError_t launch_adjoint_step(const AdjArgs &A, const double *g, int nrec, const int *rec,
                               hipStream_t st)
{
    if (nrec > 0) {
        hipLaunchKernelGGL(k_adj_inject, dim3((3 * nrec + 255) / 256), dim3(256), 0, st, A, g,
                           nrec, rec);
    }
    hipLaunchKernelGGL(k_adj_h1, dim3(A.fb_pad), dim3(ZX * ZY), 0, st, A);
    hipLaunchKernelGGL(k_adj_h2, dim3(A.fb_pad), dim3(ZX * ZY), 0, st, A);
    hipLaunchKernelGGL(k_adj_e1, dim3(A.fb_pad), dim3(ZX * ZY), 0, st, A);
    hipLaunchKernelGGL(k_adj_e2, dim3(A.fb_pad), dim3(ZX * ZY), 0, st, A);
    return hipGetLastError();
}

}  // namespace mt
