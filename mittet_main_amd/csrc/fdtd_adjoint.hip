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

__device__ __forceinline__ bool in_region(const AdjArgs &A, int i, int j, int k)
{
    return i >= A.lo[0] && i < A.hi[0] && j >= A.lo[1] && j < A.hi[1] && k >= A.lo[2] &&
           k < A.hi[2];
}

// Cell of this lane in the per-cell passes: 32 x 8 tiles of one plane per
// block, or the tiles {x0, y0, k, yend} of the cells outside the LDS region
// (16 x 16 tiles below yend when yend != 0); blocks numbered so that each XCD
// owns a contiguous run of tiles (the numbering of the fused forward kernels).
// Cells of the LDS region belong to the LDS kernels.
__device__ __forceinline__ bool cell_of(const AdjArgs &A, int &i, int &j, int &k, int &c)
{
    const int b = blockIdx.x;
    const int L = (b & 7) * (A.fb_pad >> 3) + (b >> 3);
    if (L >= A.fb)
        return false;
    if (A.shell) {
        const GLOBAL int *t = (const GLOBAL int *)(A.shell + L);
        const int x0 = t[0], y0 = t[1], yend = t[3];
        k = t[2];
        if (yend) {
            i = x0 + (threadIdx.x & 15);
            j = y0 + (threadIdx.x >> 4);
            if (j >= yend)
                return false;
        } else {
            i = x0 + (threadIdx.x & (ZX - 1));
            j = y0 + threadIdx.x / ZX;
        }
        if (k < 0)
            return false;
    } else {
        const int tx = L % A.ftx, r = L / A.ftx;
        const int ty = r % A.fty;
        k = r / A.fty;
        i = tx * ZX + (threadIdx.x & (ZX - 1));
        j = ty * ZY + threadIdx.x / ZX;
    }
    if (i >= A.ix || j >= A.iy || in_region(A, i, j, k))
        return false;
    c = (k * A.iy + j) * A.ix + i;
    return true;
}

// Weights of the transposed stencils at a cell (flat index c), component d.
// H phase: db lH with db = dt/mu0 in the LDS region, where the per-cell pass
// leaves t3 unwritten; t3 elsewhere.  E phase: cb lE with cb = dt/eps' there.
// The two bases are read into values before the per-lane choice: choosing
// between the argument struct's fields would choose their addresses and load
// the pointer per lane before the weight.
__device__ __forceinline__ double wgt_h(const AdjArgs &A, int d, int i, int j, int k, int c)
{
    const bool in = in_region(A, i, j, k);
    const GLOBAL double *fin = (const GLOBAL double *)A.lH, *fout = (const GLOBAL double *)A.t3;
    const double v = (in ? fin : fout)[(size_t)d * A.n + c];
    return in ? A.dt / A.mu0 * v : v;
}

__device__ __forceinline__ double wgt_e(const AdjArgs &A, int d, int i, int j, int k, int c)
{
    const bool in = in_region(A, i, j, k);
    const GLOBAL double *fin = (const GLOBAL double *)A.lE, *fout = (const GLOBAL double *)A.t3;
    const double v = (in ? fin : fout)[(size_t)d * A.n + c];
    const double cb = ((const GLOBAL double *)A.cbi)[c];
    return in ? cb * v : v;
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
        double l[3];
        for (int d = 0; d < 3; ++d)
            l[d] = A.lH[(size_t)d * N + c];
        for (int d = 0; d < 3; ++d)
            A.t3[(size_t)d * N + c] = dbi * l[d];
        return;
    }
    const int idx[3] = {i, j, k};
    const long long st[3] = {1, sy, sz};
    // Every store comes after the last load: the vector-memory counter counts
    // loads and stores in issue order, so a wait for a load issued after a
    // store would also wait for that store.
    const double e = A.eps[c];
    const double g0 = A.grad[c];
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
    double g = 0.0, w[3];
    bool up[3];
    const GLOBAL double *E1 = (const GLOBAL double *)A.uE1;
    const GLOBAL double *H0 = (const GLOBAL double *)A.uH0;
    for (int d = 0; d < 3; ++d) {
        const int a = (d + 1) % 3, b = (d + 2) % 3;
        up[d] = in_h(A, d, i, j, k);
        w[d] = up[d] ? db[d] * lh[d] : 0.0;
        if (up[d] && ms[d] != 0.0) {
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
    bool fc[3];
    double m0[3], m1[3], bh[3], ch[3];
    for (int a = 0; a < 3; ++a) {
        fc[a] = face_h(A, a, i, j, k);
        if (!fc[a])
            continue;
        const int p = (a + 1) % 3, q = (a + 2) % 3;
        const size_t L = (size_t)A.slab[a], s = slabq(a, A, i, j, k);
        const GLOBAL double *P = (const GLOBAL double *)A.lpsi[a];
        bh[a] = prof(A, a, PR_BH, idx[a]);
        ch[a] = prof(A, a, PR_CH, idx[a]);
        // psiH0 feeds H_p (+db_p), psiH1 feeds H_q (-db_q)
        m0[a] = P[2 * L + s] + db[p] * lh[p];
        m1[a] = P[3 * L + s] - db[q] * lh[q];
        // gradient through db in H_p += db_p psi0', H_q -= db_q psi1'
        const GLOBAL double *Q1 = (const GLOBAL double *)A.upsi1[a];
        const double dbp = dbi * km[p] * rr[p] * rr[p];
        const double dbq = dbi * km[q] * rr[q] * rr[q];
        g += lh[p] * dbp * Q1[2 * L + s] - lh[q] * dbq * Q1[3 * L + s];
    }
    for (int d = 0; d < 3; ++d)
        A.t3[(size_t)d * N + c] = w[d];
    for (int a = 0; a < 3; ++a) {
        if (!fc[a])
            continue;
        const size_t L = (size_t)A.slab[a], s = slabq(a, A, i, j, k);
        double *P = A.lpsi[a];
        double *T = A.tpsi[a];
        P[2 * L + s] = bh[a] * m0[a];
        P[3 * L + s] = bh[a] * m1[a];
        T[2 * L + s] = ch[a] * m0[a] * A.rdl[a];
        T[3 * L + s] = ch[a] * m1[a] * A.rdl[a];
    }
    for (int d = 0; d < 3; ++d) {
        if (up[d] && ms[d] != 0.0)
            A.lH[(size_t)d * N + c] = lh[d] * (e - km[d]) * rr[d];
    }
    if (g != 0.0)
        A.grad[c] = g0 + g;
}

// weighted value at a neighbour for a transposed stencil (0 outside the grid);
// h: H-phase weights and 1/khd, else E-phase weights and 1/ked.  A neighbour
// outside the grid reads the nearest cell and its value is dropped, so that
// no load waits behind a branch.
__device__ __forceinline__ double wnb(const AdjArgs &A, int comp, int ni, int nj, int nk, int axis,
                                      bool h)
{
    const bool out = ni < 0 || nj < 0 || nk < 0 || ni >= A.ix || nj >= A.iy || nk >= A.iz;
    ni = min(max(ni, 0), A.ix - 1);
    nj = min(max(nj, 0), A.iy - 1);
    nk = min(max(nk, 0), A.iz - 1);
    const int c = (nk * A.iy + nj) * A.ix + ni;
    const int id[3] = {ni, nj, nk};
    const double w = h ? wgt_h(A, comp, ni, nj, nk, c) : wgt_e(A, comp, ni, nj, nk, c);
    const double v = w * rprof(A, axis, h, id[axis]);
    return out ? 0.0 : v;
}

// Both psi weight rows (r0, r1) of axis a at the four neighbours of a cell
// along a (offsets off[0..3]), raw, and on[q]: whether the neighbour is on a
// CPML face (HF: the CPML-H faces [1,nc-1)U[n-nc,n-2), else the CPML-E faces)
// and not on the last plane of another axis; the weight is 0 where it is not
// (the caller drops the raw value, read from a slab row clamped into the
// slab, after every axis' loads are in flight).  The tests of the other two
// axes and the slab-index base are done once per cell.
template <bool HF>
__device__ __forceinline__ void psi_rows(const AdjArgs &A, int a, int i, int j, int k, int r0,
                                         int r1, const int (&off)[4], double (&w0)[4],
                                         double (&w1)[4], bool (&on)[4])
{
    const int nc = A.nc;
    int r, n;
    size_t base, stride;
    bool others;
    if (a == 0) {
        r = i;
        n = A.ix;
        base = ((size_t)k * A.iy + j) * (2 * nc);
        stride = 1;
        others = j < A.iy - 1 && k < A.iz - 1;
    } else if (a == 1) {
        r = j;
        n = A.iy;
        base = (size_t)k * (2 * nc) * A.ix + i;
        stride = A.ix;
        others = i < A.ix - 1 && k < A.iz - 1;
    } else {
        r = k;
        n = A.iz;
        base = (size_t)j * A.ix + i;
        stride = (size_t)A.ix * A.iy;
        others = i < A.ix - 1 && j < A.iy - 1;
    }
    const GLOBAL double *T = (const GLOBAL double *)A.tpsi[a];
    const size_t L = (size_t)A.slab[a];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int rr = r + off[q];
        bool face;
        if (HF)
            face = (rr >= 1 && rr < nc - 1) || (rr >= n - nc && rr < n - 2);
        else
            face = (rr >= 2 && rr < (a == 0 ? nc - 1 : nc)) || (rr >= n - nc + 1 && rr < n - 2);
        const int sp = min(max(rr < nc ? rr : rr - (n - 2 * nc), 0), 2 * nc - 1);
        const size_t qq = base + (size_t)sp * stride;
        w0[q] = T[(size_t)r0 * L + qq];
        w1[q] = T[(size_t)r1 * L + qq];
        on[q] = others && face;
    }
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
    // the cell's own lE first: loaded after the gathers' stores of earlier
    // components, each load would wait for those stores
    double l0[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        l0[d] = A.lE[(size_t)d * N + c];
    double acc[3] = {0.0, 0.0, 0.0};
    if (A.kappa_one && __all(core(A, i, j, k))) {
        const int st[3] = {1, A.ix, A.ix * A.iy};
        // D-_a of the weights of component d (stencil inside the grid)
        auto dm = [&](int d, int a) {
            int o[3] = {0, 0, 0};
            auto at = [&](int off) {
                o[a] = off;
                return wgt_h(A, d, i + o[0], j + o[1], k + o[2], c + off * st[a]);
            };
            return c1 * at(0) - c1 * at(-1) + c2 * at(1) - c2 * at(-2);
        };
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            acc[b] += dm(d, a) * A.rdl[a];
            acc[a] -= dm(d, b) * A.rdl[b];
        }
    } else {
        auto dminus = [&](int comp, int axis) {
            int o[3] = {0, 0, 0};
            auto at = [&](int off) {
                o[axis] = off;
                return wnb(A, comp, i + o[0], j + o[1], k + o[2], axis, true);
            };
            return c1 * at(0) - c1 * at(-1) + c2 * at(1) - c2 * at(-2);
        };
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            // H_d -= db (D+_a E_b/khd_a - D+_b E_a/khd_b):  lE_b += D-_a(k/khd_a), lE_a -= D-_b(k/khd_b)
            acc[b] += dminus(d, a);
            acc[a] -= dminus(d, b);
        }
    }
    const int idx[3] = {i, j, k};
    // the psi rows of every axis near a face are loaded before any is used
    bool act[3], on[3][4];
    double w2[3][4], w3[3][4];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // psiH0 <- D+_a E_q, psiH1 <- D+_a E_p; transpose of D+ is -D-
        // (offsets 0, -1, +1, -2 of D-)
        const int off[4] = {0, -1, 1, -2};
        act[a] = __any(near_face(A, a, idx[a]));
        if (act[a])
            psi_rows<true>(A, a, i, j, k, 2, 3, off, w2[a], w3[a], on[a]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!act[a])
            continue;
        const int p = (a + 1) % 3, q = (a + 2) % 3;
        double u2[4], u3[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            u2[r] = on[a][r] ? w2[a][r] : 0.0;
            u3[r] = on[a][r] ? w3[a][r] : 0.0;
        }
        acc[q] -= c1 * u2[0] - c1 * u2[1] + c2 * u2[2] - c2 * u2[3];
        acc[p] -= c1 * u3[0] - c1 * u3[1] + c2 * u3[2] - c2 * u3[3];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
        A.lE[(size_t)d * N + c] = l0[d] + acc[d];
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
        const double g0 = A.grad[c];
        double g = 0.0, le[3];
        for (int d = 0; d < 3; ++d) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            le[d] = A.lE[(size_t)d * N + c];
            const double cu = dminus_at(H0 + (size_t)b * N, c, (int)st[a], c1, c2) * A.rdl[a] -
                              dminus_at(H0 + (size_t)a * N, c, (int)st[b], c1, c2) * A.rdl[b];
            g += le[d] * cbd * cu;
        }
        for (int d = 0; d < 3; ++d)
            A.t3[(size_t)d * N + c] = cb * le[d];
        A.grad[c] = g0 + g;
        return;
    }
    const int idx[3] = {i, j, k};
    // every store after the last load (as in k_adj_h1)
    const double g0 = A.grad[c];
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
    double g = 0.0, w[3];
    bool up[3];
    const GLOBAL double *E0 = (const GLOBAL double *)A.uE0;
    for (int d = 0; d < 3; ++d) {
        const int a = (d + 1) % 3, b = (d + 2) % 3;
        up[d] = in_e(A, d, i, j, k);
        w[d] = up[d] ? cb[d] * le[d] : 0.0;
        if (up[d]) {
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
    bool fc[3];
    double m0[3], m1[3], be[3], ce[3];
    for (int a = 0; a < 3; ++a) {
        fc[a] = face_e(A, a, i, j, k);
        if (!fc[a])
            continue;
        const int p = (a + 1) % 3, q = (a + 2) % 3;
        const size_t L = (size_t)A.slab[a], s = slabq(a, A, i, j, k);
        const GLOBAL double *P = (const GLOBAL double *)A.lpsi[a];
        be[a] = prof(A, a, PR_BE, idx[a]);
        ce[a] = prof(A, a, PR_CE, idx[a]);
        // psiE0 feeds E_q (+cb_q), psiE1 feeds E_p (-cb_p)
        m0[a] = P[s] + cb[q] * le[q];
        m1[a] = P[L + s] - cb[p] * le[p];
        const GLOBAL double *Q1 = (const GLOBAL double *)A.upsi1[a];
        g += le[q] * cbd[q] * Q1[s] - le[p] * cbd[p] * Q1[L + s];
    }
    for (int d = 0; d < 3; ++d)
        A.t3[(size_t)d * N + c] = w[d];
    for (int a = 0; a < 3; ++a) {
        if (!fc[a])
            continue;
        const size_t L = (size_t)A.slab[a], s = slabq(a, A, i, j, k);
        double *P = A.lpsi[a];
        double *T = A.tpsi[a];
        P[s] = be[a] * m0[a];
        P[L + s] = be[a] * m1[a];
        T[s] = ce[a] * m0[a] * A.rdl[a];
        T[L + s] = ce[a] * m1[a] * A.rdl[a];
    }
    for (int d = 0; d < 3; ++d)
        if (up[d] && ke[d] != 0.0)
            A.lE[(size_t)d * N + c] = le[d] * (e - ke[d]) * rr[d];
    A.grad[c] = g0 + g;
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
    double l0[3];   // the cell's own lH first (as in k_adj_h2)
#pragma unroll
    for (int d = 0; d < 3; ++d)
        l0[d] = A.lH[(size_t)d * N + c];
    double acc[3] = {0.0, 0.0, 0.0};
    if (A.kappa_one && __all(core(A, i, j, k))) {
        const int st[3] = {1, A.ix, A.ix * A.iy};
        // D+_a of the weights of component d (stencil inside the grid)
        auto dp = [&](int d, int a) {
            int o[3] = {0, 0, 0};
            auto at = [&](int off) {
                o[a] = off;
                return wgt_e(A, d, i + o[0], j + o[1], k + o[2], c + off * st[a]);
            };
            return c1 * at(1) - c1 * at(0) + c2 * at(2) - c2 * at(-1);
        };
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            acc[b] -= dp(d, a) * A.rdl[a];
            acc[a] += dp(d, b) * A.rdl[b];
        }
    } else {
        auto dplus = [&](int comp, int axis) {
            int o[3] = {0, 0, 0};
            auto at = [&](int off) {
                o[axis] = off;
                return wnb(A, comp, i + o[0], j + o[1], k + o[2], axis, false);
            };
            return c1 * at(1) - c1 * at(0) + c2 * at(2) - c2 * at(-1);
        };
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int a = (d + 1) % 3, b = (d + 2) % 3;
            // E_d += cb (D-_a H_b/ked_a - D-_b H_a/ked_b): lH_b -= D+_a(nu/ked_a), lH_a += D+_b(nu/ked_b)
            acc[b] -= dplus(d, a);
            acc[a] += dplus(d, b);
        }
    }
    const int idx[3] = {i, j, k};
    // the psi rows of every axis near a face are loaded before any is used
    bool act[3], on[3][4];
    double w0[3][4], w1[3][4];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // psiE0 <- D-_a H_p, psiE1 <- D-_a H_q; transpose of D- is -D+
        // (offsets +1, 0, +2, -1 of D+)
        const int off[4] = {1, 0, 2, -1};
        act[a] = __any(near_face(A, a, idx[a]));
        if (act[a])
            psi_rows<false>(A, a, i, j, k, 0, 1, off, w0[a], w1[a], on[a]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!act[a])
            continue;
        const int p = (a + 1) % 3, q = (a + 2) % 3;
        double u0[4], u1[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            u0[r] = on[a][r] ? w0[a][r] : 0.0;
            u1[r] = on[a][r] ? w1[a][r] : 0.0;
        }
        acc[p] -= c1 * u0[0] - c1 * u0[1] + c2 * u0[2] - c2 * u0[3];
        acc[q] -= c1 * u1[0] - c1 * u1[1] + c2 * u1[2] - c2 * u1[3];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
        A.lH[(size_t)d * N + c] = l0[d] + acc[d];
}

// ------------------------------------------------------------------------
// LDS z-marching kernels of the adjoint over its region, where ca = da = 1,
// cb = dt/eps', db = dt/mu0, ked = khd = dl and no psi stencil reaches.  A
// block owns a 64 x LY column and marches lkb planes along z, as the forward
// LDS kernels.  The H-phase kernel adds the transposed curl+ of the weights
// db lH to lE (the stencil pattern of the forward E kernel); the E-phase
// kernel adds the transposed curl- of the weights cb lE to lH (the pattern
// of the forward H kernel) and accumulates the E-phase gradient
// lE . dcb/de . curl-(H^n) of the cell.  Weights of cells outside the region
// are the per-cell passes' t3, which run first.  Term order is the per-cell
// kernels'.

namespace {

constexpr int ANT = LX * LY;

template <int XL, int XR, int YL, int YR> struct ATile {
    static constexpr int W = LX + XL + XR, H = LY + YL + YR, N = W * H;
    static constexpr int R = (N + ANT - 1) / ANT;   // elements per lane
};

// Weights of component d on plane kk of a tile (halo XL/YL before the tile),
// in two parts so that a plane's loads overlap the arithmetic of the plane
// before it: wload issues the raw loads (the field inside the region, t3
// outside it, and for E weights dt/eps' as well) with no branch and no use of
// their values; wstore_w applies the region's scaling when the tile goes to
// LDS.  Coordinates beyond (xmax, ymax) are clamped onto them: those tile
// cells are read only by the idle lanes of a partial tile.
template <class T, int XL, int YL>
__device__ __forceinline__ int tile_cell(int e, int x0, int y0, int xmax, int ymax, int &gx, int &gy)
{
    gx = min(x0 - XL + e % T::W, xmax);
    gy = min(y0 - YL + e / T::W, ymax);
    return e;
}

template <class T, int XL, int YL, bool HPHASE>
__device__ __forceinline__ void wload(const AdjArgs &A, int d, int kk, int x0, int y0, int xmax,
                                      int ymax, double (&r)[T::R], double (&rc)[T::R])
{
    // the two bases as values: selecting between struct fields would select
    // their addresses and load the pointer per lane first
    const GLOBAL double *fin = (const GLOBAL double *)(HPHASE ? A.lH : A.lE) + (size_t)d * A.n;
    const GLOBAL double *fout = (const GLOBAL double *)A.t3 + (size_t)d * A.n;
    const GLOBAL double *cbi = (const GLOBAL double *)A.cbi;
#pragma unroll
    for (int q = 0; q < T::R; ++q) {
        int gx, gy;
        tile_cell<T, XL, YL>(min((int)threadIdx.x + q * ANT, T::N - 1), x0, y0, xmax, ymax, gx, gy);
        const int c = (kk * A.iy + gy) * A.ix + gx;
        const bool in = in_region(A, gx, gy, kk);
        r[q] = (in ? fin : fout)[c];
        if (!HPHASE)
            rc[q] = cbi[c];
    }
}

template <class T, int XL, int YL, bool HPHASE>
__device__ __forceinline__ void wstore_w(const AdjArgs &A, int kk, int x0, int y0, int xmax, int ymax,
                                         double *lds, const double (&r)[T::R], const double (&rc)[T::R])
{
    const double dbi = A.dt / A.mu0;
#pragma unroll
    for (int q = 0; q < T::R; ++q) {
        const int e = threadIdx.x + q * ANT;
        if (e < T::N) {
            int gx, gy;
            tile_cell<T, XL, YL>(e, x0, y0, xmax, ymax, gx, gy);
            const bool in = in_region(A, gx, gy, kk);
            const double v = r[q];
            lds[e] = in ? (HPHASE ? dbi * v : rc[q] * v) : v;
        }
    }
}

// values of array f on plane kk of a tile, coordinates clamped as in wload
template <class T, int XL, int YL>
__device__ __forceinline__ void fload(const GLOBAL double *f, int kk, int ix, int pl, int x0,
                                      int y0, int xmax, int ymax, double (&r)[T::R])
{
#pragma unroll
    for (int q = 0; q < T::R; ++q) {
        int gx, gy;
        tile_cell<T, XL, YL>(min((int)threadIdx.x + q * ANT, T::N - 1), x0, y0, xmax, ymax, gx, gy);
        r[q] = f[(size_t)kk * pl + (size_t)gy * ix + gx];
    }
}

template <class T>
__device__ __forceinline__ void wstore(double *lds, const double (&r)[T::R])
{
#pragma unroll
    for (int q = 0; q < T::R; ++q) {
        const int e = threadIdx.x + q * ANT;
        if (e < T::N)
            lds[e] = r[q];
    }
}

// block -> (region tile, z-chunk), XCD-contiguous numbering
__device__ __forceinline__ bool acol(const AdjArgs &A, int &x0, int &y0, int &k0, int &k1)
{
    const int b = blockIdx.x;
    const int L = (b & 7) * (A.lb_pad >> 3) + (b >> 3);
    if (L >= A.lb)
        return false;
    const int tx = L % A.ltx, r = L / A.ltx;
    const int ty = r % A.lty, kz = r / A.lty;
    x0 = A.lo[0] + tx * LX;
    y0 = A.lo[1] + ty * LY;
    k0 = A.lo[2] + kz * A.lkb;
    k1 = min(k0 + A.lkb, A.hi[2]);
    return true;
}

}  // namespace

__global__ __launch_bounds__(ANT) void k_adj_hl(AdjArgs A)
{
    // weights of Hz: x -2..+1, y -2..+1; of Hy: x -2..+1; of Hx: y -2..+1
    using TZ = ATile<2, 1, 2, 1>;
    using TY = ATile<2, 1, 0, 0>;
    using TX = ATile<0, 0, 2, 1>;
    __shared__ double sz[3][TZ::N], sy[3][TY::N], sx[3][TX::N];
    int x0, y0, k0, k1;
    if (!acol(A, x0, y0, k0, k1))
        return;
    const int lx = threadIdx.x % LX, ly = threadIdx.x / LX;
    const int i = x0 + lx, jj = y0 + ly;
    const bool valid = i < A.hi[0] && jj < A.hi[1];
    // idle lanes of a partial tile work on their row's last cell (in bounds)
    const int ci = min(i, A.hi[0] - 1), cj = min(jj, A.hi[1] - 1);
    const int ix = A.ix, pl = A.ix * A.iy, col = cj * ix + ci;
    const size_t N = (size_t)A.n;
    const int xmax = min(x0 + LX, A.hi[0]), ymax = min(y0 + LY, A.hi[1]);
    const double c1 = A.c1, c2 = A.c2, rx = A.rdl[0], ry = A.rdl[1], rz = A.rdl[2];
    GLOBAL double *lE = (GLOBAL double *)A.lE;
    // own column's weights of Hy and Hx at planes k-2, k-1
    double yq2 = wgt_h(A, 1, ci, cj, k0 - 2, (k0 - 2) * pl + col);
    double yq1 = wgt_h(A, 1, ci, cj, k0 - 1, (k0 - 1) * pl + col);
    double xq2 = wgt_h(A, 0, ci, cj, k0 - 2, (k0 - 2) * pl + col);
    double xq1 = wgt_h(A, 0, ci, cj, k0 - 1, (k0 - 1) * pl + col);
    double rgz[TZ::R], rgy[TY::R], rgx[TX::R];
    int b0 = 0, b1 = 1, b2 = 2;   // LDS buffers of planes k, k+1 and the one being filled
    wload<TZ, 2, 2, true>(A, 2, k0, x0, y0, xmax, ymax, rgz, rgz);
    wload<TY, 2, 0, true>(A, 1, k0, x0, y0, xmax, ymax, rgy, rgy);
    wload<TX, 0, 2, true>(A, 0, k0, x0, y0, xmax, ymax, rgx, rgx);
    wstore_w<TZ, 2, 2, true>(A, k0, x0, y0, xmax, ymax, sz[b0], rgz, rgz);
    wstore_w<TY, 2, 0, true>(A, k0, x0, y0, xmax, ymax, sy[b0], rgy, rgy);
    wstore_w<TX, 0, 2, true>(A, k0, x0, y0, xmax, ymax, sx[b0], rgx, rgx);
    wload<TZ, 2, 2, true>(A, 2, k0 + 1, x0, y0, xmax, ymax, rgz, rgz);
    wload<TY, 2, 0, true>(A, 1, k0 + 1, x0, y0, xmax, ymax, rgy, rgy);
    wload<TX, 0, 2, true>(A, 0, k0 + 1, x0, y0, xmax, ymax, rgx, rgx);
    size_t c = (size_t)k0 * pl + col;
    double l0 = lE[c], l1 = lE[N + c], l2 = lE[2 * N + c];
    for (int k = k0; k < k1; ++k) {
        wstore_w<TZ, 2, 2, true>(A, k + 1, x0, y0, xmax, ymax, sz[b1], rgz, rgz);
        wstore_w<TY, 2, 0, true>(A, k + 1, x0, y0, xmax, ymax, sy[b1], rgy, rgy);
        wstore_w<TX, 0, 2, true>(A, k + 1, x0, y0, xmax, ymax, sx[b1], rgx, rgx);
        __syncthreads();
        const bool more = k + 1 < k1;
        const size_t cn = c + pl;
        double n0 = 0.0, n1 = 0.0, n2 = 0.0;
        if (more) {
            wload<TZ, 2, 2, true>(A, 2, k + 2, x0, y0, xmax, ymax, rgz, rgz);
            wload<TY, 2, 0, true>(A, 1, k + 2, x0, y0, xmax, ymax, rgy, rgy);
            wload<TX, 0, 2, true>(A, 0, k + 2, x0, y0, xmax, ymax, rgx, rgx);
            n0 = lE[cn];
            n1 = lE[N + cn];
            n2 = lE[2 * N + cn];
        }
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
        // D-_a of the weights of component d
        const double dxy = c1 * x0v - c1 * xym1 + c2 * xyp1 - c2 * xym2;
        const double dxz = c1 * x0v - c1 * xq1 + c2 * xzp1 - c2 * xq2;
        const double dyz = c1 * y0v - c1 * yq1 + c2 * yzp1 - c2 * yq2;
        const double dyx = c1 * y0v - c1 * yxm1 + c2 * yxp1 - c2 * yxm2;
        const double dzx = c1 * z0 - c1 * zxm1 + c2 * zxp1 - c2 * zxm2;
        const double dzy = c1 * z0 - c1 * zym1 + c2 * zyp1 - c2 * zym2;
        if (valid) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            a2 += dxy * ry;
            a1 -= dxz * rz;
            a0 += dyz * rz;
            a2 -= dyx * rx;
            a1 += dzx * rx;
            a0 -= dzy * ry;
            const double e0 = l0 + a0, e1 = l1 + a1, e2 = l2 + a2;
            lE[c] = e0;
            lE[N + c] = e1;
            lE[2 * N + c] = e2;
            if ((long long)c == A.src_cell) {
                const double le = A.src_comp == 0 ? e0 : (A.src_comp == 1 ? e1 : e2);
                A.grad_src[0] += le * A.amp_n / A.sig_src;
            }
        }
        yq2 = yq1;
        yq1 = y0v;
        xq2 = xq1;
        xq1 = x0v;
        l0 = n0;
        l1 = n1;
        l2 = n2;
        c = cn;
        const int t = b0;
        b0 = b1;
        b1 = b2;
        b2 = t;
    }
}

__global__ __launch_bounds__(ANT) void k_adj_el(AdjArgs A)
{
    // weights of Ez: x -1..+2, y -1..+2; of Ey: x -1..+2; of Ex: y -1..+2
    using TZ = ATile<1, 2, 1, 2>;
    using TY = ATile<1, 2, 0, 0>;
    using TX = ATile<0, 0, 1, 2>;
    __shared__ double sz[4][TZ::N], sy[4][TY::N], sx[4][TX::N];
    int x0, y0, k0, k1;
    if (!acol(A, x0, y0, k0, k1))
        return;
    const int lx = threadIdx.x % LX, ly = threadIdx.x / LX;
    const int i = x0 + lx, jj = y0 + ly;
    const bool valid = i < A.hi[0] && jj < A.hi[1];
    const int ci = min(i, A.hi[0] - 1), cj = min(jj, A.hi[1] - 1);
    const int ix = A.ix, pl = A.ix * A.iy, col = cj * ix + ci;
    const size_t N = (size_t)A.n;
    const int xmax = min(x0 + LX, A.hi[0]) + 1, ymax = min(y0 + LY, A.hi[1]) + 1;
    const double c1 = A.c1, c2 = A.c2, rx = A.rdl[0], ry = A.rdl[1], rz = A.rdl[2];
    GLOBAL double *lH = (GLOBAL double *)A.lH;
    // own column's weights of Ey and Ex at plane k-1
    double yq1 = wgt_e(A, 1, ci, cj, k0 - 1, (k0 - 1) * pl + col);
    double xq1 = wgt_e(A, 0, ci, cj, k0 - 1, (k0 - 1) * pl + col);
    size_t c = (size_t)k0 * pl + col;
    double rgz[TZ::R], rgy[TY::R], rgx[TX::R];
    double ccz[TZ::R], ccy[TY::R], ccx[TX::R];   // dt/eps' of the weights' cells
    int b0 = 0, b1 = 1, b2 = 2, b3 = 3;   // planes k, k+1, k+2 (being filled), free
    for (int p = 0; p < 2; ++p) {
        const int bp = p == 0 ? b0 : b1;
        wload<TZ, 1, 1, false>(A, 2, k0 + p, x0, y0, xmax, ymax, rgz, ccz);
        wload<TY, 1, 0, false>(A, 1, k0 + p, x0, y0, xmax, ymax, rgy, ccy);
        wload<TX, 0, 1, false>(A, 0, k0 + p, x0, y0, xmax, ymax, rgx, ccx);
        wstore_w<TZ, 1, 1, false>(A, k0 + p, x0, y0, xmax, ymax, sz[bp], rgz, ccz);
        wstore_w<TY, 1, 0, false>(A, k0 + p, x0, y0, xmax, ymax, sy[bp], rgy, ccy);
        wstore_w<TX, 0, 1, false>(A, k0 + p, x0, y0, xmax, ymax, sx[bp], rgx, ccx);
    }
    wload<TZ, 1, 1, false>(A, 2, k0 + 2, x0, y0, xmax, ymax, rgz, ccz);
    wload<TY, 1, 0, false>(A, 1, k0 + 2, x0, y0, xmax, ymax, rgy, ccy);
    wload<TX, 0, 1, false>(A, 0, k0 + 2, x0, y0, xmax, ymax, rgx, ccx);
    double h0 = lH[c], h1 = lH[N + c], h2 = lH[2 * N + c];
    for (int k = k0; k < k1; ++k) {
        wstore_w<TZ, 1, 1, false>(A, k + 2, x0, y0, xmax, ymax, sz[b2], rgz, ccz);
        wstore_w<TY, 1, 0, false>(A, k + 2, x0, y0, xmax, ymax, sy[b2], rgy, ccy);
        wstore_w<TX, 0, 1, false>(A, k + 2, x0, y0, xmax, ymax, sx[b2], rgx, ccx);
        __syncthreads();
        const bool more = k + 1 < k1;
        const size_t cn = c + pl;
        double n0 = 0.0, n1 = 0.0, n2 = 0.0;
        if (more) {
            wload<TZ, 1, 1, false>(A, 2, k + 3, x0, y0, xmax, ymax, rgz, ccz);
            wload<TY, 1, 0, false>(A, 1, k + 3, x0, y0, xmax, ymax, rgy, ccy);
            wload<TX, 0, 1, false>(A, 0, k + 3, x0, y0, xmax, ymax, rgx, ccx);
            n0 = lH[cn];
            n1 = lH[N + cn];
            n2 = lH[2 * N + cn];
        }
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
        // D+_a of the weights of component d
        const double pxy = c1 * xyp1 - c1 * x0v + c2 * xyp2 - c2 * xym1;
        const double pxz = c1 * xzp1 - c1 * x0v + c2 * xzp2 - c2 * xq1;
        const double pyz = c1 * yzp1 - c1 * y0v + c2 * yzp2 - c2 * yq1;
        const double pyx = c1 * yxp1 - c1 * y0v + c2 * yxp2 - c2 * yxm1;
        const double pzx = c1 * zxp1 - c1 * z0 + c2 * zxp2 - c2 * zxm1;
        const double pzy = c1 * zyp1 - c1 * z0 + c2 * zyp2 - c2 * zym1;
        if (valid) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            a2 -= pxy * ry;
            a1 += pxz * rz;
            a0 -= pyz * rz;
            a2 += pyx * rx;
            a1 -= pzx * rx;
            a0 += pzy * ry;
            lH[c] = h0 + a0;
            lH[N + c] = h1 + a1;
            lH[2 * N + c] = h2 + a2;
        }
        yq1 = y0v;
        xq1 = x0v;
        h0 = n0;
        h1 = n1;
        h2 = n2;
        c = cn;
        const int t = b0;
        b0 = b1;
        b1 = b2;
        b2 = b3;
        b3 = t;
    }
}

// E-phase gradient over the region: grad += sum_d lE_d dcb/de curl-_d(H^n)
// with dcb/de = -dt/eps'^2, the forward E kernel's stencil pattern on H^n
__global__ __launch_bounds__(ANT) void k_adj_gl(AdjArgs A)
{
    // H^n_z: x -2..+1, y -2..+1; H^n_y: x -2..+1; H^n_x: y -2..+1
    using TZ = ATile<2, 1, 2, 1>;
    using TY = ATile<2, 1, 0, 0>;
    using TX = ATile<0, 0, 2, 1>;
    __shared__ double sz[3][TZ::N], sy[3][TY::N], sx[3][TX::N];
    int x0, y0, k0, k1;
    if (!acol(A, x0, y0, k0, k1))
        return;
    const int lx = threadIdx.x % LX, ly = threadIdx.x / LX;
    const int i = x0 + lx, jj = y0 + ly;
    const bool valid = i < A.hi[0] && jj < A.hi[1];
    const int ci = min(i, A.hi[0] - 1), cj = min(jj, A.hi[1] - 1);
    const int ix = A.ix, pl = A.ix * A.iy, col = cj * ix + ci;
    const size_t N = (size_t)A.n;
    const int xmax = min(x0 + LX, A.hi[0]), ymax = min(y0 + LY, A.hi[1]);
    const double c1 = A.c1, c2 = A.c2, rx = A.rdl[0], ry = A.rdl[1], rz = A.rdl[2];
    const double dt = A.dt;
    const GLOBAL double *Hx = (const GLOBAL double *)A.uH0;
    const GLOBAL double *Hy = Hx + N, *Hz = Hx + 2 * N;
    const GLOBAL double *lE = (const GLOBAL double *)A.lE;
    const GLOBAL double *eps = (const GLOBAL double *)A.eps;
    GLOBAL double *grad = (GLOBAL double *)A.grad;
    size_t c = (size_t)k0 * pl + col;
    double yq2 = Hy[c - 2 * pl], yq1 = Hy[c - pl], xq2 = Hx[c - 2 * pl], xq1 = Hx[c - pl];
    double rgz[TZ::R], rgy[TY::R], rgx[TX::R];
    int b0 = 0, b1 = 1, b2 = 2;
    fload<TZ, 2, 2>(Hz, k0, ix, pl, x0, y0, xmax, ymax, rgz);
    fload<TY, 2, 0>(Hy, k0, ix, pl, x0, y0, xmax, ymax, rgy);
    fload<TX, 0, 2>(Hx, k0, ix, pl, x0, y0, xmax, ymax, rgx);
    wstore<TZ>(sz[b0], rgz);
    wstore<TY>(sy[b0], rgy);
    wstore<TX>(sx[b0], rgx);
    fload<TZ, 2, 2>(Hz, k0 + 1, ix, pl, x0, y0, xmax, ymax, rgz);
    fload<TY, 2, 0>(Hy, k0 + 1, ix, pl, x0, y0, xmax, ymax, rgy);
    fload<TX, 0, 2>(Hx, k0 + 1, ix, pl, x0, y0, xmax, ymax, rgx);
    double e0 = lE[c], e1 = lE[N + c], e2 = lE[2 * N + c], ep = eps[c], g0 = grad[c];
    for (int k = k0; k < k1; ++k) {
        wstore<TZ>(sz[b1], rgz);
        wstore<TY>(sy[b1], rgy);
        wstore<TX>(sx[b1], rgx);
        __syncthreads();
        const bool more = k + 1 < k1;
        const size_t cn = c + pl;
        double n0 = 0.0, n1 = 0.0, n2 = 0.0, nep = 1.0, ng = 0.0;
        if (more) {
            fload<TZ, 2, 2>(Hz, k + 2, ix, pl, x0, y0, xmax, ymax, rgz);
            fload<TY, 2, 0>(Hy, k + 2, ix, pl, x0, y0, xmax, ymax, rgy);
            fload<TX, 0, 2>(Hx, k + 2, ix, pl, x0, y0, xmax, ymax, rgx);
            n0 = lE[cn];
            n1 = lE[N + cn];
            n2 = lE[2 * N + cn];
            nep = eps[cn];
            ng = grad[cn];
        }
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
        if (valid) {
            const double r = 1.0 / ep, cbd = -dt * r * r;
            const double cu0 = hzy * ry - hyz * rz;
            const double cu1 = hxz * rz - hzx * rx;
            const double cu2 = hyx * rx - hxy * ry;
            double g = 0.0;
            g += e0 * cbd * cu0;
            g += e1 * cbd * cu1;
            g += e2 * cbd * cu2;
            grad[c] = g0 + g;
        }
        yq2 = yq1;
        yq1 = y0v;
        xq2 = xq1;
        xq1 = x0v;
        e0 = n0;
        e1 = n1;
        e2 = n2;
        ep = nep;
        g0 = ng;
        c = cn;
        const int t = b0;
        b0 = b1;
        b1 = b2;
        b2 = t;
    }
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
    if (A.lb_pad > 0) {
        hipLaunchKernelGGL(k_adj_hl, dim3(A.lb_pad), dim3(ANT), 0, st, A);
        hipLaunchKernelGGL(k_adj_gl, dim3(A.lb_pad), dim3(ANT), 0, st, A);
    }
    hipLaunchKernelGGL(k_adj_e1, dim3(A.fb_pad), dim3(ZX * ZY), 0, st, A);
    hipLaunchKernelGGL(k_adj_e2, dim3(A.fb_pad), dim3(ZX * ZY), 0, st, A);
    if (A.lb_pad > 0)
        hipLaunchKernelGGL(k_adj_el, dim3(A.lb_pad), dim3(ANT), 0, st, A);
    return hipGetLastError();
}

}  // namespace mt
