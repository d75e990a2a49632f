// fdtd_compat.hip -- the reference's own FWI gradient and model update, on the
// GPU: frequency-domain correlation of forward fields with reciprocal Green's
// functions (satan_i/fwi3d_cpu_function_v1.4.c:2202-2384) and the
// step-length / update rules (:2565-2832).
//
// Spectra are complex [comp][bin][cell] (interleaved re, im).  A Green's
// function run is streamed through every source's accumulators as soon as it
// finishes, so the reference's per-receiver Green's volumes are never stored;
// the per-cell summation order (receiver, direction, bin, then x, y, z) is the
// reference's.

#include "fdtd_dev.h"

#define GLOBAL __attribute__((address_space(1)))

namespace mt {

namespace {

__device__ __forceinline__ cplx ld(const double *p, size_t q)
{
    const GLOBAL double *g = (const GLOBAL double *)p;
    return {g[2 * q], g[2 * q + 1]};
}

__device__ __forceinline__ void st(double *p, size_t q, cplx v)
{
    GLOBAL double *g = (GLOBAL double *)p;
    g[2 * q] = v.re;
    g[2 * q + 1] = v.im;
}

}  // namespace

// laplaceToFreq_2 / _back per cell: G = E/J per bin; with tw, E_ff = G * T
// (T = FFTW_BACKWARD of the true waveform at that bin), else G itself
__global__ __launch_bounds__(256) void k_compat_eff(const double *__restrict__ acc, long long astride,
                                                    int n, int nb, const double *__restrict__ jw,
                                                    const double *__restrict__ tw,
                                                    double *__restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n)
        return;
    for (int d = 0; d < 3; ++d)
        for (int b = 0; b < nb; ++b) {
            const cplx g = cdiv(ld(acc, ((size_t)d * nb + b) * astride + c), ld(jw, b));
            st(out, ((size_t)d * nb + b) * n + c, tw ? cmul(g, ld(tw, b)) : g);
        }
}

// sensitivity + conv_GdelE for one (receiver, direction) Green's function G,
// for every source s:  sense_s += |E_ff,s . G| (bins, then x, y, z) and
// E_back,s += G conj(dE_s)
__global__ __launch_bounds__(256) void k_compat_green(int n, int nb, int nsrc,
                                                      const double *__restrict__ G,
                                                      const double *__restrict__ eff,
                                                      const double *__restrict__ dconj,
                                                      double *__restrict__ sense,
                                                      double *__restrict__ eback)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n)
        return;
    const size_t vol = (size_t)3 * nb * n;
    for (int s = 0; s < nsrc; ++s) {
        double sen = sense[(size_t)s * n + c];
        for (int b = 0; b < nb; ++b) {
            const cplx dc = ld(dconj, (size_t)s * nb + b);
            for (int d = 0; d < 3; ++d) {
                const size_t q = ((size_t)d * nb + b) * n + c;
                const cplx g = ld(G, q);
                const cplx p = cmul(ld(eff, s * vol + q), g);
                sen += hypot(p.re, p.im);
                const cplx e = ld(eback, s * vol + q), t = cmul(g, dc);
                st(eback, s * vol + q, {e.re + t.re, e.im + t.im});
            }
        }
        sense[(size_t)s * n + c] = sen;
    }
}

// calc_gamma: grad += sum_s sum_bins Re(E_ff . E_back)/sense_s on the cells
// ncpml <= i,j,k < n - ncpml; `nlow` bins below the first computed one add
// 0/sense as the reference's loop over [0, upperF) does
__global__ __launch_bounds__(256) void k_compat_gamma(int ix, int iy, int iz, int nc, int nb,
                                                      int nlow, int nsrc,
                                                      const double *__restrict__ eff,
                                                      const double *__restrict__ eback,
                                                      const double *__restrict__ sense,
                                                      double *__restrict__ grad)
{
    const int n = ix * iy * iz;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n)
        return;
    const int i = c % ix, r = c / ix, j = r % iy, k = r / iy;
    if (i < nc || i >= ix - nc || j < nc || j >= iy - nc || k < nc || k >= iz - nc)
        return;
    const size_t vol = (size_t)3 * nb * n;
    double g = grad[c];
    for (int s = 0; s < nsrc; ++s) {
        const double sen = sense[(size_t)s * n + c];
        for (int b = 0; b < nlow; ++b)
            g += 0.0 / sen;
        for (int b = 0; b < nb; ++b) {
            double re = 0.0;
            for (int d = 0; d < 3; ++d) {
                const size_t q = s * vol + ((size_t)d * nb + b) * n + c;
                const cplx p = cmul(ld(eff, q), ld(eback, q));
                re = d == 0 ? p.re : re + p.re;
            }
            g += re / sen;
        }
    }
    grad[c] = g;
}

// raw accumulators at the receiver cells: out [3][nrec][nb] complex
__global__ void k_compat_gather(const double *__restrict__ acc, long long astride, int nb, int nrec,
                                const int *__restrict__ rec, double *__restrict__ out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3 * nrec * nb)
        return;
    const int b = q % nb, r = (q / nb) % nrec, d = q / (nb * nrec);
    st(out, q, ld(acc, ((size_t)d * nb + b) * astride + rec[r]));
}

// calc_phi, first pass: per-block maxima of sigma and grad over k <= iseabed
// (-inf where a block has no such cell); the host finishes the reduction
__global__ __launch_bounds__(256) void k_phi_max(const double *__restrict__ sig,
                                                 const double *__restrict__ grad, int n, int nxy,
                                                 int iseabed, double *__restrict__ blockmax)
{
    __shared__ double ms[256], mg[256];
    const int t = threadIdx.x;
    double a = -__builtin_inf(), b = -__builtin_inf();
    for (int c = blockIdx.x * 256 + t; c < n; c += gridDim.x * 256)
        if (c / nxy <= iseabed) {
            a = sig[c] > a ? sig[c] : a;
            b = grad[c] > b ? grad[c] : b;
        }
    ms[t] = a;
    mg[t] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            ms[t] = ms[t + h] > ms[t] ? ms[t + h] : ms[t];
            mg[t] = mg[t + h] > mg[t] ? mg[t + h] : mg[t];
        }
        __syncthreads();
    }
    if (t == 0) {
        blockmax[2 * blockIdx.x] = ms[0];
        blockmax[2 * blockIdx.x + 1] = mg[0];
    }
}

// calc_phi, second pass: sig2 = sig - scaler * grad on every cell
__global__ __launch_bounds__(256) void k_phi_apply(const double *__restrict__ sig,
                                                   const double *__restrict__ grad, double scaler,
                                                   int n, double *__restrict__ sig2)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n)
        sig2[c] = sig[c] - scaler * grad[c];
}

// update_para2 on the cells ncpml <= i,j,k < n - ncpml: sig += step*grad where
// k <= iseabed (step = alpha*scaler), then clamp to [minval, maxval]
__global__ __launch_bounds__(256) void k_update(double *__restrict__ sig,
                                                const double *__restrict__ grad, double step,
                                                double minval, double maxval, int ix, int iy,
                                                int iz, int nc, int iseabed)
{
    const int n = ix * iy * iz;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n)
        return;
    const int i = c % ix, r = c / ix, j = r % iy, k = r / iy;
    if (i < nc || i >= ix - nc || j < nc || j >= iy - nc || k < nc || k >= iz - nc)
        return;
    double v = sig[c];
    if (k <= iseabed)
        v += step * grad[c];
    if (v < minval)
        v = minval;
    if (v > maxval)
        v = maxval;
    sig[c] = v;
}

static inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_compat_eff(const double *acc, long long astride, int n, int nb,
                             const double *jw, const double *tw, double *eff, hipStream_t st)
{
    hipLaunchKernelGGL(k_compat_eff, dim3(blocks_for(n)), dim3(256), 0, st, acc, astride, n, nb,
                       jw, tw, eff);
    return hipGetLastError();
}

hipError_t launch_compat_green(const double *acc, long long astride, int n, int nb,
                               const double *jw, int nsrc, const double *eff,
                               const double *dconj, double *G, double *sense, double *eback,
                               hipStream_t st)
{
    hipLaunchKernelGGL(k_compat_eff, dim3(blocks_for(n)), dim3(256), 0, st, acc, astride, n, nb,
                       jw, (const double *)nullptr, G);
    hipLaunchKernelGGL(k_compat_green, dim3(blocks_for(n)), dim3(256), 0, st, n, nb, nsrc,
                       (const double *)G, eff, dconj, sense, eback);
    return hipGetLastError();
}

hipError_t launch_compat_gamma(int ix, int iy, int iz, int nc, int nb, int nlow, int nsrc,
                               const double *eff, const double *eback, const double *sense,
                               double *grad, hipStream_t st)
{
    hipLaunchKernelGGL(k_compat_gamma, dim3(blocks_for((long long)ix * iy * iz)), dim3(256), 0, st,
                       ix, iy, iz, nc, nb, nlow, nsrc, eff, eback, sense, grad);
    return hipGetLastError();
}

hipError_t launch_compat_gather(const double *acc, long long astride, int nb, int nrec,
                                const int *rec, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_compat_gather, dim3(blocks_for(3LL * nrec * nb)), dim3(256), 0, st, acc,
                       astride, nb, nrec, rec, out);
    return hipGetLastError();
}

hipError_t launch_phi_max(const double *sig, const double *grad, int n, int nxy, int iseabed,
                          double *blockmax, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_phi_max, dim3(nblocks), dim3(256), 0, st, sig, grad, n, nxy, iseabed,
                       blockmax);
    return hipGetLastError();
}

hipError_t launch_phi_apply(const double *sig, const double *grad, double scaler, int n,
                            double *sig2, hipStream_t st)
{
    hipLaunchKernelGGL(k_phi_apply, dim3(blocks_for(n)), dim3(256), 0, st, sig, grad, scaler, n,
                       sig2);
    return hipGetLastError();
}

hipError_t launch_update(double *sig, const double *grad, double step, double minval,
                         double maxval, int ix, int iy, int iz, int nc, int iseabed,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_update, dim3(blocks_for((long long)ix * iy * iz)), dim3(256), 0, st, sig,
                       grad, step, minval, maxval, ix, iy, iz, nc, iseabed);
    return hipGetLastError();
}

}  // namespace mt
