// fdtd_transform.hip -- complex-row DFTs for the reference's post-processing:
// the diffusive-time back-transform (x_t[k] = 2/(nd dt) sum_n X[n]
// exp(-+ 2 pi i k n / nd): satan_i/fwi3d_cpu_function_v1.4.c:1507-1527 with
// FFTW_FORWARD, ver0040/cpu_f_to_d_e.f90:281-320 with FFTW_BACKWARD) and the
// spectra product of convolution_GJ_to_E (function.c:1566-1635).
//
// A direct O(nd^2) transform per row: one wavefront per output bin, lanes take
// the inputs in a fixed stride order and a fixed butterfly reduces them, so the
// result does not depend on scheduling.  Twiddles come from one table of nd
// entries indexed by k n mod nd.

#include "fdtd_transform.h"

#define GLOBAL __attribute__((address_space(1)))

namespace mt {

namespace {

constexpr int kDftWaves = 4;   // output bins per 256-thread block

__global__ __launch_bounds__(256) void k_twiddle(int nd, double *__restrict__ tw)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nd)
        return;
    double s, c;
    sincospi(2.0 * m / nd, &s, &c);
    GLOBAL double *t = (GLOBAL double *)tw;
    t[2 * m] = c;
    t[2 * m + 1] = s;
}

__global__ __launch_bounds__(64 * kDftWaves) void k_dft_rows(const double *__restrict__ in, int nd,
                                                             int sign, double scale,
                                                             const double *__restrict__ tw,
                                                             double *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * kDftWaves + (threadIdx.x >> 6);
    const int r = blockIdx.y;
    if (k >= nd)
        return;   // uniform over the wavefront
    const GLOBAL double *x = (const GLOBAL double *)in + 2 * (size_t)r * nd;
    const GLOBAL double *T = (const GLOBAL double *)tw;
    const int step = (int)((64LL * k) % nd);
    int q = (int)(((long long)lane * k) % nd);
    double re = 0.0, im = 0.0;
    for (int n = lane; n < nd; n += 64) {
        const double c = T[2 * q], s = sign > 0 ? T[2 * q + 1] : -T[2 * q + 1];
        const double a = x[2 * n], b = x[2 * n + 1];
        re += a * c - b * s;
        im += a * s + b * c;
        q += step;
        if (q >= nd)
            q -= nd;
    }
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
    }
    if (lane == 0) {
        GLOBAL double *o = (GLOBAL double *)out;
        o[2 * ((size_t)r * nd + k)] = scale * re;
        o[2 * ((size_t)r * nd + k) + 1] = scale * im;
    }
}

__global__ __launch_bounds__(256) void k_cmul_rows(double *__restrict__ a,
                                                   const double *__restrict__ b, int rows, int nd)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)rows * nd)
        return;
    const int k = (int)(q % nd);
    GLOBAL double *A = (GLOBAL double *)a;
    const GLOBAL double *B = (const GLOBAL double *)b;
    const double ar = A[2 * q], ai = A[2 * q + 1], br = B[2 * k], bi = B[2 * k + 1];
    A[2 * q] = ar * br - ai * bi;
    A[2 * q + 1] = ar * bi + ai * br;
}

}  // namespace

hipError_t launch_twiddle(int nd, double *tw, hipStream_t st)
{
    hipLaunchKernelGGL(k_twiddle, dim3((nd + 255) / 256), dim3(256), 0, st, nd, tw);
    return hipGetLastError();
}

hipError_t launch_dft_rows(const double *in, int rows, int nd, int sign, double scale,
                           const double *tw, double *out, hipStream_t st)
{
    if (rows <= 0 || nd <= 0)
        return hipSuccess;
    if (rows > 65535)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dft_rows, dim3((nd + kDftWaves - 1) / kDftWaves, rows), dim3(64 * kDftWaves),
                       0, st, in, nd, sign, scale, tw, out);
    return hipGetLastError();
}

hipError_t launch_cmul_rows(double *a, const double *b, int rows, int nd, hipStream_t st)
{
    const size_t n = (size_t)rows * nd;
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_cmul_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, rows,
                       nd);
    return hipGetLastError();
}

}  // namespace mt
