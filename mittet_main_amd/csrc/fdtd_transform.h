// fdtd_transform.h -- device DFTs of complex rows: the diffusive-time
// back-transform and the G*J convolution of the reference's post-processing
// (fdtd_transform.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace mt {

// tw[m] = exp(2 pi i m / nd) for m < nd (complex, interleaved)
hipError_t launch_twiddle(int nd, double *tw, hipStream_t st);
// out[r][k] = scale * sum_n in[r][n] exp(sign 2 pi i k n / nd), r < rows (<= 65535)
hipError_t launch_dft_rows(const double *in, int rows, int nd, int sign, double scale,
                           const double *tw, double *out, hipStream_t st);
// a[r][k] *= b[k] (complex)
hipError_t launch_cmul_rows(double *a, const double *b, int rows, int nd, hipStream_t st);

}  // namespace mt
