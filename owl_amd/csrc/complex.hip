// Complex convolutions (Owl's complex32 / complex64, owl_ndarray_conv_stub.c:91-122)
// as real ones.  A complex array [..][C] is the real array [..][2C] with the
// real and imaginary parts interleaved, so with x = xr + i xi and
// w = wr + i wi:
//   y[o]   = sum_q x[q] w[q][o]            (forward)
//   dx[q]  = sum_o dy[o] w[q][o]           (backward input: Owl's complex
//                                           GEMMs transpose without conjugating)
//   dw[q][o] = sum_pixels x[q] dy[o]       (backward kernel)
// become real convolutions over 2 Ci -> 2 Co channels with the kernel's
// 2 x 2 blocks
//   forward:        [ wr  wi ]   rows 2q, 2q+1; columns 2o, 2o+1
//                   [-wi  wr ]
//   backward input: [ wr -wi ]
//                   [ wi  wr ]
// and the weight gradient of the real convolution, dW[2q+a][2o+b], folds
// back as dw = (dW[2q][2o] - dW[2q+1][2o+1]) + i (dW[2q][2o+1] + dW[2q+1][2o]).
#include "owl_amd_common.h"
#include "conv_kernels.h"

namespace owl_amd {

// w: [n = T Ci Co] complex elements -> wr: [T][2 Ci][2 Co] real
template <typename R>
__global__ __launch_bounds__(256) void cplx_expand_kernel(const R *__restrict__ w, R *__restrict__ wr, int64_t n,
                                                          int Ci, int Co, int pattern) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int o = (int)(e % Co);
    const int64_t r = e / Co;
    const int q = (int)(r % Ci);
    const int64_t t = r / Ci;
    const R re = w[2 * e], im = w[2 * e + 1];
    R *row0 = wr + ((t * 2 * Ci + 2 * q) * (int64_t)(2 * Co) + 2 * o);
    R *row1 = row0 + 2 * Co;
    row0[0] = re;
    row1[1] = re;
    if (pattern == 0) {
      row0[1] = im;
      row1[0] = -im;
    } else {
      row0[1] = -im;
      row1[0] = im;
    }
  }
}

// dwr: [T][2 Ci][2 Co] real -> dw: [n = T Ci Co] complex elements
template <typename R>
__global__ __launch_bounds__(256) void cplx_fold_kernel(const R *__restrict__ dwr, R *__restrict__ dw, int64_t n,
                                                        int Ci, int Co) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int o = (int)(e % Co);
    const int64_t r = e / Co;
    const int q = (int)(r % Ci);
    const int64_t t = r / Ci;
    const R *row0 = dwr + ((t * 2 * Ci + 2 * q) * (int64_t)(2 * Co) + 2 * o);
    const R *row1 = row0 + 2 * Co;
    dw[2 * e] = row0[0] - row1[1];
    dw[2 * e + 1] = row0[1] + row1[0];
  }
}

static dim3 elem_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

template <typename R>
void run_cplx_expand(const R *w, R *wr, int64_t T, int Ci, int Co, int pattern, hipStream_t st) {
  const int64_t n = T * Ci * (int64_t)Co;
  launch_k((cplx_expand_kernel<R>), elem_grid(n), dim3(256), st, w, wr, n, Ci, Co, pattern);
}

template <typename R>
void run_cplx_fold(const R *dwr, R *dw, int64_t T, int Ci, int Co, hipStream_t st) {
  const int64_t n = T * Ci * (int64_t)Co;
  launch_k((cplx_fold_kernel<R>), elem_grid(n), dim3(256), st, dwr, dw, n, Ci, Co);
}

template void run_cplx_expand<float>(const float *, float *, int64_t, int, int, int, hipStream_t);
template void run_cplx_expand<double>(const double *, double *, int64_t, int, int, int, hipStream_t);
template void run_cplx_fold<float>(const float *, float *, int64_t, int, int, hipStream_t);
template void run_cplx_fold<double>(const double *, double *, int64_t, int, int, hipStream_t);

}  // namespace owl_amd
