// Kernel-side parameter block and launcher declarations for pool.hip.
#pragma once
#include "owl_amd_common.h"

namespace owl_amd {

struct PoolParams {
  int B, H, W, D, C;
  int kh, kw, kd;
  int Ho, Wo, Do;
  int sh, sw, sd;
  int ph, pw, pd;
};

// mode: 0 average, 1 max, 2 max with int64 argmax
template <typename T>
void run_pool_fwd(const PoolParams &p, int mode, const T *x, T *y, int64_t *arg, hipStream_t st);
template <typename T>
void run_pool_bwd(const PoolParams &p, bool is_max, const T *x, const T *dy, T *dx, int *flag,
                  hipStream_t st);
void run_set_flag(int *flag, hipStream_t st);

}  // namespace owl_amd
