// Error reporting and small C-ABI helpers of librgp.
#include <stdarg.h>
#include <stdio.h>

#include "rgp_internal.h"

static thread_local char g_err[512] = "";

// The host packs cameras as 32 float32 words; arrays of them have that stride.
static_assert(sizeof(rgp_camera) == 128, "rgp_camera is 32 words");

namespace rgp {

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* who) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(RGP_ERR_LAUNCH, "%s: %s", who, hipGetErrorString(e));
  return RGP_OK;
}

}  // namespace rgp

extern "C" const char* rgp_last_error(void) { return g_err; }

extern "C" int rgp_version(void) { return 1; }

// Single-image mix: a batch of one, with the per-image arrays taken from a
// small device scratch passed in the palette argument's neighbourhood is not
// possible without an allocation, so the single form uses a kernel that takes
// its arguments by value.
namespace rgp {
template <int K>
__global__ __launch_bounds__(256) void k_mix1(int64_t hw, const float* __restrict__ Wp,
                                              const float* __restrict__ pal,
                                              float* __restrict__ out) {
  float P[K][3];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) P[k][c] = pal[k * 3 + c];
  const int64_t p4 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (p4 >= hw) return;
  if ((hw & 3) == 0) {
    float4 wv[K];
#pragma unroll
    for (int k = 0; k < K; ++k) wv[k] = *(const float4*)(Wp + k * hw + p4);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float4 r;
      r.x = P[0][c] * wv[0].x; r.y = P[0][c] * wv[0].y;
      r.z = P[0][c] * wv[0].z; r.w = P[0][c] * wv[0].w;
#pragma unroll
      for (int k = 1; k < K; ++k) {
        r.x = r.x + P[k][c] * wv[k].x; r.y = r.y + P[k][c] * wv[k].y;
        r.z = r.z + P[k][c] * wv[k].z; r.w = r.w + P[k][c] * wv[k].w;
      }
      *(float4*)(out + c * hw + p4) = r;
    }
  } else {
    for (int64_t p = p4; p < p4 + 4 && p < hw; ++p)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float r = P[0][c] * Wp[p];
#pragma unroll
        for (int k = 1; k < K; ++k) r = r + P[k][c] * Wp[k * hw + p];
        out[c * hw + p] = r;
      }
  }
}
}  // namespace rgp

extern "C" int rgp_mix(int K, int64_t hw, const float* W, const float* palette, float* rgb,
                       hipStream_t s) {
  using namespace rgp;
  if (K < 1 || K > RGP_MAX_DIM || hw < 0) return set_error(RGP_ERR_ARG, "rgp_mix: bad K or size");
  if (hw == 0) return RGP_OK;
  const unsigned blocks = (unsigned)((hw + 1023) / 1024);
  switch (K) {
    case 1: hipLaunchKernelGGL(k_mix1<1>, dim3(blocks), dim3(256), 0, s, hw, W, palette, rgb); break;
    case 2: hipLaunchKernelGGL(k_mix1<2>, dim3(blocks), dim3(256), 0, s, hw, W, palette, rgb); break;
    case 3: hipLaunchKernelGGL(k_mix1<3>, dim3(blocks), dim3(256), 0, s, hw, W, palette, rgb); break;
    case 4: hipLaunchKernelGGL(k_mix1<4>, dim3(blocks), dim3(256), 0, s, hw, W, palette, rgb); break;
    case 5: hipLaunchKernelGGL(k_mix1<5>, dim3(blocks), dim3(256), 0, s, hw, W, palette, rgb); break;
    case 6: hipLaunchKernelGGL(k_mix1<6>, dim3(blocks), dim3(256), 0, s, hw, W, palette, rgb); break;
    case 7: hipLaunchKernelGGL(k_mix1<7>, dim3(blocks), dim3(256), 0, s, hw, W, palette, rgb); break;
    case 8: hipLaunchKernelGGL(k_mix1<8>, dim3(blocks), dim3(256), 0, s, hw, W, palette, rgb); break;
  }
  return check_launch("rgp_mix");
}
