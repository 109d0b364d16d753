// Fused (1 - lambda) L1 + lambda (1 - SSIM) image loss of the finetune step
// for gfx950, forward and backward.  SSIM is the 11x11 Gaussian-window form
// of gsplat/pytorch_ssim.py:17-66 with zero padding; the window is applied as
// two separable passes over LDS tiles (it is an outer product).
//
// Forward, per 32x32 output tile and channel: stage the image and target with
// a 5-pixel halo in LDS, blur the five maps x, y, x^2, y^2, xy horizontally
// then vertically, evaluate the SSIM map and, per pixel, the partials of the
// SSIM map with respect to the blurred maps that depend on the image:
//   a = ds/dmu_x, b = ds/dE[x^2], c = ds/dE[xy]
// (s = N/D, N = (2 mu_x mu_y + C1)(2 sxy + C2), D = (mu_x^2 + mu_y^2 + C1)
// (sx^2 + sy^2 + C2)).  Each block writes its sums of the SSIM map and of
// |x - y| to its own pair of float64 slots; one block then adds the slots in
// a fixed order, so the loss value does not depend on the block schedule.
// Backward: the adjoint of a zero-padded blur with a symmetric window is the
// same blur, so dS/dx = G*a + 2 x (G*b) + y (G*c), evaluated with the same
// LDS tiling; the L1 part adds sign(x - y).
#include "rgp_internal.h"

namespace rgp {

constexpr int kST = 32;            // output tile edge
constexpr int kSR = 5;             // window radius
constexpr int kSI = kST + 2 * kSR;  // staged tile edge

struct Win11 {
  float g[11];
};

__device__ __forceinline__ float block_sum_f(float v, float* lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v = wave_sum_to63(v);
  if (lane == 63) lds[w] = v;
  __syncthreads();
  float s = 0.f;
  if (threadIdx.x == 0)
    for (int j = 0; j < (int)(blockDim.x >> 6); ++j) s += lds[j];
  return s;
}

// Horizontal pass items: a row r of the staged tile and a group of eight
// output columns (42 x 4 items, one per thread of the first 168); vertical
// pass: a group of four output rows and a column.  Each item loads the inputs
// its outputs share once (18 for eight outputs, 14 for four), so products and
// LDS reads are not repeated per tap; sums use fmaf (the loss is compared
// with a float64 reference within a tolerance, not bit for bit).
constexpr int kSG = kST / 4;   // four-output groups per tile edge (vertical)
constexpr int kHO = 8;         // outputs per horizontal item
constexpr int kHG = kST / kHO;  // horizontal items per staged row
constexpr int kSB = 256;       // threads per block
static_assert(kSI * kHG <= kSB, "one horizontal item per thread");

// Stages the (kSI x kSI) windows at (x0, y0) of NP planes, zero outside the
// image, into d[p][r][q]: every load of a thread (7 per plane) is issued
// before the first LDS store.  With kSB = 6 kSI + 4 the row and column of a
// thread's next element follow from the current ones without a division.
template <int NP>
__device__ __forceinline__ void ssim_stage(const float* const (&P)[NP], int H, int W, int x0, int y0,
                                           float (*const (&d)[NP])[kSI + 1]) {
  static_assert(kSB == 6 * kSI + 4, "stepping below assumes 256 = 6 * 42 + 4");
  constexpr int IT = (kSI * kSI + kSB - 1) / kSB;
  const int t = threadIdx.x;
  float v[NP][IT];
  int r = t / kSI, q = t - r * kSI;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int gy = y0 + r, gx = x0 + q;
    const bool in = t + it * kSB < kSI * kSI && gy >= 0 && gy < H && gx >= 0 && gx < W;
#pragma unroll
    for (int p = 0; p < NP; ++p) v[p][it] = in ? P[p][gy * W + gx] : 0.f;
    r += 6;
    q += 4;
    if (q >= kSI) {
      q -= kSI;
      ++r;
    }
  }
  r = t / kSI;
  q = t - r * kSI;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    if (t + it * kSB < kSI * kSI) {
#pragma unroll
      for (int p = 0; p < NP; ++p) d[p][r][q] = v[p][it];
    }
    r += 6;
    q += 4;
    if (q >= kSI) {
      q -= kSI;
      ++r;
    }
  }
}

// The block's output tile (bx, by) of channel c.  The grid is one dimension,
// tiles in row-major order per channel, and each XCD takes a contiguous band
// of them: the 5-pixel halos a tile shares with its neighbours are then read
// through the same L2 (blocks go round-robin to the 8 XCDs).
__device__ __forceinline__ void ssim_block(int H, int W, int* bx, int* by, int* c) {
  const int nbx = (W + kST - 1) / kST, nby = (H + kST - 1) / kST;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  *bx = t % nbx;
  *by = (t / nbx) % nby;
  *c = t / (nbx * nby);
}

// Blurs three maps of the staged tile: horizontally into hb (the products of
// an input pair formed once per input, not once per tap), then vertically
// into mu[m][o] for this thread's four outputs (column q, rows r0..r0+3).
// MAPS selects (x, y, x^2) or (y^2, x y).
template <int MAPS>
__device__ __forceinline__ void ssim_blur_phase(const float (*sx)[kSI + 1], const float (*sy)[kSI + 1],
                                                float (*hb)[kSI][kST + 1], const Win11& win, int q, int r0,
                                                float (*mu)[4]) {
  constexpr int NM = MAPS == 0 ? 3 : 2;
  constexpr int NI = kHO + 10;
  if (threadIdx.x < kSI * kHG) {
    const int r = threadIdx.x % kSI, c0 = kHO * (threadIdx.x / kSI);
    float in[NM][NI];
#pragma unroll
    for (int j = 0; j < NI; ++j) {
      const float a = sx[r][c0 + j], b = sy[r][c0 + j];
      if (MAPS == 0) {
        in[0][j] = a;
        in[1][j] = b;
        in[2][j] = a * a;
      } else {
        in[0][j] = b * b;
        in[1][j] = a * b;
      }
    }
#pragma unroll
    for (int o = 0; o < kHO; ++o) {
      float acc[NM];
#pragma unroll
      for (int m = 0; m < NM; ++m) acc[m] = 0.f;
#pragma unroll
      for (int j = 0; j < 11; ++j)
#pragma unroll
        for (int m = 0; m < NM; ++m) acc[m] = fmaf(win.g[j], in[m][o + j], acc[m]);
#pragma unroll
      for (int m = 0; m < NM; ++m) hb[m][r][c0 + o] = acc[m];
    }
  }
  __syncthreads();
  float v[NM][14];
#pragma unroll
  for (int m = 0; m < NM; ++m)
#pragma unroll
    for (int j = 0; j < 14; ++j) v[m][j] = hb[m][r0 + j][q];
#pragma unroll
  for (int o = 0; o < 4; ++o) {
#pragma unroll
    for (int m = 0; m < NM; ++m) mu[m][o] = 0.f;
#pragma unroll
    for (int j = 0; j < 11; ++j)
#pragma unroll
      for (int m = 0; m < NM; ++m) mu[m][o] = fmaf(win.g[j], v[m][o + j], mu[m][o]);
  }
  __syncthreads();  // hb is reused by the next phase
}

// One 32x32 output tile per block of 256 threads; thread t owns column
// t % 32, rows 4 (t / 32) .. + 3.  The five blurred maps are formed in two
// phases through one three-map LDS buffer (31 KB per block in all, five
// blocks per CU).
__global__ __launch_bounds__(256) void k_ssim_fwd(int H, int W, const float* __restrict__ img,
                                                  const float* __restrict__ tgt, Win11 win,
                                                  float C1, float C2, float* __restrict__ abc,
                                                  double* __restrict__ sums) {
  __shared__ float sx[kSI][kSI + 1], sy[kSI][kSI + 1];
  __shared__ float hb[3][kSI][kST + 1];
  __shared__ float red[8];
  int bx, by, c;
  ssim_block(H, W, &bx, &by, &c);
  const size_t HW = (size_t)H * W;
  const float* X = img + c * HW;
  const float* Y = tgt + c * HW;
  const int x0 = bx * kST - kSR, y0 = by * kST - kSR;
  {
    const float* const planes[2] = {X, Y};
    float (*const dst[2])[kSI + 1] = {sx, sy};
    ssim_stage<2>(planes, H, W, x0, y0, dst);
  }
  __syncthreads();
  const int q = threadIdx.x % kST, r0 = 4 * (threadIdx.x / kST);
  float m3[3][4], m2[2][4];
  ssim_blur_phase<0>(sx, sy, hb, win, q, r0, m3);
  ssim_blur_phase<1>(sx, sy, hb, win, q, r0, m2);
  float s_sum = 0.f, l1_sum = 0.f;
  const int gx = x0 + kSR + q;
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    const int r = r0 + o, gy = y0 + kSR + r;
    if (gy >= H || gx >= W) continue;
    const float mx = m3[0][o], my = m3[1][o], exx = m3[2][o], eyy = m2[0][o], exy = m2[1][o];
    const float A1 = 2.f * mx * my + C1, A2 = 2.f * (exy - mx * my) + C2;
    const float B1 = mx * mx + my * my + C1, B2 = (exx - mx * mx) + (eyy - my * my) + C2;
    // D >= C1 (C2 - rounding) > 0 is normal: v_rcp_f32 (1 ulp) in place of a
    // correctly rounded division, within the loss's float64-reference bar
    const float D = B1 * B2, Dinv = __builtin_amdgcn_rcpf(D);
    const float sv = A1 * A2 * Dinv;
    const float dN = 2.f * my * (A2 - A1), dD = 2.f * mx * (B2 - B1);
    const size_t o_ = c * HW + (size_t)gy * W + gx;
    abc[o_] = (dN - sv * dD) * Dinv;     // ds/dmu_x
    abc[3 * HW + o_] = -sv * B1 * Dinv;  // ds/dE[x^2]
    abc[6 * HW + o_] = 2.f * A1 * Dinv;  // ds/dE[xy]
    s_sum += sv;
    l1_sum += fabsf(sx[r + kSR][q + kSR] - sy[r + kSR][q + kSR]);
  }
  const float bs = block_sum_f(s_sum, red);
  __syncthreads();
  const float bl = block_sum_f(l1_sum, red);
  if (threadIdx.x == 0) {
    sums[2 * (size_t)blockIdx.x] = (double)bs;
    sums[2 * (size_t)blockIdx.x + 1] = (double)bl;
  }
}

// The loss from the blocks' slot pairs: thread t adds the slots t, t + 256,
// ... in order (all its loads issued first), then a pairwise tree over the
// 256 partial sums, the same tree on every call.
__global__ __launch_bounds__(256) void k_ssim_finish(const double* __restrict__ sums, int nblocks, double n,
                                                     float lam, float* __restrict__ loss) {
  __shared__ double ps[256], pl[256];
  const int t = threadIdx.x;
  double s = 0.0, l = 0.0;
  int i = t;
  for (; i + 3 * 256 < nblocks; i += 4 * 256) {
    const double2 a = reinterpret_cast<const double2*>(sums)[i];
    const double2 b = reinterpret_cast<const double2*>(sums)[i + 256];
    const double2 c = reinterpret_cast<const double2*>(sums)[i + 512];
    const double2 d = reinterpret_cast<const double2*>(sums)[i + 768];
    s += a.x; l += a.y;
    s += b.x; l += b.y;
    s += c.x; l += c.y;
    s += d.x; l += d.y;
  }
  for (; i < nblocks; i += 256) {
    const double2 a = reinterpret_cast<const double2*>(sums)[i];
    s += a.x;
    l += a.y;
  }
  ps[t] = s;
  pl[t] = l;
  __syncthreads();
#pragma unroll
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      ps[t] += ps[t + h];
      pl[t] += pl[t + h];
    }
    __syncthreads();
  }
  if (t == 0) *loss = (float)((1.0 - lam) * (pl[0] / n) + lam * (1.0 - ps[0] / n));
}

__global__ __launch_bounds__(256) void k_ssim_bwd(int H, int W, const float* __restrict__ img,
                                                  const float* __restrict__ tgt,
                                                  const float* __restrict__ abc, Win11 win,
                                                  float lam, double n,
                                                  const float* __restrict__ gout,
                                                  float* __restrict__ dimg) {
  // The horizontal pass writes its sums over the staged maps they were read
  // from (every thread holds its inputs in registers across a barrier), so
  // the block needs 22 KB of LDS instead of 38: seven blocks per CU, not four.
  __shared__ float sa[3][kSI][kSI + 1];
  float (*const hb)[kSI][kST + 1] = reinterpret_cast<float (*)[kSI][kST + 1]>(&sa[0][0][0]);
  static_assert(3 * kSI * (kST + 1) <= 3 * kSI * (kSI + 1), "the sums fit in the staged maps' space");
  int bx, by, c;
  ssim_block(H, W, &bx, &by, &c);
  const size_t HW = (size_t)H * W;
  const int x0 = bx * kST - kSR, y0 = by * kST - kSR;
  {
    const float* const planes[3] = {abc + c * HW, abc + 3 * HW + c * HW, abc + 6 * HW + c * HW};
    float (*const dst[3])[kSI + 1] = {sa[0], sa[1], sa[2]};
    ssim_stage<3>(planes, H, W, x0, y0, dst);
  }
  __syncthreads();
  const bool hitem = threadIdx.x < kSI * kHG;
  const int hr = threadIdx.x % kSI, q0 = kHO * (threadIdx.x / kSI);
  float hv[3][kHO + 10];
  if (hitem) {
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int j = 0; j < kHO + 10; ++j) hv[m][j] = sa[m][hr][q0 + j];
  }
  __syncthreads();  // every staged value read before the sums overwrite them
  if (hitem) {
    const int r = hr;
#pragma unroll
    for (int o = 0; o < kHO; ++o) {
      float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int j = 0; j < 11; ++j) {
        const float g = win.g[j];
#pragma unroll
        for (int m = 0; m < 3; ++m) acc[m] = fmaf(g, hv[m][o + j], acc[m]);
      }
#pragma unroll
      for (int m = 0; m < 3; ++m) hb[m][r][q0 + o] = acc[m];
    }
  }
  __syncthreads();
  const float go = *gout;
  const float kl1 = (float)((1.0 - lam) / n) * go, kss = (float)(-lam / n) * go;
  for (int i = threadIdx.x; i < kSG * kST; i += blockDim.x) {
    const int q = i % kST, r0 = 4 * (i / kST);
    float v[3][14];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int j = 0; j < 14; ++j) v[m][j] = hb[m][r0 + j][q];
    const int gx = x0 + kSR + q;
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      const int gy = y0 + kSR + r0 + o;
      if (gy >= H || gx >= W) continue;
      float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int j = 0; j < 11; ++j) {
        const float g = win.g[j];
#pragma unroll
        for (int m = 0; m < 3; ++m) acc[m] = fmaf(g, v[m][o + j], acc[m]);
      }
      const size_t o_ = c * HW + (size_t)gy * W + gx;
      const float x = img[o_], y = tgt[o_];
      const float sg = x > y ? 1.f : (x < y ? -1.f : 0.f);
      dimg[o_] = kl1 * sg + kss * (acc[0] + 2.f * x * acc[1] + y * acc[2]);
    }
  }
}

}  // namespace rgp

using namespace rgp;

static Win11 make_win(const float* w) {
  Win11 v;
  for (int j = 0; j < 11; ++j) v.g[j] = w[j];
  return v;
}

extern "C" int64_t rgp_l1_ssim_sums(int channels, int height, int width) {
  if (channels <= 0 || height <= 0 || width <= 0) return 0;
  return 2 * (int64_t)((width + kST - 1) / kST) * ((height + kST - 1) / kST) * channels;
}

extern "C" int rgp_l1_ssim_forward(int channels, int height, int width, const float* img,
                                   const float* target, const float* window11, float C1, float C2,
                                   float lam, float* abc, double* sums, float* loss,
                                   hipStream_t s) {
  if (channels <= 0 || height <= 0 || width <= 0 || channels > 65535 || !window11)
    return set_error(RGP_ERR_ARG, "rgp_l1_ssim_forward: bad size or window");
  if ((int64_t)height * width > 0x7fffffff) return set_error(RGP_ERR_ARG, "rgp_l1_ssim_forward: image too large");
  const int64_t blocks = (int64_t)((width + kST - 1) / kST) * ((height + kST - 1) / kST) * channels;
  if (blocks > 0x7fffffff) return set_error(RGP_ERR_ARG, "rgp_l1_ssim_forward: image too large");
  if (((uintptr_t)sums & 15) != 0) return set_error(RGP_ERR_ARG, "rgp_l1_ssim_forward: sums must be 16-byte aligned");
  const dim3 grid((unsigned)blocks);
  hipLaunchKernelGGL(k_ssim_fwd, grid, dim3(256), 0, s, height, width, img, target,
                     make_win(window11), C1, C2, abc, sums);
  hipLaunchKernelGGL(k_ssim_finish, dim3(1), dim3(256), 0, s, (const double*)sums, (int)blocks,
                     (double)channels * height * width, lam, loss);
  return check_launch("rgp_l1_ssim_forward");
}

extern "C" int rgp_l1_ssim_backward(int channels, int height, int width, const float* img,
                                    const float* target, const float* abc, const float* window11,
                                    float lam, const float* grad_loss, float* dimg,
                                    hipStream_t s) {
  if (channels <= 0 || height <= 0 || width <= 0 || channels > 65535 || !window11)
    return set_error(RGP_ERR_ARG, "rgp_l1_ssim_backward: bad size or window");
  if ((int64_t)height * width > 0x7fffffff) return set_error(RGP_ERR_ARG, "rgp_l1_ssim_backward: image too large");
  const int64_t blocks = (int64_t)((width + kST - 1) / kST) * ((height + kST - 1) / kST) * channels;
  if (blocks > 0x7fffffff) return set_error(RGP_ERR_ARG, "rgp_l1_ssim_backward: image too large");
  const dim3 grid((unsigned)blocks);
  hipLaunchKernelGGL(k_ssim_bwd, grid, dim3(256), 0, s, height, width, img, target, abc,
                     make_win(window11), lam, (double)channels * height * width, grad_loss, dimg);
  return check_launch("rgp_l1_ssim_backward");
}
