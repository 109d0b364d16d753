// Per-Gaussian kernels of the palette-weight rasterizer for gfx950: the
// unfused gsplatcu operators (project, computeCov3D, computeCov2D,
// inverseCov2D, sh2Weight, sh2Weight_fast, sh2Weight_fast_B, sh2Color), the
// fused preprocess that replaces their chain in WeightGSFunctionTest/Fast
// (gsplat/gsmodel.py:107-118, 155-170), record packing, and the SH backward.
// One thread per Gaussian; every stage is HBM-bound.
#include <cmath>

#include "rgp_internal.h"

namespace rgp {

// ---------------------------------------------------------------------------
// unfused operators
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_project(int n, const float* __restrict__ pws,
                                                 const float* __restrict__ Rcw,
                                                 const float* __restrict__ tcw, float fx, float fy,
                                                 float cx, float cy, float* __restrict__ us,
                                                 float* __restrict__ pcs,
                                                 float* __restrict__ depths) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float R[9], t[3];
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = Rcw[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) t[k] = tcw[k];
  const Proj p = project_one(R, t, pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], fx, fy, cx, cy);
  if (!p.ok) { depths[i] = kBadMarker; return; }
  us[2 * i] = p.ux; us[2 * i + 1] = p.uy;
  pcs[3 * i] = p.x; pcs[3 * i + 1] = p.y; pcs[3 * i + 2] = p.z;
  depths[i] = p.z;
}

__global__ __launch_bounds__(256) void k_cov3d(int n, const float4* __restrict__ rots,
                                               const float* __restrict__ scales,
                                               const float* __restrict__ depths,
                                               float* __restrict__ cov3ds) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if ((double)depths[i] < kMinDepth) return;
  const float4 q = rots[i];
  float c[6];
  cov3d_one(q.x, q.y, q.z, q.w, scales[3 * i], scales[3 * i + 1], scales[3 * i + 2], c);
#pragma unroll
  for (int k = 0; k < 6; ++k) cov3ds[6 * i + k] = c[k];
}

__global__ __launch_bounds__(256) void k_cov2d(int n, const float* __restrict__ cov3ds,
                                               const float* __restrict__ pcs,
                                               const float* __restrict__ Rcw,
                                               const float* __restrict__ depths, float fx,
                                               float fy, float tfx, float tfy,
                                               float* __restrict__ cov2ds) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if ((double)depths[i] < kMinDepth) return;
  float R[9], c3[6], c2[3];
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = Rcw[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) c3[k] = cov3ds[6 * i + k];
  cov2d_one(c3, pcs[3 * i], pcs[3 * i + 1], pcs[3 * i + 2], R, fx, fy, tfx, tfy, c2);
#pragma unroll
  for (int k = 0; k < 3; ++k) cov2ds[3 * i + k] = c2[k];
}

__global__ __launch_bounds__(256) void k_inverse(int n, const float* __restrict__ cov2ds,
                                                 float* __restrict__ depths,
                                                 float* __restrict__ cinv,
                                                 int32_t* __restrict__ areas) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if ((double)depths[i] < kMinDepth) return;
  float ci[3];
  int ar[2];
  if (!inverse_one(cov2ds[3 * i], cov2ds[3 * i + 1], cov2ds[3 * i + 2], ci, ar)) {
    depths[i] = kBadMarker;
    return;
  }
  cinv[3 * i] = ci[0]; cinv[3 * i + 1] = ci[1]; cinv[3 * i + 2] = ci[2];
  areas[2 * i] = ar[0]; areas[2 * i + 1] = ar[1];
}

template <int K>
__global__ __launch_bounds__(256) void k_sh2weight(int n, int sh_dim, const float* __restrict__ shs,
                                                   const float* __restrict__ pws,
                                                   const float* __restrict__ twc,
                                                   float* __restrict__ weights,
                                                   float* __restrict__ basis) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float Y[16];
#pragma unroll
  for (int l = 0; l < 16; ++l) Y[l] = 0.f;
  const float tw[3] = {twc[0], twc[1], twc[2]};
  if (sh_dim > 1) sh_basis(pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], tw, sh_dim, Y);
  float w[K];
  sh_weights_full<K>(shs + (size_t)i * sh_dim * K, Y, sh_dim, w);
#pragma unroll
  for (int k = 0; k < K; ++k) weights[(size_t)i * K + k] = w[k];
  if (basis) {
    basis[(size_t)i * sh_dim] = (float)SH_C0_0;
    for (int l = 1; l < sh_dim; ++l) basis[(size_t)i * sh_dim + l] = Y[l];
  }
}

template <int K>
__global__ __launch_bounds__(256) void k_sh2weight_fast(int n, int P, int Q, int rank,
                                                        const float* __restrict__ dc,
                                                        const float* __restrict__ A,
                                                        const float* __restrict__ B,
                                                        const float* __restrict__ pws,
                                                        const float* __restrict__ twc,
                                                        float* __restrict__ weights,
                                                        float* __restrict__ dweight_ddc,
                                                        float* __restrict__ ylm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int L = P * Q / K;
  float Y[16];
#pragma unroll
  for (int l = 0; l < 16; ++l) Y[l] = 0.f;
  const float tw[3] = {twc[0], twc[1], twc[2]};
  if (L > 0) sh_basis(pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], tw, L + 1, Y);
  float w[K];
  sh_weights_lowrank<K>(dc + (size_t)i * K, A + (size_t)i * P * rank, B + (size_t)i * rank * Q, Y,
                        L, Q, rank, w);
#pragma unroll
  for (int k = 0; k < K; ++k) weights[(size_t)i * K + k] = w[k];
  if (dweight_ddc) {
#pragma unroll
    for (int k = 0; k < K; ++k) dweight_ddc[(size_t)i * K + k] = (float)SH_C0_0;
  }
  if (ylm)
    for (int l = 0; l < L; ++l) ylm[(size_t)i * L + l] = Y[l + 1];
}

// sh2Color's calc_J outputs (kernel.cu:1112-1167).  dcolor_dshs (N, sh_dim):
// SH_C0_0 and the basis values the forward multiplies the coefficients by.
// dcolor_dpws (N, 3, 3) = dc_dr (colour c, direction component r) times
// dr_dpw, accumulated from 0 (matrix.cuh:53-70).  dc_dr follows the float3
// arithmetic of common.cuh: a double constant times a float is formed in
// double and rounded to float when it meets a float3 (operator* takes
// const float&), the float3 products and sums are float, left to right.
// Its Y_9 term is C30 (3x^2 - 2y^2) (kernel.cu:1152), as in sh2Weight.
__global__ __launch_bounds__(256) void k_sh2color_jac(int n, int sh_dim, const float* __restrict__ shs,
                                                      const float* __restrict__ pws,
                                                      const float* __restrict__ twc,
                                                      float* __restrict__ dcolor_dshs,
                                                      float* __restrict__ dcolor_dpws) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* sh = shs + (size_t)i * sh_dim * 3;
  float Y[16];
#pragma unroll
  for (int l = 0; l < 16; ++l) Y[l] = 0.f;
  const float tw[3] = {twc[0], twc[1], twc[2]};
  if (sh_dim > 1) sh_basis(pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], tw, sh_dim, Y);
  dcolor_dshs[(size_t)i * sh_dim] = (float)SH_C0_0;
  for (int l = 1; l < sh_dim; ++l) dcolor_dshs[(size_t)i * sh_dim + l] = Y[l];
  float out[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) out[e] = 0.f;
  if (sh_dim > 1) {
    const float d0 = pws[3 * i] - tw[0], d1 = pws[3 * i + 1] - tw[1], d2 = pws[3 * i + 2] - tw[2];
    const float normd_inv = 1.f / sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
    const float x = d0 * normd_inv, y = d1 * normd_inv, z = d2 * normd_inv;
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
    const float n3 = (normd_inv * normd_inv) * normd_inv;
    const float r00 = ((-d0) * d0) * n3 + normd_inv, r11 = ((-d1) * d1) * n3 + normd_inv;
    const float r22 = ((-d2) * d2) * n3 + normd_inv, r01 = ((-d0) * d1) * n3;
    const float r02 = ((-d0) * d2) * n3, r12 = ((-d1) * d2) * n3;
    const float R[9] = {r00, r01, r02, r01, r11, r12, r02, r12, r22};
    auto F = [](double v) { return (float)v; };
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float* h = sh + c;  // coefficient l of channel c: h[3 * l]
      float dr0 = 0.f, dr1 = 0.f, dr2 = 0.f;
      dr0 = dr0 + h[9] * F(SH_C1_2);
      dr1 = dr1 + h[3] * F(SH_C1_0);
      dr2 = dr2 + h[6] * F(SH_C1_1);
      if (sh_dim > 4) {
        dr0 = dr0 + (((h[12] * F(SH_C2_0 * (double)y) - h[18] * F((SH_C2_2 * 2) * (double)x)) +
                      h[21] * F(SH_C2_3 * (double)z)) +
                     h[24] * F((SH_C2_4 * 2) * (double)x));
        dr1 = dr1 + (((h[12] * F(SH_C2_0 * (double)x) + h[15] * F(SH_C2_1 * (double)z)) -
                      h[18] * F((SH_C2_2 * 2.0) * (double)y)) -
                     h[24] * F((SH_C2_4 * 2) * (double)y));
        dr2 = dr2 + ((h[15] * F(SH_C2_1 * (double)y) + h[18] * F((SH_C2_2 * 4.0) * (double)z)) +
                     h[21] * F(SH_C2_3 * (double)x));
        if (sh_dim > 9) {
          const float a0 = ((h[27] * F(6.0 * SH_C3_0)) * x) * y;
          const float a1 = (h[30] * F(SH_C3_1)) * yz;
          const float a2 = (h[33] * F(2 * SH_C3_2)) * xy;
          const float a3 = (h[36] * F(6.0 * SH_C3_3)) * xz;
          const float a4 = (h[39] * F(SH_C3_4)) * ((4.f * zz - 3.f * xx) - yy);
          const float a5 = (h[42] * F(2 * SH_C3_5)) * xz;
          const float a6 = (h[45] * F(SH_C3_6)) * (3 * xx - 3 * yy);
          dr0 = dr0 + ((((((a0 + a1) - a2) - a3) + a4) + a5) + a6);
          const float b0 = (h[27] * F(SH_C3_0)) * (3.f * xx - 2.f * yy);
          const float b1 = (h[30] * F(SH_C3_1)) * xz;
          const float b2 = (h[33] * F(SH_C3_2)) * ((4.f * zz - xx) - 3.f * yy);
          const float b3 = (h[36] * F(6.0 * SH_C3_3)) * yz;
          const float b4 = (h[39] * F(SH_C3_4)) * (-2 * xy);
          const float b5 = (h[42] * F(2 * SH_C3_5)) * yz;
          const float b6 = (h[45] * F(6.0 * SH_C3_6)) * xy;
          dr1 = dr1 + ((((((b0 + b1) + b2) - b3) + b4) - b5) - b6);
          const float c0 = (h[30] * F(SH_C3_1)) * xy;
          const float c1 = (h[33] * F(8.0 * SH_C3_2)) * yz;
          const float c2 = (h[36] * F(SH_C3_3)) * ((6.f * zz - 3.f * xx) - 3.f * yy);
          const float c3 = (h[39] * F(8.0 * SH_C3_4)) * xz;
          const float c4 = (h[42] * F(SH_C3_5)) * (xx - yy);
          dr2 = dr2 + ((((c0 + c1) + c2) + c3) + c4);
        }
      }
      const float D[3] = {dr0, dr1, dr2};
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) acc = acc + D[k] * R[3 * k + j];
        out[3 * c + j] = acc;
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) dcolor_dpws[(size_t)i * 9 + e] = out[e];
}

__global__ __launch_bounds__(256) void k_sh2color(int n, int sh_dim, const float* __restrict__ shs,
                                                  const float* __restrict__ pws,
                                                  const float* __restrict__ twc,
                                                  float* __restrict__ colors) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* sh = shs + (size_t)i * sh_dim * 3;
  const float c0 = (float)SH_C0_0;
  float col[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) col[c] = c0 * sh[c];
  if (sh_dim > 1) {
    float Y[16];
    const float tw[3] = {twc[0], twc[1], twc[2]};
    sh_basis(pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], tw, sh_dim, Y);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      col[c] = col[c] + ((Y[1] * sh[3 + c] + Y[2] * sh[6 + c]) + Y[3] * sh[9 + c]);
    if (sh_dim > 4) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
        col[c] = col[c] + ((((Y[4] * sh[12 + c] + Y[5] * sh[15 + c]) + Y[6] * sh[18 + c]) +
                            Y[7] * sh[21 + c]) +
                           Y[8] * sh[24 + c]);
      if (sh_dim > 9) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          col[c] = col[c] + ((((((Y[9] * sh[27 + c] + Y[10] * sh[30 + c]) + Y[11] * sh[33 + c]) +
                                Y[12] * sh[36 + c]) +
                               Y[13] * sh[39 + c]) +
                              Y[14] * sh[42 + c]) +
                             Y[15] * sh[45 + c]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) colors[(size_t)i * 3 + c] = col[c];
}

// Low-rank factor backward (kernel.cu:1019-1031, rank r).  The reference's
// atomics hit rows private to the thread, so its sums run in loop order (k
// outer, lm inner); this kernel keeps that order with plain read-modify-write.
// ylm may be NULL: Y is then recomputed from pws and twc.
__global__ __launch_bounds__(256) void k_lowrank_backward(
    int n, int K, int P, int Q, int rank, const float* __restrict__ A, const float* __restrict__ B,
    const float* __restrict__ ylm, const float* __restrict__ pws, const float* __restrict__ twc,
    const float* __restrict__ dw, float* __restrict__ ddc, float* __restrict__ dA,
    float* __restrict__ dB) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int L = P * Q / K;
  float Y[16];
#pragma unroll
  for (int l = 0; l < 16; ++l) Y[l] = 0.f;
  if (ylm) {
    for (int l = 0; l < L; ++l) Y[l + 1] = ylm[(size_t)i * L + l];
  } else if (L > 0) {
    const float tw[3] = {twc[0], twc[1], twc[2]};
    sh_basis(pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], tw, L + 1, Y);
  }
  const float* a = A + (size_t)i * P * rank;
  const float* b = B + (size_t)i * rank * Q;
  float* da = dA + (size_t)i * P * rank;
  float* db = dB + (size_t)i * rank * Q;
  for (int j = 0; j < P * rank; ++j) da[j] = 0.f;
  for (int j = 0; j < rank * Q; ++j) db[j] = 0.f;
  for (int k = 0; k < K; ++k) {
    const float dwk = dw[(size_t)i * K + k];
    if (ddc) ddc[(size_t)i * K + k] = dwk * (float)SH_C0_0;
    int row = k / Q, col = k - (k / Q) * Q;
    for (int lm = 0; lm < L; ++lm) {
      const float y = Y[lm + 1];
      for (int r = 0; r < rank; ++r) {
        da[row * rank + r] = da[row * rank + r] + dwk * b[r * Q + col] * y;
        db[r * Q + col] = db[r * Q + col] + dwk * a[row * rank + r] * y;
      }
      col += K;
      while (col >= Q) { col -= Q; ++row; }
    }
  }
}

// The same with Q = 15, P = K and rank R known at compile time: factors and
// gradient accumulators in registers, one store per output.  Same operations
// in the same order.
// dL/d(dc, A, B) of one Gaussian for Q = 15, P = K: Y[1..15] its basis, a
// (P x R) and b (R x Q) its factors, dw its K weight gradients.
template <int K, int R>
__device__ __forceinline__ void lowrank_q15_grad(const float* Y, const float* a, const float* b,
                                                 const float* dw, float* ddc, float* da, float* db) {
  constexpr int Q = 15, P = K;
#pragma unroll
  for (int j = 0; j < P * R; ++j) da[j] = 0.f;
#pragma unroll
  for (int j = 0; j < R * Q; ++j) db[j] = 0.f;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float dwk = dw[k];
    ddc[k] = dwk * (float)SH_C0_0;
#pragma unroll
    for (int lm = 0; lm < 15; ++lm) {
      const int g = lm * K + k, row = g / Q, col = g % Q;
      const float y = Y[lm + 1];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        da[row * R + r] = da[row * R + r] + dwk * b[r * Q + col] * y;
        db[r * Q + col] = db[r * Q + col] + dwk * a[row * R + r] * y;
      }
    }
  }
}

template <int K, int R>
__global__ __launch_bounds__(256) void k_lowrank_backward_q15(
    int n, const float* __restrict__ A, const float* __restrict__ B,
    const float* __restrict__ ylm, const float* __restrict__ pws, const float* __restrict__ twc,
    const float* __restrict__ dw, float* __restrict__ ddc, float* __restrict__ dA,
    float* __restrict__ dB) {
  constexpr int Q = 15, P = K;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float Y[16];
  if (ylm) {
#pragma unroll
    for (int l = 0; l < 15; ++l) Y[l + 1] = ylm[(size_t)i * 15 + l];
  } else {
    const float tw[3] = {twc[0], twc[1], twc[2]};
    sh_basis(pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], tw, 16, Y);
  }
  float a[P * R], b[R * Q], dwr[K], gdc[K], da[P * R], db[R * Q];
#pragma unroll
  for (int j = 0; j < P * R; ++j) a[j] = A[(size_t)i * P * R + j];
#pragma unroll
  for (int j = 0; j < R * Q; ++j) b[j] = B[(size_t)i * R * Q + j];
#pragma unroll
  for (int k = 0; k < K; ++k) dwr[k] = dw[(size_t)i * K + k];
  lowrank_q15_grad<K, R>(Y, a, b, dwr, gdc, da, db);
  if (ddc) {
#pragma unroll
    for (int k = 0; k < K; ++k) ddc[(size_t)i * K + k] = gdc[k];
  }
#pragma unroll
  for (int j = 0; j < P * R; ++j) dA[(size_t)i * P * R + j] = da[j];
#pragma unroll
  for (int j = 0; j < R * Q; ++j) dB[(size_t)i * R * Q + j] = db[j];
}

// Multi-view forms: the gradient of the factors is the sum over `views`
// cameras of the single-view chain, each with its own basis Y(pw, twc_v) and
// weight gradient dw + v*n*K.  Sums run view by view into the same
// accumulators (dL/ddc: sum_v dw_v * SH_C0).
template <int K, int R>
__global__ __launch_bounds__(256) void k_lowrank_backward_q15_views(
    int views, int n, const float* __restrict__ A, const float* __restrict__ B, const float* __restrict__ pws,
    const rgp_camera* __restrict__ cams, const float* __restrict__ dw, float* __restrict__ ddc,
    float* __restrict__ dA, float* __restrict__ dB) {
  constexpr int Q = 15, P = K;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float a[P * R], b[R * Q], da[P * R], db[R * Q], dd[K];
#pragma unroll
  for (int j = 0; j < P * R; ++j) { a[j] = A[(size_t)i * P * R + j]; da[j] = 0.f; }
#pragma unroll
  for (int j = 0; j < R * Q; ++j) { b[j] = B[(size_t)i * R * Q + j]; db[j] = 0.f; }
#pragma unroll
  for (int k = 0; k < K; ++k) dd[k] = 0.f;
  const float px = pws[3 * i], py = pws[3 * i + 1], pz = pws[3 * i + 2];
  for (int v = 0; v < views; ++v) {
    float Y[16];
    sh_basis(px, py, pz, cams[v].twc, 16, Y);
    const float* dwv = dw + ((size_t)v * n + i) * K;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const float dwk = dwv[k];
      dd[k] = dd[k] + dwk * (float)SH_C0_0;
#pragma unroll
      for (int lm = 0; lm < 15; ++lm) {
        const int g = lm * K + k, row = g / Q, col = g % Q;
        const float y = Y[lm + 1];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          da[row * R + r] = da[row * R + r] + dwk * b[r * Q + col] * y;
          db[r * Q + col] = db[r * Q + col] + dwk * a[row * R + r] * y;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) ddc[(size_t)i * K + k] = dd[k];
#pragma unroll
  for (int j = 0; j < P * R; ++j) dA[(size_t)i * P * R + j] = da[j];
#pragma unroll
  for (int j = 0; j < R * Q; ++j) dB[(size_t)i * R * Q + j] = db[j];
}

__global__ __launch_bounds__(256) void k_lowrank_backward_views(
    int views, int n, int K, int P, int Q, int rank, const float* __restrict__ A, const float* __restrict__ B,
    const float* __restrict__ pws, const rgp_camera* __restrict__ cams, const float* __restrict__ dw,
    float* __restrict__ ddc, float* __restrict__ dA, float* __restrict__ dB) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int L = P * Q / K;
  const float* a = A + (size_t)i * P * rank;
  const float* b = B + (size_t)i * rank * Q;
  float* da = dA + (size_t)i * P * rank;
  float* db = dB + (size_t)i * rank * Q;
  for (int j = 0; j < P * rank; ++j) da[j] = 0.f;
  for (int j = 0; j < rank * Q; ++j) db[j] = 0.f;
  for (int k = 0; k < K; ++k) ddc[(size_t)i * K + k] = 0.f;
  for (int v = 0; v < views; ++v) {
    float Y[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) Y[l] = 0.f;
    if (L > 0) sh_basis(pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], cams[v].twc, L + 1, Y);
    const float* dwv = dw + ((size_t)v * n + i) * K;
    for (int k = 0; k < K; ++k) {
      const float dwk = dwv[k];
      ddc[(size_t)i * K + k] = ddc[(size_t)i * K + k] + dwk * (float)SH_C0_0;
      int row = k / Q, col = k - (k / Q) * Q;
      for (int lm = 0; lm < L; ++lm) {
        const float y = Y[lm + 1];
        for (int r = 0; r < rank; ++r) {
          da[row * rank + r] = da[row * rank + r] + dwk * b[r * Q + col] * y;
          db[r * Q + col] = db[r * Q + col] + dwk * a[row * rank + r] * y;
        }
        col += K;
        while (col >= Q) { col -= Q; ++row; }
      }
    }
  }
}

// Launch the low-rank backward, picking the register form when it applies.
static void launch_lowrank_backward(int n, int K, int P, int Q, int rank, const float* A,
                                    const float* B, const float* ylm, const float* pws,
                                    const float* twc, const float* dw, float* ddc, float* dA,
                                    float* dB, hipStream_t s);

// ---------------------------------------------------------------------------
// records
// ---------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ void store_record(float4* __restrict__ rec, float ux, float uy,
                                             const float* ci, float alpha, const float* f) {
  constexpr int F4 = RecStride<DIM>::f4;
  rec[0] = make_float4(ux, uy, ci[0], 2.f * ci[1]);
  const float mc = mcut_of(alpha);
  rec[1] = make_float4(ci[2], alpha, mc, ry_of(mc, ci[0], ci[1], ci[2]));
#pragma unroll
  for (int q = 0; q < F4 - 2; ++q) {
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = (4 * q + e < DIM) ? f[4 * q + e] : 0.f;
    rec[2 + q] = make_float4(v[0], v[1], v[2], v[3]);
  }
}

template <int DIM>
__global__ __launch_bounds__(256) void k_pack(int n, const float* __restrict__ us,
                                              const float* __restrict__ cinv,
                                              const float* __restrict__ alphas,
                                              const float* __restrict__ feat,
                                              float4* __restrict__ recs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) f[k] = feat[(size_t)i * DIM + k];
  const float ci[3] = {cinv[3 * i], cinv[3 * i + 1], cinv[3 * i + 2]};
  store_record<DIM>(recs + (size_t)i * RecStride<DIM>::f4, us[2 * i], us[2 * i + 1], ci, alphas[i],
                    f);
}

// getRects (kernel.cu:127-162): returns the tile count (0 = culled) and the
// packed rect (x0 | y0 << 16, x1 | y1 << 16).
__device__ __forceinline__ uint32_t rect_of(float ux, float uy, int ax, int ay, int gx, int gy,
                                            uint2* rect) {
  const float xs = (float)ax, ys = (float)ay;
  auto cl = [](int32_t v, int g) -> uint32_t {
    const int32_t m = v > 0 ? v : 0;
    return (uint32_t)m < (uint32_t)g ? (uint32_t)m : (uint32_t)g;
  };
  const uint32_t x0 = cl(sat_i32((ux - xs) / 16.f), gx);
  const uint32_t y0 = cl(sat_i32((uy - ys) / 16.f), gy);
  const uint32_t x1 = cl(sat_i32((((ux + xs) + 16.f) - 1.f) / 16.f), gx);
  const uint32_t y1 = cl(sat_i32((((uy + ys) + 16.f) - 1.f) / 16.f), gy);
  rect->x = x0 | (y0 << 16);
  rect->y = x1 | (y1 << 16);
  return (y1 - y0) * (x1 - x0);
}

__global__ __launch_bounds__(256) void k_rects(int n, const float* __restrict__ us,
                                               int32_t* __restrict__ areas,
                                               float* __restrict__ depths, int gx, int gy,
                                               uint2* __restrict__ rects,
                                               uint32_t* __restrict__ counts,
                                               uint32_t* __restrict__ dqw,
                                               uint32_t* __restrict__ dqs, uint32_t* __restrict__ clr,
                                               int clr_words) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0)  // the binning header (and the super-tile totals), cleared for the scan
    for (int q = threadIdx.x; q < clr_words; q += blockDim.x) clr[q] = 0u;
  uint32_t dq = 0;
  if (i < n) {
    uint32_t cnt = 0;
    const float depth = depths[i];
    if (!((double)depth < kMinDepth)) {
      uint2 r;
      cnt = rect_of(us[2 * i], us[2 * i + 1], areas[2 * i], areas[2 * i + 1], gx, gy, &r);
      if (cnt == 0) {
        depths[i] = kBadMarker;
        areas[2 * i] = 0; areas[2 * i + 1] = 0;
      } else {
        rects[i] = r;
        dq = sat_u32(depth * 1000.f);
      }
    }
    counts[i] = cnt;
    dqs[i] = dq;
  }
  // one plain store per group of 64 Gaussians; the scan reduces them
  dq = wave_max_u32(dq);
  if ((threadIdx.x & 63) == 0 && i < n) dqw[i >> 6] = dq;
}

// ---------------------------------------------------------------------------
// fused preprocess (WeightGSFunctionTest/Fast forward up to getRects)
// ---------------------------------------------------------------------------
// WMODE: 0 full rank; 1 low rank, generic P, Q, rank; 2 / 3 low rank with
// Q = 15, P = K and rank 1 / 2 (factors in registers); 4 full rank with 16
// coefficients per channel and 16-byte aligned rows, staged through LDS.
// One 64-thread block per 64 Gaussians; the block's quantised-depth maximum
// goes to dqw[block] with a plain store (a same-address atomic per wavefront
// cost more than the rest of the kernel).
//
// Staging (WMODE 4): the block's 64 rows of 4K float4 are one contiguous
// span, read with 4K coalesced 16-byte loads per lane (lanes reading their
// own rows, 64*K bytes apart, touch 64 cache lines per load instead of 8).  Float4
// q of row g is kept at LDS slot g*4K + (q ^ sw(g)); the swizzle permutes
// aligned groups of four slots, so the stores stay contiguous per 16 lanes
// and, for odd K, the row reads of 8 or 16 lanes hit distinct banks.
__device__ __forceinline__ int sh_swizzle(int g) { return ((g >> 1) ^ (g >> 3)) & 3; }

// Stages the block's 64 coefficient rows (WMODE 4) in LDS.
template <int K, int WMODE>
__device__ __forceinline__ void stage_sh_rows(float4* s_sh, int n, const float* __restrict__ shs) {
  constexpr int R4 = 4 * K;  // float4 per staged row
  const int t = threadIdx.x;
  const int tot = min(64, n - (int)blockIdx.x * 64) * R4;
  const float4* src = reinterpret_cast<const float4*>(shs) + (size_t)blockIdx.x * 64 * R4;
  float4 v[R4];
  if (tot == 64 * R4) {
#pragma clang loop unroll(full)
    for (int m = 0; m < R4; ++m) v[m] = src[t + 64 * m];
#pragma clang loop unroll(full)
    for (int m = 0; m < R4; ++m) {
      const int e = t + 64 * m;
      s_sh[e ^ sh_swizzle(e / R4)] = v[m];
    }
  } else {
    for (int e = t; e < tot; e += 64) s_sh[e ^ sh_swizzle(e / R4)] = src[e];
  }
  __syncthreads();
}

// Stages the block's 64 rows of the higher bands (15*K floats each, WMODE 5:
// the reference's second coefficient leaf) in LDS as one contiguous span.
// Row g's floats sit at s[g*15K ..]: with 15K odd the row reads of a
// wavefront fall in distinct banks.
template <int K>
__device__ __forceinline__ void stage_high_rows(float* s_hi, int n, const float* __restrict__ high) {
  constexpr int ROW = 15 * K;
  const int t = threadIdx.x;
  const int tot = min(64, n - (int)blockIdx.x * 64) * ROW;
  const float* src = high + (size_t)blockIdx.x * 64 * ROW;
  if (tot == 64 * ROW) {
    // a full block: 16*ROW float4, every load issued before the first store
    // (a load-store loop waits out one HBM latency per iteration)
    constexpr int F4 = 16 * ROW, PER = (F4 + 63) / 64;
    const float4* src4 = reinterpret_cast<const float4*>(src);
    const bool tail = t + 64 * (PER - 1) < F4;
    float4 v[PER - 1], last = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int m = 0; m < PER - 1; ++m) v[m] = src4[t + 64 * m];
    if (tail) last = src4[t + 64 * (PER - 1)];
#pragma unroll
    for (int m = 0; m < PER - 1; ++m) reinterpret_cast<float4*>(s_hi)[t + 64 * m] = v[m];
    if (tail) reinterpret_cast<float4*>(s_hi)[t + 64 * (PER - 1)] = last;
  } else {
    const int t4 = tot >> 2;
    for (int e = t; e < t4; e += 64) reinterpret_cast<float4*>(s_hi)[e] = reinterpret_cast<const float4*>(src)[e];
    for (int e = 4 * t4 + t; e < tot; e += 64) s_hi[e] = src[e];
  }
  __syncthreads();
}

// The whole per-Gaussian chain of one view for Gaussian i (lane t of its
// block); returns uint32(depth * 1000) of a binned Gaussian, else 0.
template <int K, int WMODE>
__device__ __forceinline__ uint32_t preprocess_gaussian(
    int i, int t, const float4* s_sh, int sh_dim, int P, int Q, int rank, const rgp_camera* __restrict__ cam,
    const float* __restrict__ pws, const float4* __restrict__ rots, const float* __restrict__ scales,
    const float* __restrict__ alphas, const float* __restrict__ shs, const float* __restrict__ dc,
    const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ depths,
    float4* __restrict__ recs, uint2* __restrict__ rects, uint32_t* __restrict__ counts, int gx, int gy,
    float* __restrict__ us_out, float* __restrict__ cinv_out, int32_t* __restrict__ areas_out,
    float* __restrict__ w_out) {
  constexpr int R = WMODE == 2 ? 1 : 2;
  constexpr int R4 = 4 * K;
  uint32_t dqv = 0;
  const float px = pws[3 * i], py = pws[3 * i + 1], pz = pws[3 * i + 2];
  const Proj p = project_one(cam->Rcw, cam->tcw, px, py, pz, cam->fx, cam->fy, cam->cx, cam->cy);
  // palette weights are evaluated for every Gaussian (kernel.cu:670-767
  // has no depth test)
  float Y[16];
#pragma unroll
  for (int l = 0; l < 16; ++l) Y[l] = 0.f;
  float w[K];
  if (WMODE == 0) {
    if (sh_dim > 1) sh_basis(px, py, pz, cam->twc, sh_dim, Y);
    sh_weights_full<K>(shs + (size_t)i * sh_dim * K, Y, sh_dim, w);
  } else if (WMODE == 4) {
    sh_basis(px, py, pz, cam->twc, 16, Y);
    float row[16 * K];
    const int sw = sh_swizzle(t);
#pragma unroll
    for (int q = 0; q < R4; ++q) {
      const float4 c = s_sh[t * R4 + (q ^ sw)];
      row[4 * q] = c.x; row[4 * q + 1] = c.y; row[4 * q + 2] = c.z; row[4 * q + 3] = c.w;
    }
    sh_weights_full<K>(row, Y, 16, w);
  } else if (WMODE == 5 || WMODE == 6) {
    // the reference's two leaves, low (N,K) in dc and the higher bands
    // (N,(sh_dim-1)K) in shs: the row they concatenate to (gsplat/utils.py:157)
    // Both leaves are read where sh_weights_full uses them, at constant
    // offsets (a local copy of the 16K-float row would live in scratch).
    if (sh_dim > 1) sh_basis(px, py, pz, cam->twc, sh_dim, Y);
    float lo[K];
#pragma unroll
    for (int k = 0; k < K; ++k) lo[k] = dc[(size_t)i * K + k];
    const float* hi = WMODE == 5 ? reinterpret_cast<const float*>(s_sh) + t * 15 * K
                                 : shs + (size_t)i * (sh_dim - 1) * K;
    sh_weights_full<K>(SplitRow<K>{lo, hi}, Y, sh_dim, w);
  } else if (WMODE == 1) {
    const int L = P * Q / K;
    if (L > 0) sh_basis(px, py, pz, cam->twc, L + 1, Y);
    sh_weights_lowrank<K>(dc + (size_t)i * K, A + (size_t)i * P * rank,
                          B + (size_t)i * rank * Q, Y, L, Q, rank, w);
  } else {
    sh_basis(px, py, pz, cam->twc, 16, Y);
    sh_weights_lowrank_q15<K, R>(dc + (size_t)i * K, A + (size_t)i * K * R, B + (size_t)i * R * 15, Y,
                                 w);
  }
  if (w_out) {
#pragma unroll
    for (int k = 0; k < K; ++k) w_out[(size_t)i * K + k] = w[k];
  }
  float depth = kBadMarker;
  uint32_t cnt = 0;
  if (p.ok) {
    depth = p.z;
    if (us_out) { us_out[2 * i] = p.ux; us_out[2 * i + 1] = p.uy; }
    const float4 q = rots[i];
    float c3[6], c2[3], ci[3];
    int ar[2];
    cov3d_one(q.x, q.y, q.z, q.w, scales[3 * i], scales[3 * i + 1], scales[3 * i + 2], c3);
    cov2d_one(c3, p.x, p.y, p.z, cam->Rcw, cam->fx, cam->fy, cam->tan_fovx, cam->tan_fovy, c2);
    if (!inverse_one(c2[0], c2[1], c2[2], ci, ar)) {
      depth = kBadMarker;
    } else {
      if (cinv_out) { cinv_out[3 * i] = ci[0]; cinv_out[3 * i + 1] = ci[1]; cinv_out[3 * i + 2] = ci[2]; }
      uint2 r;
      cnt = rect_of(p.ux, p.uy, ar[0], ar[1], gx, gy, &r);
      if (cnt == 0) {
        depth = kBadMarker;
        ar[0] = 0; ar[1] = 0;
      } else {
        rects[i] = r;
        dqv = sat_u32(depth * 1000.f);
        store_record<K>(recs + (size_t)i * RecStride<K>::f4, p.ux, p.uy, ci, alphas[i], w);
      }
      if (areas_out) { areas_out[2 * i] = ar[0]; areas_out[2 * i + 1] = ar[1]; }
    }
  }
  depths[i] = depth;
  counts[i] = cnt;
  return dqv;
}

template <int K, int WMODE>
__global__ __launch_bounds__(64) void k_preprocess(
    int n, int sh_dim, int P, int Q, int rank, const rgp_camera* __restrict__ cam,
    const float* __restrict__ pws, const float4* __restrict__ rots,
    const float* __restrict__ scales, const float* __restrict__ alphas,
    const float* __restrict__ shs, const float* __restrict__ dc, const float* __restrict__ A,
    const float* __restrict__ B, float* __restrict__ depths, float4* __restrict__ recs,
    uint2* __restrict__ rects, uint32_t* __restrict__ counts, uint32_t* __restrict__ dqw,
    uint32_t* __restrict__ dqs, uint32_t* __restrict__ clr, int clr_words, int gx, int gy,
    float* __restrict__ us_out, float* __restrict__ cinv_out, int32_t* __restrict__ areas_out,
    float* __restrict__ w_out) {
  __shared__ float4 s_sh[WMODE == 4 ? 64 * 4 * K : WMODE == 5 ? (64 * 15 * K + 3) / 4 : 1];
  const int t = threadIdx.x;
  const int i = blockIdx.x * 64 + t;
  if (blockIdx.x == 0)  // the binning header (and the super-tile totals), cleared for the scan
    for (int q = t; q < clr_words; q += 64) clr[q] = 0u;
  if constexpr (WMODE == 4) stage_sh_rows<K, WMODE>(s_sh, n, shs);
  if constexpr (WMODE == 5) stage_high_rows<K>(reinterpret_cast<float*>(s_sh), n, shs);
  uint32_t dqv = 0;
  if (i < n) {
    dqv = preprocess_gaussian<K, WMODE>(i, t, s_sh, sh_dim, P, Q, rank, cam, pws, rots, scales, alphas, shs, dc,
                                        A, B, depths, recs, rects, counts, gx, gy, us_out, cinv_out, areas_out,
                                        w_out);
    dqs[i] = dqv;
  }
  dqv = wave_max_u32(dqv);
  if (t == 0) dqw[blockIdx.x] = dqv;
}

// The same chain for `views` cameras of one resolution in one pass: the
// block's parameters come from HBM once (the coefficient rows staged in LDS
// once; the other per-Gaussian loads of later views hit the caches), and
// view v writes depths + v*n, recs + v*n*F4 and the binning regions of the
// workspace at ws + v*ws_stride.
template <int K, int WMODE>
__global__ __launch_bounds__(64) void k_preprocess_views(
    int views, int n, int sh_dim, int P, int Q, int rank, const rgp_camera* __restrict__ cams,
    const float* __restrict__ pws, const float4* __restrict__ rots, const float* __restrict__ scales,
    const float* __restrict__ alphas, const float* __restrict__ shs, const float* __restrict__ dc,
    const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ depths,
    float4* __restrict__ recs, char* __restrict__ ws, size_t ws_stride, size_t off_rects, size_t off_counts,
    size_t off_dqw, size_t off_dqs, int clr_words, int gx, int gy) {
  __shared__ float4 s_sh[WMODE == 4 ? 64 * 4 * K : WMODE == 5 ? (64 * 15 * K + 3) / 4 : 1];
  const int t = threadIdx.x;
  const int i = blockIdx.x * 64 + t;
  if (blockIdx.x == 0)  // every view's binning header (and super-tile totals), cleared for the scan
    for (int v = 0; v < views; ++v)
      for (int q = t; q < clr_words; q += 64) ((uint32_t*)(ws + (size_t)v * ws_stride))[q] = 0u;
  if constexpr (WMODE == 4) stage_sh_rows<K, WMODE>(s_sh, n, shs);
  if constexpr (WMODE == 5) stage_high_rows<K>(reinterpret_cast<float*>(s_sh), n, shs);
  for (int v = 0; v < views; ++v) {
    char* w = ws + (size_t)v * ws_stride;
    uint32_t dqv = 0;
    if (i < n) {
      dqv = preprocess_gaussian<K, WMODE>(i, t, s_sh, sh_dim, P, Q, rank, cams + v, pws, rots, scales, alphas,
                                          shs, dc, A, B, depths + (size_t)v * n,
                                          recs + (size_t)v * n * RecStride<K>::f4, (uint2*)(w + off_rects),
                                          (uint32_t*)(w + off_counts), gx, gy, nullptr, nullptr, nullptr,
                                          nullptr);
      ((uint32_t*)(w + off_dqs))[i] = dqv;
    }
    dqv = wave_max_u32(dqv);
    if (t == 0) ((uint32_t*)(w + off_dqw))[blockIdx.x] = dqv;
  }
}

// ---------------------------------------------------------------------------
// full-rank SH backward: dshs[n, lm*K + k] = dw[n,k] * Y_lm(n)
// (gsmodel.py:140-143).  A block owns 128 Gaussians; their output rows are
// one contiguous span written with coalesced stores from LDS-held factors.
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(256) void k_sh_backward_full(int n, int sh_dim,
                                                          const rgp_camera* __restrict__ cam,
                                                          const float* __restrict__ pws,
                                                          const float* __restrict__ dw,
                                                          float* __restrict__ dshs) {
  constexpr int G = 128;
  __shared__ float sY[G][17];
  __shared__ float sW[G][K];
  const int g0 = blockIdx.x * G;
  const int t = threadIdx.x;
  if (t < G) {
    const int i = g0 + t;
    float Y[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) Y[l] = 0.f;
    if (i < n) {
      if (sh_dim > 1) sh_basis(pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], cam->twc, sh_dim, Y);
#pragma unroll
      for (int k = 0; k < K; ++k) sW[t][k] = dw[(size_t)i * K + k];
    }
    Y[0] = (float)SH_C0_0;
#pragma unroll
    for (int l = 0; l < 16; ++l) sY[t][l] = Y[l];
  }
  __syncthreads();
  const int cnt = min(G, n - g0);
  const int row = sh_dim * K;
  const int64_t total = (int64_t)cnt * row;
  float* out = dshs + (size_t)g0 * row;
  for (int64_t e = t; e < total; e += blockDim.x) {
    const int gl = (int)(e / row);
    const int r = (int)(e - (int64_t)gl * row);
    const int lm = r / K, k = r - lm * K;
    out[e] = sW[gl][k] * sY[gl][lm];
  }
}

// dshs[n, lm*K + k] = sum_v dw_v[n,k] * Y_lm(pw, twc_v) over `views`
// cameras (view by view), with dw_v at dw + v*n*K.
template <int K>
__global__ __launch_bounds__(256) void k_sh_backward_full_views(int views, int n, int sh_dim,
                                                                const rgp_camera* __restrict__ cams,
                                                                const float* __restrict__ pws,
                                                                const float* __restrict__ dw,
                                                                float* __restrict__ dshs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int row = sh_dim * K;
  float acc[16 * K];
#pragma unroll
  for (int e = 0; e < 16 * K; ++e) acc[e] = 0.f;
  const float px = pws[3 * i], py = pws[3 * i + 1], pz = pws[3 * i + 2];
  for (int v = 0; v < views; ++v) {
    float Y[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) Y[l] = 0.f;
    if (sh_dim > 1) sh_basis(px, py, pz, cams[v].twc, sh_dim, Y);
    Y[0] = (float)SH_C0_0;
    const float* dwv = dw + ((size_t)v * n + i) * K;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const float d = dwv[k];
#pragma unroll
      for (int lm = 0; lm < 16; ++lm) acc[lm * K + k] = acc[lm * K + k] + d * Y[lm];
    }
  }
  float* out = dshs + (size_t)i * row;
#pragma unroll
  for (int e = 0; e < 16 * K; ++e)
    if (e < row) out[e] = acc[e];
}

// ---------------------------------------------------------------------------
// calc_J outputs of the unfused operators (kernel.cu:402-414, 463-500,
// 557-607, 648-660, 769-869).  Small matrix products accumulate from 0 over
// k ascending with every term (Matrix::operator*, matrix.cuh:53-70); other
// expressions are evaluated left to right as the reference writes them.
// Entries the reference does not write are left to the caller's zero fill.
// ---------------------------------------------------------------------------
template <int I, int K, int J>
__device__ __forceinline__ void small_matmul(const float* A, const float* B, float* C) {
#pragma unroll
  for (int i = 0; i < I; ++i)
#pragma unroll
    for (int j = 0; j < J; ++j) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < K; ++k) acc = acc + A[i * K + k] * B[k * J + j];
      C[i * J + j] = acc;
    }
}

__global__ __launch_bounds__(256) void k_project_jac(int n, const float* __restrict__ pws,
                                                     const float* __restrict__ Rcw,
                                                     const float* __restrict__ tcw, float fx,
                                                     float fy, float* __restrict__ du) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float R[9], t[3];
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = Rcw[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) t[k] = tcw[k];
  const Proj p = project_one(R, t, pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], fx, fy, 0.f, 0.f);
  if (!p.ok) return;
  const float z_inv = 1.f / p.z;
  const float z2_inv = z_inv * z_inv;
  float* o = du + (size_t)i * 6;
  o[0] = fx * z_inv;
  o[2] = -(p.x * fx) * z2_inv;
  o[4] = fy * z_inv;
  o[5] = -(p.y * fy) * z2_inv;
}

// dcov3d/drot (N,6,4), dcov3d/dscale (N,6,3): dSigma/dM (6x9) times dM/dq
// and dM/ds, M = R(q) diag(s) with the unnormalised quaternion (w, x, y, z).
__global__ __launch_bounds__(256) void k_cov3d_jac(int n, const float4* __restrict__ rots,
                                                   const float* __restrict__ scales,
                                                   const float* __restrict__ depths,
                                                   float* __restrict__ drot,
                                                   float* __restrict__ dscale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if ((double)depths[i] < kMinDepth) return;
  const float4 q = rots[i];
  const float w = q.x, x = q.y, y = q.z, z = q.w;
  const float s0 = scales[3 * i], s1 = scales[3 * i + 1], s2 = scales[3 * i + 2];
  const float xx = x * x, yy = y * y, zz = z * z;
  const float xy = x * y, xz = x * z, yz = y * z;
  const float xw = x * w, yw = y * w, zw = z * w;
  const float R[9] = {1.f - 2.f * (yy + zz), 2.f * (xy - zw), 2.f * (xz + yw),
                      2.f * (xy + zw), 1.f - 2.f * (xx + zz), 2.f * (yz - xw),
                      2.f * (xz - yw), 2.f * (yz + xw), 1.f - 2.f * (xx + yy)};
  const float s[3] = {s0, s1, s2};
  float M[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) M[r * 3 + c] = R[r * 3 + c] * s[c];
  float dS[54];
#pragma unroll
  for (int k = 0; k < 54; ++k) dS[k] = 0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    dS[0 * 9 + c] = 2.f * M[c];
    dS[1 * 9 + c] = M[3 + c];
    dS[1 * 9 + 3 + c] = M[c];
    dS[2 * 9 + c] = M[6 + c];
    dS[2 * 9 + 6 + c] = M[c];
    dS[3 * 9 + 3 + c] = 2.f * M[3 + c];
    dS[4 * 9 + 3 + c] = M[6 + c];
    dS[4 * 9 + 6 + c] = M[3 + c];
    dS[5 * 9 + 6 + c] = 2.f * M[6 + c];
  }
  const float dMq[36] = {
      0.f, 0.f, -4.f * s0 * y, -4.f * s0 * z,
      -2.f * s1 * z, 2.f * s1 * y, 2.f * s1 * x, -2.f * s1 * w,
      2.f * s2 * y, 2.f * s2 * z, 2.f * s2 * w, 2.f * s2 * x,
      2.f * s0 * z, 2.f * s0 * y, 2.f * s0 * x, 2.f * s0 * w,
      0.f, -4.f * s1 * x, 0.f, -4.f * s1 * z,
      -2.f * s2 * x, -2.f * s2 * w, 2.f * s2 * z, 2.f * s2 * y,
      -2.f * s0 * y, 2.f * s0 * z, -2.f * s0 * w, 2.f * s0 * x,
      2.f * s1 * x, 2.f * s1 * w, 2.f * s1 * z, 2.f * s1 * y,
      0.f, -4.f * s2 * x, -4.f * s2 * y, 0.f};
  float dMs[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) dMs[k] = 0.f;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) dMs[(r * 3 + c) * 3 + c] = R[r * 3 + c];
  float out_r[24], out_s[18];
  small_matmul<6, 9, 4>(dS, dMq, out_r);
  small_matmul<6, 9, 3>(dS, dMs, out_s);
#pragma unroll
  for (int k = 0; k < 24; ++k) drot[(size_t)i * 24 + k] = out_r[k];
#pragma unroll
  for (int k = 0; k < 18; ++k) dscale[(size_t)i * 18 + k] = out_s[k];
}

// dcov2d/dcov3d (N,3,6), dcov2d/dpc (N,3,3); the clamped x, y enter J and
// dM/dpc (the clamp is not differentiated, as in the reference).
__global__ __launch_bounds__(256) void k_cov2d_jac(int n, const float* __restrict__ cov3ds,
                                                   const float* __restrict__ pcs,
                                                   const float* __restrict__ Rcw,
                                                   const float* __restrict__ depths, float fx,
                                                   float fy, float tfx, float tfy,
                                                   float* __restrict__ d3, float* __restrict__ dp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if ((double)depths[i] < kMinDepth) return;
  float R[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = Rcw[k];
  float x = pcs[3 * i], y = pcs[3 * i + 1];
  const float z = pcs[3 * i + 2];
  const float* c3 = cov3ds + (size_t)i * 6;
  const float a = c3[0], b = c3[1], c = c3[2], d = c3[3], e = c3[4], f = c3[5];
  const float limx = 1.3f * tfx;
  const float limy = 1.3f * tfy;
  x = fminf(limx, fmaxf(-limx, x / z)) * z;
  y = fminf(limy, fmaxf(-limy, y / z)) * z;
  const float z2 = z * z;
  const float J[6] = {fx / z, 0.0f, -(fx * x) / z2, 0.0f, fy / z, -(fy * y) / z2};
  float M[6];
  small_matmul<2, 3, 3>(J, R, M);
  float o[18];
  // rows: (M0, M0), (M0, M1), (M1, M1)
  o[0] = M[0] * M[0];
  o[1] = 2.f * M[0] * M[1];
  o[2] = 2.f * M[0] * M[2];
  o[3] = M[1] * M[1];
  o[4] = 2.f * M[1] * M[2];
  o[5] = M[2] * M[2];
  o[6] = M[0] * M[3];
  o[7] = M[0] * M[4] + M[1] * M[3];
  o[8] = M[0] * M[5] + M[2] * M[3];
  o[9] = M[1] * M[4];
  o[10] = M[1] * M[5] + M[2] * M[4];
  o[11] = M[2] * M[5];
  o[12] = M[3] * M[3];
  o[13] = 2.f * M[3] * M[4];
  o[14] = 2.f * M[3] * M[5];
  o[15] = M[4] * M[4];
  o[16] = 2.f * M[4] * M[5];
  o[17] = M[5] * M[5];
#pragma unroll
  for (int k = 0; k < 18; ++k) d3[(size_t)i * 18 + k] = o[k];
  const float S[9] = {a, b, c, b, d, e, c, e, f};
  float dM[18];
#pragma unroll
  for (int k = 0; k < 18; ++k) dM[k] = 0.f;
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    const float* Sc = S + 3 * cc;
    dM[cc] = 2.f * Sc[0] * M[0] + 2.f * Sc[1] * M[1] + 2.f * Sc[2] * M[2];
    dM[6 + cc] = Sc[0] * M[3] + Sc[1] * M[4] + Sc[2] * M[5];
    dM[6 + 3 + cc] = Sc[0] * M[0] + Sc[1] * M[1] + Sc[2] * M[2];
    dM[12 + 3 + cc] = 2.f * Sc[0] * M[3] + 2.f * Sc[1] * M[4] + 2.f * Sc[2] * M[5];
  }
  const float z2_inv = 1.f / (z * z);
  const float z3_inv = z2_inv / z;
  float dMp[18];
#pragma unroll
  for (int k = 0; k < 18; ++k) dMp[k] = 0.f;
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    const float R0 = R[cc], R1 = R[3 + cc], R2 = R[6 + cc];
    dMp[cc * 3 + 0] = -fx * R2 * z2_inv;
    dMp[cc * 3 + 2] = -fx * R0 * z2_inv + 2.f * fx * R2 * x * z3_inv;
    dMp[(3 + cc) * 3 + 1] = -fy * R2 * z2_inv;
    dMp[(3 + cc) * 3 + 2] = -fy * R1 * z2_inv + 2.f * fy * R2 * y * z3_inv;
  }
  float out[9];
  small_matmul<3, 6, 3>(dM, dMp, out);
#pragma unroll
  for (int k = 0; k < 9; ++k) dp[(size_t)i * 9 + k] = out[k];
}

__global__ __launch_bounds__(256) void k_inverse_jac(int n, const float* __restrict__ cov2ds,
                                                     const float* __restrict__ depths,
                                                     float* __restrict__ dcinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if ((double)depths[i] < kMinDepth) return;
  const float a = cov2ds[3 * i], b = cov2ds[3 * i + 1], c = cov2ds[3 * i + 2];
  const float di = 1.f / (a * c - b * b);
  if (di != di) return;
  const float d2 = di * di;
  float* o = dcinv + (size_t)i * 9;
  o[0] = -c * c * d2;
  o[1] = 2.f * b * c * d2;
  o[2] = -a * c * d2 + di;
  o[3] = b * c * d2;
  o[4] = -2.f * b * b * d2 - di;
  o[5] = a * b * d2;
  o[6] = -a * c * d2 + di;
  o[7] = 2.f * a * b * d2;
  o[8] = -a * a * d2;
}

// dW/dpw (N,K,3): dW/ddir from the SH polynomials (double constants, each
// band's sum added to the float accumulator, as in the reference; including
// its dY_9/dy = C30 (3x^2 - 2y^2), kernel.cu:840) times
// ddir/dpw = (I - dir dir^T) / |d|.
template <int K>
__global__ __launch_bounds__(256) void k_sh2weight_jac(int n, int sh_dim,
                                                       const float* __restrict__ shs,
                                                       const float* __restrict__ pws,
                                                       const float* __restrict__ twc,
                                                       float* __restrict__ out_all) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float* out = out_all + (size_t)i * K * 3;
  if (sh_dim <= 1) {
#pragma unroll
    for (int q = 0; q < K * 3; ++q) out[q] = 0.f;
    return;
  }
  const float* sh = shs + (size_t)i * sh_dim * K;
  const float d0 = pws[3 * i] - twc[0], d1 = pws[3 * i + 1] - twc[1], d2 = pws[3 * i + 2] - twc[2];
  const float ni = 1.f / sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
  const float x = d0 * ni, y = d1 * ni, z = d2 * ni;
  const float ni3 = ni * ni * ni;
  const float r00 = -d0 * d0 * ni3 + ni, r11 = -d1 * d1 * ni3 + ni, r22 = -d2 * d2 * ni3 + ni;
  const float r01 = -d0 * d1 * ni3, r02 = -d0 * d2 * ni3, r12 = -d1 * d2 * ni3;
  const float dr[9] = {r00, r01, r02, r01, r11, r12, r02, r12, r22};
  const float xx = x * x, yy = y * y, zz = z * z;
  const double X = x, Yd = y, Z = z;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    auto S = [&](int l) { return (double)sh[l * K + k]; };
    float g0 = (float)(0.0 + SH_C1_2 * S(3));
    float g1 = (float)(0.0 + SH_C1_0 * S(1));
    float g2 = (float)(0.0 + SH_C1_1 * S(2));
    if (sh_dim > 4) {
      g0 = (float)((double)g0 + ((((SH_C2_0 * Yd) * S(4) - ((SH_C2_2 * 2) * X) * S(6)) +
                                 (SH_C2_3 * Z) * S(7)) + ((SH_C2_4 * 2) * X) * S(8)));
      g1 = (float)((double)g1 + ((((SH_C2_0 * X) * S(4) + (SH_C2_1 * Z) * S(5)) -
                                 ((SH_C2_2 * 2) * Yd) * S(6)) - ((SH_C2_4 * 2) * Yd) * S(8)));
      g2 = (float)((double)g2 + (((SH_C2_1 * Yd) * S(5) + ((SH_C2_2 * 4) * Z) * S(6)) +
                                 (SH_C2_3 * X) * S(7)));
      if (sh_dim > 9) {
        const float p0 = 4.f * zz - 3.f * xx - yy, p1 = 3.f * xx - 3.f * yy;
        const float q0 = 3.f * xx - 2.f * yy, q1 = 4.f * zz - xx - 3.f * yy;
        const float u0 = 6.f * zz - 3.f * xx - 3.f * yy, u1 = xx - yy;
        double a0 = (((6.0 * SH_C3_0) * X) * Yd) * S(9);
        a0 = a0 + ((SH_C3_1 * Yd) * Z) * S(10);
        a0 = a0 - (((2.0 * SH_C3_2) * X) * Yd) * S(11);
        a0 = a0 - (((6.0 * SH_C3_3) * X) * Z) * S(12);
        a0 = a0 + (SH_C3_4 * S(13)) * (double)p0;
        a0 = a0 + (((2.0 * SH_C3_5) * X) * Z) * S(14);
        a0 = a0 + (SH_C3_6 * S(15)) * (double)p1;
        double a1 = (SH_C3_0 * S(9)) * (double)q0;
        a1 = a1 + ((SH_C3_1 * X) * Z) * S(10);
        a1 = a1 + (SH_C3_2 * S(11)) * (double)q1;
        a1 = a1 - (((6.0 * SH_C3_3) * Yd) * Z) * S(12);
        a1 = a1 - (((2.0 * SH_C3_4) * X) * Yd) * S(13);
        a1 = a1 - (((2.0 * SH_C3_5) * Yd) * Z) * S(14);
        a1 = a1 - (((6.0 * SH_C3_6) * X) * Yd) * S(15);
        double a2 = ((SH_C3_1 * X) * Yd) * S(10);
        a2 = a2 + (((8.0 * SH_C3_2) * Yd) * Z) * S(11);
        a2 = a2 + (SH_C3_3 * S(12)) * (double)u0;
        a2 = a2 + (((8.0 * SH_C3_4) * X) * Z) * S(13);
        a2 = a2 + (SH_C3_5 * S(14)) * (double)u1;
        g0 = (float)((double)g0 + a0);
        g1 = (float)((double)g1 + a1);
        g2 = (float)((double)g2 + a2);
      }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) out[k * 3 + j] = g0 * dr[j] + g1 * dr[3 + j] + g2 * dr[6 + j];
  }
}

// ---------------------------------------------------------------------------
// Full-rank SH gradient fused with the Adam update of the coefficients: every
// coefficient's gradient g = dw[n,k] * Y_lm(n) (the product gsmodel.py:135-144
// forms) is used where it is made, so dL/dshs is never stored.  The update is
// torch.optim.Adam's (no amsgrad), element by element:
//   g += wd * p;  m += (1 - b1) * (g - m);  v = v * b2 + ((1 - b2) * g) * g;
//   p += -(lr / bc1) * (m / (sqrt(v) / sqrt(bc2) + eps)).
// A block owns 64 Gaussians: their basis and dw in LDS, then p, m, v as
// float4 runs (a row is SH*K floats, the block's span 64 rows).
// ---------------------------------------------------------------------------
struct AdamArgs {
  float neg_step, omb1, b2, omb2, bc2_sqrt, eps, wd;
};

__device__ __forceinline__ void adam_elem(float& p, float& m, float& v, float g, const AdamArgs& a) {
  if (a.wd != 0.f) g = g + a.wd * p;
  m = m + a.omb1 * (g - m);
  v = v * a.b2 + (a.omb2 * g) * g;
  const float denom = sqrtf(v) / a.bc2_sqrt + a.eps;
  p = p + a.neg_step * (m / denom);
}

// The update of one coefficient tensor holding bands LM0 .. LM0 + NB - 1 of
// each Gaussian (rows of NB*K floats): the full-rank tensor is LM0 = 0,
// NB = SH; the reference's two leaves are the DC band (0, 1) and the rest
// (1, 15).
template <int K, int LM0, int NB>
__device__ __forceinline__ void sh_adam_span(int n, const rgp_camera* __restrict__ cam,
                                             const float* __restrict__ pws, const float* __restrict__ dw,
                                             float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                             const AdamArgs& a) {
  constexpr int G = 64, ROW = NB * K, SH = LM0 + NB;
  __shared__ float sY[G][NB + 1];
  __shared__ float sW[G][K];
  const int g0 = blockIdx.x * G, t = threadIdx.x;
  if (t < G) {
    const int i = g0 + t;
    float Y[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) Y[l] = 0.f;
    if (i < n) {
      if (SH > 1) sh_basis(pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], cam->twc, SH, Y);
#pragma unroll
      for (int k = 0; k < K; ++k) sW[t][k] = dw[(size_t)i * K + k];
    }
    Y[0] = (float)SH_C0_0;
#pragma unroll
    for (int l = 0; l < NB; ++l) sY[t][l] = Y[LM0 + l];
  }
  __syncthreads();
  const int cnt = min(G, n - g0);
  const int total = cnt * ROW;
  const size_t base = (size_t)g0 * ROW;
  auto grad = [&](int e) {
    const int gl = e / ROW, r = e - gl * ROW, lm = r / K, k = r - lm * K;
    return sW[gl][k] * sY[gl][lm];
  };
  float4* p4 = (float4*)(p + base);
  float4* m4 = (float4*)(m + base);
  float4* v4 = (float4*)(v + base);
  {
    // a full block is F4 float4 of each array, PER per thread (the last one
    // on the first F4 - 256 (PER-1) threads): all loads are issued before
    // the first update.  Non-temporal: p, m and v are touched once per step,
    // and stores left dirty in the caches would be written back during the
    // next kernel.
    constexpr int F4 = G * ROW / 4, PER = (F4 + 255) / 256;
    if (cnt == G) {
      const bool tail = t + 256 * (PER - 1) < F4;
      const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
      float4 pp[PER], mm[PER], vv[PER];
#pragma unroll
      for (int j = 0; j < PER - 1; ++j) {
        pp[j] = ld_nt(&p4[t + 256 * j]);
        mm[j] = ld_nt(&m4[t + 256 * j]);
        vv[j] = ld_nt(&v4[t + 256 * j]);
      }
      pp[PER - 1] = mm[PER - 1] = vv[PER - 1] = z;
      if (tail) {
        pp[PER - 1] = ld_nt(&p4[t + 256 * (PER - 1)]);
        mm[PER - 1] = ld_nt(&m4[t + 256 * (PER - 1)]);
        vv[PER - 1] = ld_nt(&v4[t + 256 * (PER - 1)]);
      }
      asm volatile("" ::: "memory");  // keep the loads ahead of the updates
#pragma unroll
      for (int j = 0; j < PER; ++j) {
        if (j == PER - 1 && !tail) break;
        const int q = t + 256 * j;
        adam_elem(pp[j].x, mm[j].x, vv[j].x, grad(4 * q), a);
        adam_elem(pp[j].y, mm[j].y, vv[j].y, grad(4 * q + 1), a);
        adam_elem(pp[j].z, mm[j].z, vv[j].z, grad(4 * q + 2), a);
        adam_elem(pp[j].w, mm[j].w, vv[j].w, grad(4 * q + 3), a);
        st_nt(&p4[q], pp[j]);
        st_nt(&m4[q], mm[j]);
        st_nt(&v4[q], vv[j]);
      }
      return;
    }
  }
  const int n4 = total >> 2;
  for (int q = t; q < n4; q += blockDim.x) {
    float4 pp = p4[q], mm = m4[q], vv = v4[q];
    adam_elem(pp.x, mm.x, vv.x, grad(4 * q), a);
    adam_elem(pp.y, mm.y, vv.y, grad(4 * q + 1), a);
    adam_elem(pp.z, mm.z, vv.z, grad(4 * q + 2), a);
    adam_elem(pp.w, mm.w, vv.w, grad(4 * q + 3), a);
    p4[q] = pp;
    m4[q] = mm;
    v4[q] = vv;
  }
  for (int e = 4 * n4 + t; e < total; e += blockDim.x)
    adam_elem(p[base + e], m[base + e], v[base + e], grad(e), a);
}

template <int K, int SH>
__global__ __launch_bounds__(256) void k_sh_adam_full(int n, const rgp_camera* __restrict__ cam,
                                                      const float* __restrict__ pws,
                                                      const float* __restrict__ dw, float* __restrict__ p,
                                                      float* __restrict__ m, float* __restrict__ v,
                                                      AdamArgs a) {
  sh_adam_span<K, 0, SH>(n, cam, pws, dw, p, m, v, a);
}

template <int K, int LM0, int NB>
__global__ __launch_bounds__(256) void k_sh_adam_bands(int n, const rgp_camera* __restrict__ cam,
                                                       const float* __restrict__ pws,
                                                       const float* __restrict__ dw, float* __restrict__ p,
                                                       float* __restrict__ m, float* __restrict__ v,
                                                       AdamArgs a) {
  sh_adam_span<K, LM0, NB>(n, cam, pws, dw, p, m, v, a);
}

// Adam over one factor tensor's span of a 64-Gaussian block (rows of ROW
// floats) with the gradients the block formed in LDS (row stride GS, this
// tensor's at offset GOFF): a full block as PER float4 per thread, every
// load issued first, non-temporal as in sh_adam_span; a partial last block
// element by element.
template <int ROW, int GOFF, int GS>
__device__ __forceinline__ void lowrank_adam_span(float* __restrict__ p, float* __restrict__ m,
                                                  float* __restrict__ v, int cnt, const float* sG,
                                                  const AdamArgs& a) {
  constexpr int F4 = 64 * ROW / 4, PER = (F4 + 63) / 64;
  const int t = threadIdx.x;
  auto grad = [&](int e) {
    const int gl = e / ROW;
    return sG[gl * GS + GOFF + (e - gl * ROW)];
  };
  float4* p4 = reinterpret_cast<float4*>(p);
  float4* m4 = reinterpret_cast<float4*>(m);
  float4* v4 = reinterpret_cast<float4*>(v);
  if (cnt == 64) {
    const bool tail = t + 64 * (PER - 1) < F4;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 pp[PER], mm[PER], vv[PER];
#pragma unroll
    for (int j = 0; j < PER - 1; ++j) {
      pp[j] = ld_nt(&p4[t + 64 * j]);
      mm[j] = ld_nt(&m4[t + 64 * j]);
      vv[j] = ld_nt(&v4[t + 64 * j]);
    }
    pp[PER - 1] = mm[PER - 1] = vv[PER - 1] = z;
    if (tail) {
      pp[PER - 1] = ld_nt(&p4[t + 64 * (PER - 1)]);
      mm[PER - 1] = ld_nt(&m4[t + 64 * (PER - 1)]);
      vv[PER - 1] = ld_nt(&v4[t + 64 * (PER - 1)]);
    }
    asm volatile("" ::: "memory");  // keep the loads ahead of the updates
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      if (j == PER - 1 && !tail) break;
      const int q = t + 64 * j;
      adam_elem(pp[j].x, mm[j].x, vv[j].x, grad(4 * q), a);
      adam_elem(pp[j].y, mm[j].y, vv[j].y, grad(4 * q + 1), a);
      adam_elem(pp[j].z, mm[j].z, vv[j].z, grad(4 * q + 2), a);
      adam_elem(pp[j].w, mm[j].w, vv[j].w, grad(4 * q + 3), a);
      st_nt(&p4[q], pp[j]);
      st_nt(&m4[q], mm[j]);
      st_nt(&v4[q], vv[j]);
    }
    return;
  }
  for (int e = t; e < cnt * ROW; e += 64) adam_elem(p[e], m[e], v[e], grad(e), a);
}

// The low-rank SH backward (gsmodel.py:194-198, sh2Weight_fast_B) fused with
// torch.optim.Adam's update of dc (N,K), A (N,K,R) and B (N,R,15): a block
// of 64 threads forms the gradients of its 64 Gaussians (one per thread,
// lowrank_q15_grad, from the factors before the update) into LDS, then
// updates the block's rows of the three tensors.  The gradients never reach
// HBM.  Blocks own disjoint rows, so reading A and B before updating them
// needs only the block barrier.
template <int K, int R>
__global__ __launch_bounds__(64) void k_lowrank_adam_q15(
    int n, const rgp_camera* __restrict__ cam, const float* __restrict__ pws, const float* __restrict__ dw,
    float* __restrict__ dc, float* __restrict__ mdc, float* __restrict__ vdc, float* __restrict__ A,
    float* __restrict__ mA, float* __restrict__ vA, float* __restrict__ B, float* __restrict__ mB,
    float* __restrict__ vB, AdamArgs a) {
  constexpr int Q = 15, P = K, GS = K + P * R + R * Q + 1;
  __shared__ float sG[64 * GS];
  const int g0 = blockIdx.x * 64, t = threadIdx.x, i = g0 + t;
  if (i < n) {
    float Y[16];
    sh_basis(pws[3 * i], pws[3 * i + 1], pws[3 * i + 2], cam->twc, 16, Y);
    float ar[P * R], br[R * Q], dwr[K], gdc[K], da[P * R], db[R * Q];
#pragma unroll
    for (int j = 0; j < P * R; ++j) ar[j] = A[(size_t)i * P * R + j];
#pragma unroll
    for (int j = 0; j < R * Q; ++j) br[j] = B[(size_t)i * R * Q + j];
#pragma unroll
    for (int k = 0; k < K; ++k) dwr[k] = dw[(size_t)i * K + k];
    lowrank_q15_grad<K, R>(Y, ar, br, dwr, gdc, da, db);
    float* row = sG + t * GS;
#pragma unroll
    for (int k = 0; k < K; ++k) row[k] = gdc[k];
#pragma unroll
    for (int j = 0; j < P * R; ++j) row[K + j] = da[j];
#pragma unroll
    for (int j = 0; j < R * Q; ++j) row[K + P * R + j] = db[j];
  }
  __syncthreads();
  const int cnt = min(64, n - g0);
  lowrank_adam_span<K, 0, GS>(dc + (size_t)g0 * K, mdc + (size_t)g0 * K, vdc + (size_t)g0 * K, cnt, sG, a);
  lowrank_adam_span<P * R, K, GS>(A + (size_t)g0 * P * R, mA + (size_t)g0 * P * R, vA + (size_t)g0 * P * R,
                                  cnt, sG, a);
  lowrank_adam_span<R * Q, K + P * R, GS>(B + (size_t)g0 * R * Q, mB + (size_t)g0 * R * Q,
                                          vB + (size_t)g0 * R * Q, cnt, sG, a);
}

// The multi-view form: the gradients are the sums over `views` cameras of the
// single-view chain (as k_lowrank_backward_q15_views forms them, view by view
// into the same accumulators), then the same update.
template <int K, int R>
__global__ __launch_bounds__(64) void k_lowrank_adam_q15_views(
    int views, int n, const rgp_camera* __restrict__ cams, const float* __restrict__ pws,
    const float* __restrict__ dw, float* __restrict__ dc, float* __restrict__ mdc, float* __restrict__ vdc,
    float* __restrict__ A, float* __restrict__ mA, float* __restrict__ vA, float* __restrict__ B,
    float* __restrict__ mB, float* __restrict__ vB, AdamArgs a) {
  constexpr int Q = 15, P = K, GS = K + P * R + R * Q + 1;
  __shared__ float sG[64 * GS];
  const int g0 = blockIdx.x * 64, t = threadIdx.x, i = g0 + t;
  if (i < n) {
    float ar[P * R], br[R * Q], da[P * R], db[R * Q], dd[K];
#pragma unroll
    for (int j = 0; j < P * R; ++j) { ar[j] = A[(size_t)i * P * R + j]; da[j] = 0.f; }
#pragma unroll
    for (int j = 0; j < R * Q; ++j) { br[j] = B[(size_t)i * R * Q + j]; db[j] = 0.f; }
#pragma unroll
    for (int kk = 0; kk < K; ++kk) dd[kk] = 0.f;
    const float px = pws[3 * i], py = pws[3 * i + 1], pz = pws[3 * i + 2];
    for (int v = 0; v < views; ++v) {
      float Y[16];
      sh_basis(px, py, pz, cams[v].twc, 16, Y);
      const float* dwv = dw + ((size_t)v * n + i) * K;
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const float dwk = dwv[kk];
        dd[kk] = dd[kk] + dwk * (float)SH_C0_0;
#pragma unroll
        for (int lm = 0; lm < 15; ++lm) {
          const int g = lm * K + kk, row = g / Q, col = g % Q;
          const float y = Y[lm + 1];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            da[row * R + r] = da[row * R + r] + dwk * br[r * Q + col] * y;
            db[r * Q + col] = db[r * Q + col] + dwk * ar[row * R + r] * y;
          }
        }
      }
    }
    float* row = sG + t * GS;
#pragma unroll
    for (int kk = 0; kk < K; ++kk) row[kk] = dd[kk];
#pragma unroll
    for (int j = 0; j < P * R; ++j) row[K + j] = da[j];
#pragma unroll
    for (int j = 0; j < R * Q; ++j) row[K + P * R + j] = db[j];
  }
  __syncthreads();
  const int cnt = min(64, n - g0);
  lowrank_adam_span<K, 0, GS>(dc + (size_t)g0 * K, mdc + (size_t)g0 * K, vdc + (size_t)g0 * K, cnt, sG, a);
  lowrank_adam_span<P * R, K, GS>(A + (size_t)g0 * P * R, mA + (size_t)g0 * P * R, vA + (size_t)g0 * P * R,
                                  cnt, sG, a);
  lowrank_adam_span<R * Q, K + P * R, GS>(B + (size_t)g0 * R * Q, mB + (size_t)g0 * R * Q,
                                          vB + (size_t)g0 * R * Q, cnt, sG, a);
}

}  // namespace rgp

using namespace rgp;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static inline int grid1(int n) { return (n + 255) / 256; }

extern "C" int rgp_project(int n, const float* pws, const float* Rcw, const float* tcw, float fx,
                           float fy, float cx, float cy, float* us, float* pcs, float* depths,
                           float* du_dpcs, hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_project: n < 0");
  if (n == 0) return RGP_OK;
  hipLaunchKernelGGL(k_project, dim3(grid1(n)), dim3(256), 0, s, n, pws, Rcw, tcw, fx, fy, cx, cy,
                     us, pcs, depths);
  if (du_dpcs)
    hipLaunchKernelGGL(k_project_jac, dim3(grid1(n)), dim3(256), 0, s, n, pws, Rcw, tcw, fx, fy,
                       du_dpcs);
  return check_launch("rgp_project");
}

extern "C" int rgp_compute_cov3d(int n, const float* rots, const float* scales,
                                 const float* depths, float* cov3ds, float* dcov3d_drots,
                                 float* dcov3d_dscales, hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_compute_cov3d: n < 0");
  if (!dcov3d_drots != !dcov3d_dscales)
    return set_error(RGP_ERR_ARG, "rgp_compute_cov3d: give both Jacobians or neither");
  if (n == 0) return RGP_OK;
  hipLaunchKernelGGL(k_cov3d, dim3(grid1(n)), dim3(256), 0, s, n, (const float4*)rots, scales,
                     depths, cov3ds);
  if (dcov3d_drots)
    hipLaunchKernelGGL(k_cov3d_jac, dim3(grid1(n)), dim3(256), 0, s, n, (const float4*)rots, scales,
                       depths, dcov3d_drots, dcov3d_dscales);
  return check_launch("rgp_compute_cov3d");
}

extern "C" int rgp_compute_cov2d(int n, const float* cov3ds, const float* pcs, const float* Rcw,
                                 const float* depths, float fx, float fy, float width,
                                 float height, float* cov2ds, float* dcov2d_dcov3ds,
                                 float* dcov2d_dpcs, hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_compute_cov2d: n < 0");
  if (!dcov2d_dcov3ds != !dcov2d_dpcs)
    return set_error(RGP_ERR_ARG, "rgp_compute_cov2d: give both Jacobians or neither");
  if (n == 0) return RGP_OK;
  // gausplat.cu:357-358, in float
  const float tfx = width / (2.f * fx);
  const float tfy = height / (2.f * fy);
  hipLaunchKernelGGL(k_cov2d, dim3(grid1(n)), dim3(256), 0, s, n, cov3ds, pcs, Rcw, depths, fx,
                     fy, tfx, tfy, cov2ds);
  if (dcov2d_dcov3ds)
    hipLaunchKernelGGL(k_cov2d_jac, dim3(grid1(n)), dim3(256), 0, s, n, cov3ds, pcs, Rcw, depths, fx,
                       fy, tfx, tfy, dcov2d_dcov3ds, dcov2d_dpcs);
  return check_launch("rgp_compute_cov2d");
}

extern "C" int rgp_inverse_cov2d(int n, const float* cov2ds, float* depths, float* cinv2ds,
                                 int32_t* areas, float* dcinv2d_dcov2ds, hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_inverse_cov2d: n < 0");
  if (n == 0) return RGP_OK;
  hipLaunchKernelGGL(k_inverse, dim3(grid1(n)), dim3(256), 0, s, n, cov2ds, depths, cinv2ds,
                     areas);
  // after the cull: depths already holds -1 for a NaN determinant
  if (dcinv2d_dcov2ds)
    hipLaunchKernelGGL(k_inverse_jac, dim3(grid1(n)), dim3(256), 0, s, n, cov2ds, depths,
                       dcinv2d_dcov2ds);
  return check_launch("rgp_inverse_cov2d");
}

#define RGP_K_SWITCH(K, CALL)                    \
  switch (K) {                                   \
    case 1: { constexpr int KK = 1; CALL; } break; \
    case 2: { constexpr int KK = 2; CALL; } break; \
    case 3: { constexpr int KK = 3; CALL; } break; \
    case 4: { constexpr int KK = 4; CALL; } break; \
    case 5: { constexpr int KK = 5; CALL; } break; \
    case 6: { constexpr int KK = 6; CALL; } break; \
    case 7: { constexpr int KK = 7; CALL; } break; \
    case 8: { constexpr int KK = 8; CALL; } break; \
    default: break;                              \
  }

namespace rgp {
static void launch_lowrank_backward(int n, int K, int P, int Q, int rank, const float* A,
                                    const float* B, const float* ylm, const float* pws,
                                    const float* twc, const float* dw, float* ddc, float* dA,
                                    float* dB, hipStream_t s) {
  if (Q == 15 && P == K && P * Q / K == 15 && (rank == 1 || rank == 2)) {
    if (rank == 1) {
      RGP_K_SWITCH(K, hipLaunchKernelGGL((k_lowrank_backward_q15<KK, 1>), dim3(grid1(n)), dim3(256),
                                         0, s, n, A, B, ylm, pws, twc, dw, ddc, dA, dB));
    } else {
      RGP_K_SWITCH(K, hipLaunchKernelGGL((k_lowrank_backward_q15<KK, 2>), dim3(grid1(n)), dim3(256),
                                         0, s, n, A, B, ylm, pws, twc, dw, ddc, dA, dB));
    }
    return;
  }
  hipLaunchKernelGGL(k_lowrank_backward, dim3(grid1(n)), dim3(256), 0, s, n, K, P, Q, rank, A, B,
                     ylm, pws, twc, dw, ddc, dA, dB);
}
}  // namespace rgp

extern "C" int rgp_sh2weight(int n, int K, int sh_dim, const float* shs, const float* pws,
                             const float* twc, float* weights, float* basis, float* dweight_dpws,
                             hipStream_t s) {
  if (n < 0 || K < 1 || K > RGP_MAX_DIM)
    return set_error(RGP_ERR_ARG, "rgp_sh2weight: need n >= 0 and 1 <= K <= 8");
  if (!(sh_dim == 1 || sh_dim == 4 || sh_dim == 9 || sh_dim == 16))
    return set_error(RGP_ERR_ARG, "rgp_sh2weight: sh_dim must be 1, 4, 9 or 16");
  if (n == 0) return RGP_OK;
  RGP_K_SWITCH(K, hipLaunchKernelGGL(k_sh2weight<KK>, dim3(grid1(n)), dim3(256), 0, s, n, sh_dim,
                                     shs, pws, twc, weights, basis));
  if (dweight_dpws)
    RGP_K_SWITCH(K, hipLaunchKernelGGL(k_sh2weight_jac<KK>, dim3(grid1(n)), dim3(256), 0, s, n,
                                       sh_dim, shs, pws, twc, dweight_dpws));
  return check_launch("rgp_sh2weight");
}

static int check_lowrank(const char* who, int K, int P, int Q, int rank) {
  if (K < 1 || K > RGP_MAX_DIM) return set_error(RGP_ERR_ARG, "%s: need 1 <= K <= 8", who);
  if (P < 1 || Q < 1 || rank < 1) return set_error(RGP_ERR_ARG, "%s: P, Q, rank must be >= 1", who);
  if ((P * Q) % K != 0) return set_error(RGP_ERR_ARG, "%s: P*Q must be a multiple of K", who);
  const int L = P * Q / K;
  if (!(L == 0 || L == 3 || L == 8 || L == 15))
    return set_error(RGP_ERR_ARG, "%s: P*Q/K must be 3, 8 or 15 (SH degree 1-3)", who);
  return RGP_OK;
}

extern "C" int rgp_sh2weight_fast(int n, int K, int P, int Q, int rank, const float* dc,
                                  const float* A, const float* B, const float* pws,
                                  const float* twc, float* weights, float* dweight_ddc,
                                  float* ylm, hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_sh2weight_fast: n < 0");
  int rc = check_lowrank("rgp_sh2weight_fast", K, P, Q, rank);
  if (rc) return rc;
  if (n == 0) return RGP_OK;
  RGP_K_SWITCH(K, hipLaunchKernelGGL(k_sh2weight_fast<KK>, dim3(grid1(n)), dim3(256), 0, s, n, P,
                                     Q, rank, dc, A, B, pws, twc, weights, dweight_ddc, ylm));
  return check_launch("rgp_sh2weight_fast");
}

extern "C" int rgp_sh2weight_fast_backward(int n, int K, int P, int Q, int rank, const float* A,
                                           const float* B, const float* ylm,
                                           const float* dloss_dw, float* dA, float* dB,
                                           hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_sh2weight_fast_backward: n < 0");
  int rc = check_lowrank("rgp_sh2weight_fast_backward", K, P, Q, rank);
  if (rc) return rc;
  if (!ylm) return set_error(RGP_ERR_ARG, "rgp_sh2weight_fast_backward: ylm is required");
  if (n == 0) return RGP_OK;
  rgp::launch_lowrank_backward(n, K, P, Q, rank, A, B, ylm, nullptr, nullptr, dloss_dw, nullptr, dA,
                               dB, s);
  return check_launch("rgp_sh2weight_fast_backward");
}

extern "C" int rgp_sh2color(int n, int sh_dim, const float* shs, const float* pws,
                            const float* twc, float* colors, float* dcolor_dshs, float* dcolor_dpws,
                            hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_sh2color: n < 0");
  if (!(sh_dim == 1 || sh_dim == 4 || sh_dim == 9 || sh_dim == 16))
    return set_error(RGP_ERR_ARG, "rgp_sh2color: sh_dim must be 1, 4, 9 or 16");
  if (!dcolor_dshs != !dcolor_dpws)
    return set_error(RGP_ERR_ARG, "rgp_sh2color: give both Jacobian outputs or neither");
  if (n == 0) return RGP_OK;
  hipLaunchKernelGGL(k_sh2color, dim3(grid1(n)), dim3(256), 0, s, n, sh_dim, shs, pws, twc,
                     colors);
  if (dcolor_dshs)
    hipLaunchKernelGGL(k_sh2color_jac, dim3(grid1(n)), dim3(256), 0, s, n, sh_dim, shs, pws, twc,
                       dcolor_dshs, dcolor_dpws);
  return check_launch("rgp_sh2color");
}

extern "C" int rgp_record_floats(int dim) {
  if (dim < 1 || dim > RGP_MAX_DIM) return 0;
  return 8 + 4 * ((dim + 3) / 4);
}

extern "C" int rgp_pack_records(int n, int dim, const float* us, const float* cinv2ds,
                                const float* alphas, const float* feat, float* records,
                                hipStream_t s) {
  if (n < 0 || dim < 1 || dim > RGP_MAX_DIM)
    return set_error(RGP_ERR_ARG, "rgp_pack_records: need n >= 0 and 1 <= dim <= 8");
  if (n == 0) return RGP_OK;
  RGP_K_SWITCH(dim, hipLaunchKernelGGL(k_pack<KK>, dim3(grid1(n)), dim3(256), 0, s, n, us,
                                       cinv2ds, alphas, feat, (float4*)records));
  return check_launch("rgp_pack_records");
}

extern "C" int rgp_bin_rects(int n, const float* us, int32_t* areas, float* depths, int width,
                             int height, void* ws, int64_t capacity, hipStream_t s) {
  if (n < 0 || width <= 0 || height <= 0)
    return set_error(RGP_ERR_ARG, "rgp_bin_rects: bad size");
  BinLayout L = bin_layout(n, capacity, width, height);
  char* base = (char*)ws;
  if (n == 0) return bin_reset_header(ws, L.reset_words, s);
  hipLaunchKernelGGL(k_rects, dim3(grid1(n)), dim3(256), 0, s, n, us, areas, depths, L.gx, L.gy,
                     (uint2*)(base + L.rects), (uint32_t*)(base + L.counts),
                     (uint32_t*)(base + L.dqw), (uint32_t*)(base + L.dqs), (uint32_t*)(base + L.header),
                     L.reset_words);
  return check_launch("rgp_bin_rects");
}

extern "C" int rgp_preprocess_weight(int n, int K, int mode, int P, int Q, int rank,
                                     const rgp_camera* cam, const float* pws, const float* rots,
                                     const float* scales, const float* alphas, const float* shs,
                                     const float* dc, const float* A, const float* B,
                                     float* depths, float* records, void* ws, int64_t capacity,
                                     int width, int height, float* us, float* cinv2ds,
                                     int32_t* areas, float* weights, hipStream_t s) {
  if (n < 0 || K < 1 || K > RGP_MAX_DIM || width <= 0 || height <= 0)
    return set_error(RGP_ERR_ARG, "rgp_preprocess_weight: bad size or K");
  if (mode == 1) {
    int rc = check_lowrank("rgp_preprocess_weight", K, P, Q, rank);
    if (rc) return rc;
  } else if (mode != 0 && mode != 2) {
    return set_error(RGP_ERR_ARG, "rgp_preprocess_weight: mode must be 0, 1 or 2");
  }
  // modes 0 and 2: P carries the number of SH coefficients per channel
  const int sh_dim = mode != 1 ? P : 16;
  if (mode == 2 && sh_dim > 1 && !shs) return set_error(RGP_ERR_ARG, "rgp_preprocess_weight: mode 2 needs shs (higher bands)");
  if (mode != 1 && !(sh_dim == 1 || sh_dim == 4 || sh_dim == 9 || sh_dim == 16))
    return set_error(RGP_ERR_ARG, "rgp_preprocess_weight: sh_dim (P) must be 1, 4, 9 or 16");
  BinLayout L = bin_layout(n, capacity, width, height);
  char* base = (char*)ws;
  if (n == 0) return bin_reset_header(ws, L.reset_words, s);
  int wmode = mode;
  if (mode == 1 && Q == 15 && P == K && (rank == 1 || rank == 2)) wmode = rank + 1;
  if (mode == 0 && sh_dim == 16 && ((uintptr_t)shs & 15) == 0) wmode = 4;
  if (mode == 2) wmode = (sh_dim == 16 && ((uintptr_t)shs & 15) == 0) ? 5 : 6;
#define RGP_PRE(WM)                                                                               \
  RGP_K_SWITCH(K, hipLaunchKernelGGL((k_preprocess<KK, WM>), dim3((n + 63) / 64), dim3(64), 0, s, n, \
                                     sh_dim, P, Q, rank, cam, pws, (const float4*)rots, scales,   \
                                     alphas, shs, dc, A, B, depths, (float4*)records,             \
                                     (uint2*)(base + L.rects), (uint32_t*)(base + L.counts),      \
                                     (uint32_t*)(base + L.dqw), (uint32_t*)(base + L.dqs),        \
                                     (uint32_t*)(base + L.header), L.reset_words, L.gx, L.gy, us, \
                                     cinv2ds, areas, weights))
  switch (wmode) {
    case 0: RGP_PRE(0); break;
    case 1: RGP_PRE(1); break;
    case 2: RGP_PRE(2); break;
    case 3: RGP_PRE(3); break;
    case 5: RGP_PRE(5); break;
    case 6: RGP_PRE(6); break;
    default: RGP_PRE(4); break;
  }
#undef RGP_PRE
  return check_launch("rgp_preprocess_weight");
}

extern "C" int rgp_preprocess_weight_views(int views, int n, int K, int mode, int P, int Q, int rank,
                                           const rgp_camera* cams, const float* pws, const float* rots,
                                           const float* scales, const float* alphas, const float* shs,
                                           const float* dc, const float* A, const float* B, float* depths,
                                           float* records, void* ws, size_t ws_stride, int64_t capacity,
                                           int width, int height, hipStream_t s) {
  if (views < 1 || views > RGP_MAX_VIEWS || n < 0 || K < 1 || K > RGP_MAX_DIM || width <= 0 || height <= 0)
    return set_error(RGP_ERR_ARG, "rgp_preprocess_weight_views: bad views, size or K");
  if (mode == 1) {
    int rc = check_lowrank("rgp_preprocess_weight_views", K, P, Q, rank);
    if (rc) return rc;
  } else if (mode != 0 && mode != 2) {
    return set_error(RGP_ERR_ARG, "rgp_preprocess_weight_views: mode must be 0, 1 or 2");
  }
  const int sh_dim = mode != 1 ? P : 16;
  if (mode == 2 && sh_dim > 1 && !shs) return set_error(RGP_ERR_ARG, "rgp_preprocess_weight_views: mode 2 needs shs (higher bands)");
  if (mode != 1 && !(sh_dim == 1 || sh_dim == 4 || sh_dim == 9 || sh_dim == 16))
    return set_error(RGP_ERR_ARG, "rgp_preprocess_weight_views: sh_dim (P) must be 1, 4, 9 or 16");
  BinLayout L = bin_layout(n, capacity, width, height);
  if (ws_stride < L.total || (ws_stride & 255) != 0)
    return set_error(RGP_ERR_ARG, "rgp_preprocess_weight_views: ws_stride %zu below the workspace size %zu "
                     "or not a multiple of 256", ws_stride, L.total);
  char* base = (char*)ws;
  if (n == 0) {
    for (int v = 0; v < views; ++v) {
      int rc = bin_reset_header(base + (size_t)v * ws_stride, L.reset_words, s);
      if (rc) return rc;
    }
    return RGP_OK;
  }
  int wmode = mode;
  if (mode == 1 && Q == 15 && P == K && (rank == 1 || rank == 2)) wmode = rank + 1;
  if (mode == 0 && sh_dim == 16 && ((uintptr_t)shs & 15) == 0) wmode = 4;
  if (mode == 2) wmode = (sh_dim == 16 && ((uintptr_t)shs & 15) == 0) ? 5 : 6;
#define RGP_PREV(WM)                                                                                    \
  RGP_K_SWITCH(K, hipLaunchKernelGGL((k_preprocess_views<KK, WM>), dim3((n + 63) / 64), dim3(64), 0, s,  \
                                     views, n, sh_dim, P, Q, rank, cams, pws, (const float4*)rots, scales, \
                                     alphas, shs, dc, A, B, depths, (float4*)records, base, ws_stride,    \
                                     L.rects, L.counts, L.dqw, L.dqs, L.reset_words, L.gx, L.gy))
  switch (wmode) {
    case 0: RGP_PREV(0); break;
    case 1: RGP_PREV(1); break;
    case 2: RGP_PREV(2); break;
    case 3: RGP_PREV(3); break;
    case 5: RGP_PREV(5); break;
    case 6: RGP_PREV(6); break;
    default: RGP_PREV(4); break;
  }
#undef RGP_PREV
  return check_launch("rgp_preprocess_weight_views");
}

extern "C" int rgp_sh_backward_full(int n, int K, int sh_dim, const rgp_camera* cam,
                                    const float* pws, const float* dw, float* dshs,
                                    hipStream_t s) {
  if (n < 0 || K < 1 || K > RGP_MAX_DIM)
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_full: bad n or K");
  if (!(sh_dim == 1 || sh_dim == 4 || sh_dim == 9 || sh_dim == 16))
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_full: sh_dim must be 1, 4, 9 or 16");
  if (n == 0) return RGP_OK;
  const int blocks = (n + 127) / 128;
  RGP_K_SWITCH(K, hipLaunchKernelGGL(k_sh_backward_full<KK>, dim3(blocks), dim3(256), 0, s, n,
                                     sh_dim, cam, pws, dw, dshs));
  return check_launch("rgp_sh_backward_full");
}

// Adam scalars as torch.optim.Adam forms them: in double from the double
// hyper-parameters, then rounded once to float.
static AdamArgs adam_args(int64_t step, double lr, double beta1, double beta2, double eps, double wd) {
  const double bc1 = 1.0 - std::pow(beta1, (double)step);
  const double bc2 = 1.0 - std::pow(beta2, (double)step);
  AdamArgs a;
  a.neg_step = (float)(-lr / bc1);
  a.omb1 = (float)(1.0 - beta1);
  a.b2 = (float)beta2;
  a.omb2 = (float)(1.0 - beta2);
  a.bc2_sqrt = (float)std::sqrt(bc2);
  a.eps = (float)eps;
  a.wd = (float)wd;
  return a;
}

namespace rgp {
// One torch.optim.Adam step over a flat tensor with its gradient: one pass,
// float4 when the four arrays are 16-byte aligned.
__global__ __launch_bounds__(256) void k_adam4(int64_t n4, float4* __restrict__ p, const float4* __restrict__ g,
                                               float4* __restrict__ m, float4* __restrict__ v, AdamArgs a) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    float4 pp = ld_nt(&p[i]), mm = ld_nt(&m[i]), vv = ld_nt(&v[i]);
    const float4 gg = ld_nt(&g[i]);
    adam_elem(pp.x, mm.x, vv.x, gg.x, a);
    adam_elem(pp.y, mm.y, vv.y, gg.y, a);
    adam_elem(pp.z, mm.z, vv.z, gg.z, a);
    adam_elem(pp.w, mm.w, vv.w, gg.w, a);
    st_nt(&p[i], pp);
    st_nt(&m[i], mm);
    st_nt(&v[i], vv);
  }
}

__global__ __launch_bounds__(256) void k_adam1(int64_t lo, int64_t n, float* __restrict__ p,
                                               const float* __restrict__ g, float* __restrict__ m,
                                               float* __restrict__ v, AdamArgs a) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    adam_elem(p[i], m[i], v[i], g[i], a);
}
}  // namespace rgp

namespace rgp {
// The vanilla training step's backward chain (vanilla_3dgs/gsplat/gsmodel.py:
// 59-73) for one Gaussian per thread: the reference's batched products, each
// accumulated from 0 in index order.
//   dcov2d = dcinv J_inv;  dcov3d = dcov2d J_c2;  drot = dcov3d J_rot;
//   dscale = dcov3d J_scale;  dshs[l, c] = dcolor[c] dcolor_dshs[l];
//   dpw = ((du J_u) R + dcolor J_colpw) + (dcov2d J_c2pc) R
template <int I, int J>
__device__ __forceinline__ void vec_mat(const float* v, const float* __restrict__ M, float* out) {
#pragma unroll
  for (int j = 0; j < J; ++j) {
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < I; ++i) acc = acc + v[i] * M[i * J + j];
    out[j] = acc;
  }
}

__global__ __launch_bounds__(256) void k_vanilla_chain(
    int n, int sh_dim, const float* __restrict__ Rcw, const float* __restrict__ dus,
    const float* __restrict__ dcinv, const float* __restrict__ dcolors, const float* __restrict__ du_dpcs,
    const float* __restrict__ dcinv_dcov2d, const float* __restrict__ dcov2d_dcov3d,
    const float* __restrict__ dcov2d_dpc, const float* __restrict__ dcov3d_drot,
    const float* __restrict__ dcov3d_dscale, const float* __restrict__ dcolor_dshs,
    const float* __restrict__ dcolor_dpw, float* __restrict__ dpws, float* __restrict__ dshs,
    float* __restrict__ drots, float* __restrict__ dscales) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float R[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) R[e] = Rcw[e];
  const float du[2] = {dus[2 * i], dus[2 * i + 1]};
  const float dci[3] = {dcinv[3 * i], dcinv[3 * i + 1], dcinv[3 * i + 2]};
  const float dco[3] = {dcolors[3 * i], dcolors[3 * i + 1], dcolors[3 * i + 2]};
  float dc2[3], dc3[6], dr[4], dsc[3];
  vec_mat<3, 3>(dci, dcinv_dcov2d + (size_t)i * 9, dc2);
  vec_mat<3, 6>(dc2, dcov2d_dcov3d + (size_t)i * 18, dc3);
  vec_mat<6, 4>(dc3, dcov3d_drot + (size_t)i * 24, dr);
  vec_mat<6, 3>(dc3, dcov3d_dscale + (size_t)i * 18, dsc);
#pragma unroll
  for (int e = 0; e < 4; ++e) drots[(size_t)i * 4 + e] = dr[e];
#pragma unroll
  for (int e = 0; e < 3; ++e) dscales[(size_t)i * 3 + e] = dsc[e];
  for (int l = 0; l < sh_dim; ++l) {
    const float y = dcolor_dshs[(size_t)i * sh_dim + l];
#pragma unroll
    for (int c = 0; c < 3; ++c) dshs[((size_t)i * sh_dim + l) * 3 + c] = dco[c] * y;
  }
  float a[3], b[3], c[3], d[3], e3[3];
  vec_mat<2, 3>(du, du_dpcs + (size_t)i * 6, a);
  vec_mat<3, 3>(a, R, b);
  vec_mat<3, 3>(dco, dcolor_dpw + (size_t)i * 9, c);
  vec_mat<3, 3>(dc2, dcov2d_dpc + (size_t)i * 9, d);
  vec_mat<3, 3>(d, R, e3);
#pragma unroll
  for (int k = 0; k < 3; ++k) dpws[(size_t)i * 3 + k] = (b[k] + c[k]) + e3[k];
}
}  // namespace rgp

extern "C" int rgp_vanilla_chain(int n, int sh_dim, const float* Rcw, const float* dus, const float* dcinv,
                                 const float* dcolors, const float* du_dpcs, const float* dcinv_dcov2d,
                                 const float* dcov2d_dcov3d, const float* dcov2d_dpc, const float* dcov3d_drot,
                                 const float* dcov3d_dscale, const float* dcolor_dshs, const float* dcolor_dpw,
                                 float* dpws, float* dshs, float* drots, float* dscales, hipStream_t s) {
  if (n < 0 || sh_dim < 1 || sh_dim > 16) return set_error(RGP_ERR_ARG, "rgp_vanilla_chain: bad n or sh_dim");
  if (n == 0) return RGP_OK;
  hipLaunchKernelGGL(k_vanilla_chain, dim3(grid1(n)), dim3(256), 0, s, n, sh_dim, Rcw, dus, dcinv, dcolors,
                     du_dpcs, dcinv_dcov2d, dcov2d_dcov3d, dcov2d_dpc, dcov3d_drot, dcov3d_dscale, dcolor_dshs,
                     dcolor_dpw, dpws, dshs, drots, dscales);
  return check_launch("rgp_vanilla_chain");
}

extern "C" int rgp_adam(int64_t n, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                        int64_t step, double lr, double beta1, double beta2, double eps, double weight_decay,
                        hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_adam: n < 0");
  if (step < 1) return set_error(RGP_ERR_ARG, "rgp_adam: step must be >= 1");
  if (n == 0) return RGP_OK;
  const AdamArgs a = adam_args(step, lr, beta1, beta2, eps, weight_decay);
  const bool al = ((((uintptr_t)param) | ((uintptr_t)grad) | ((uintptr_t)exp_avg) | ((uintptr_t)exp_avg_sq)) & 15) == 0;
  const int64_t n4 = al ? n >> 2 : 0;
  auto grid = [](int64_t work) {
    const int64_t b = (work + 255) / 256;
    return (int)(b < 1 ? 1 : b > 8192 ? 8192 : b);
  };
  if (n4 > 0)
    hipLaunchKernelGGL(k_adam4, dim3(grid(n4)), dim3(256), 0, s, n4, (float4*)param, (const float4*)grad,
                       (float4*)exp_avg, (float4*)exp_avg_sq, a);
  if (4 * n4 < n)
    hipLaunchKernelGGL(k_adam1, dim3(grid(n - 4 * n4)), dim3(256), 0, s, 4 * n4, n, param, grad, exp_avg,
                       exp_avg_sq, a);
  return check_launch("rgp_adam");
}

extern "C" int rgp_sh_backward_adam_full(int n, int K, int sh_dim, const rgp_camera* cam,
                                         const float* pws, const float* dw, float* shs, float* exp_avg,
                                         float* exp_avg_sq, int64_t step, double lr, double beta1,
                                         double beta2, double eps, double weight_decay, hipStream_t s) {
  if (n < 0 || K < 1 || K > RGP_MAX_DIM)
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_full: bad n or K");
  if (!(sh_dim == 1 || sh_dim == 4 || sh_dim == 9 || sh_dim == 16))
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_full: sh_dim must be 1, 4, 9 or 16");
  if (step < 1) return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_full: step must be >= 1");
  if ((((uintptr_t)shs) | ((uintptr_t)exp_avg) | ((uintptr_t)exp_avg_sq)) & 15)
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_full: shs and moments must be 16-byte aligned");
  if (n == 0) return RGP_OK;
  const AdamArgs a = adam_args(step, lr, beta1, beta2, eps, weight_decay);
  const int blocks = (n + 63) / 64;
#define RGP_ADAM(SH)                                                                              \
  RGP_K_SWITCH(K, hipLaunchKernelGGL((k_sh_adam_full<KK, SH>), dim3(blocks), dim3(256), 0, s, n, cam, \
                                     pws, dw, shs, exp_avg, exp_avg_sq, a))
  switch (sh_dim) {
    case 1: RGP_ADAM(1); break;
    case 4: RGP_ADAM(4); break;
    case 9: RGP_ADAM(9); break;
    default: RGP_ADAM(16); break;
  }
#undef RGP_ADAM
  return check_launch("rgp_sh_backward_adam_full");
}

extern "C" int rgp_sh_backward_adam_split(int n, int K, const rgp_camera* cam, const float* pws, const float* dw,
                                          float* low, float* low_exp_avg, float* low_exp_avg_sq, float* high,
                                          float* high_exp_avg, float* high_exp_avg_sq, int64_t step, double lr,
                                          double beta1, double beta2, double eps, double weight_decay,
                                          hipStream_t s) {
  if (n < 0 || K < 1 || K > RGP_MAX_DIM)
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_split: bad n or K");
  if (step < 1) return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_split: step must be >= 1");
  if ((((uintptr_t)low) | ((uintptr_t)low_exp_avg) | ((uintptr_t)low_exp_avg_sq) | ((uintptr_t)high) |
       ((uintptr_t)high_exp_avg) | ((uintptr_t)high_exp_avg_sq)) & 15)
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_split: leaves and moments must be 16-byte aligned");
  if (n == 0) return RGP_OK;
  const AdamArgs a = adam_args(step, lr, beta1, beta2, eps, weight_decay);
  const int blocks = (n + 63) / 64;
  RGP_K_SWITCH(K, hipLaunchKernelGGL((k_sh_adam_bands<KK, 0, 1>), dim3(blocks), dim3(256), 0, s, n, cam, pws, dw,
                                     low, low_exp_avg, low_exp_avg_sq, a));
  RGP_K_SWITCH(K, hipLaunchKernelGGL((k_sh_adam_bands<KK, 1, 15>), dim3(blocks), dim3(256), 0, s, n, cam, pws, dw,
                                     high, high_exp_avg, high_exp_avg_sq, a));
  return check_launch("rgp_sh_backward_adam_split");
}

extern "C" int rgp_sh_backward_adam_lowrank(int n, int K, int P, int Q, int rank, const rgp_camera* cam,
                                            const float* pws, const float* dw, float* dc, float* dc_exp_avg,
                                            float* dc_exp_avg_sq, float* A, float* A_exp_avg, float* A_exp_avg_sq,
                                            float* B, float* B_exp_avg, float* B_exp_avg_sq, int64_t step,
                                            double lr, double beta1, double beta2, double eps, double weight_decay,
                                            hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_lowrank: n < 0");
  int rc = check_lowrank("rgp_sh_backward_adam_lowrank", K, P, Q, rank);
  if (rc) return rc;
  if (!(Q == 15 && P == K && (rank == 1 || rank == 2)))
    return set_error(RGP_ERR_ARG,
                     "rgp_sh_backward_adam_lowrank: needs Q = 15, P = K and rank 1 or 2 (else "
                     "rgp_sh_backward_lowrank + rgp_adam)");
  if (step < 1) return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_lowrank: step must be >= 1");
  if ((((uintptr_t)dc) | ((uintptr_t)dc_exp_avg) | ((uintptr_t)dc_exp_avg_sq) | ((uintptr_t)A) |
       ((uintptr_t)A_exp_avg) | ((uintptr_t)A_exp_avg_sq) | ((uintptr_t)B) | ((uintptr_t)B_exp_avg) |
       ((uintptr_t)B_exp_avg_sq)) & 15)
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_lowrank: tensors and moments must be 16-byte aligned");
  if (n == 0) return RGP_OK;
  const AdamArgs a = adam_args(step, lr, beta1, beta2, eps, weight_decay);
  const int blocks = (n + 63) / 64;
  if (rank == 1) {
    RGP_K_SWITCH(K, hipLaunchKernelGGL((k_lowrank_adam_q15<KK, 1>), dim3(blocks), dim3(64), 0, s, n, cam, pws, dw,
                                       dc, dc_exp_avg, dc_exp_avg_sq, A, A_exp_avg, A_exp_avg_sq, B, B_exp_avg,
                                       B_exp_avg_sq, a));
  } else {
    RGP_K_SWITCH(K, hipLaunchKernelGGL((k_lowrank_adam_q15<KK, 2>), dim3(blocks), dim3(64), 0, s, n, cam, pws, dw,
                                       dc, dc_exp_avg, dc_exp_avg_sq, A, A_exp_avg, A_exp_avg_sq, B, B_exp_avg,
                                       B_exp_avg_sq, a));
  }
  return check_launch("rgp_sh_backward_adam_lowrank");
}

extern "C" int rgp_sh_backward_adam_lowrank_views(int views, int n, int K, int P, int Q, int rank,
                                                  const rgp_camera* cams, const float* pws, const float* dw,
                                                  float* dc, float* dc_exp_avg, float* dc_exp_avg_sq, float* A,
                                                  float* A_exp_avg, float* A_exp_avg_sq, float* B, float* B_exp_avg,
                                                  float* B_exp_avg_sq, int64_t step, double lr, double beta1,
                                                  double beta2, double eps, double weight_decay, hipStream_t s) {
  if (views < 1 || views > RGP_MAX_VIEWS || n < 0)
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_lowrank_views: bad views or n");
  int rc = check_lowrank("rgp_sh_backward_adam_lowrank_views", K, P, Q, rank);
  if (rc) return rc;
  if (!(Q == 15 && P == K && (rank == 1 || rank == 2)))
    return set_error(RGP_ERR_ARG,
                     "rgp_sh_backward_adam_lowrank_views: needs Q = 15, P = K and rank 1 or 2 (else "
                     "rgp_sh_backward_lowrank_views + rgp_adam)");
  if (step < 1) return set_error(RGP_ERR_ARG, "rgp_sh_backward_adam_lowrank_views: step must be >= 1");
  if ((((uintptr_t)dc) | ((uintptr_t)dc_exp_avg) | ((uintptr_t)dc_exp_avg_sq) | ((uintptr_t)A) |
       ((uintptr_t)A_exp_avg) | ((uintptr_t)A_exp_avg_sq) | ((uintptr_t)B) | ((uintptr_t)B_exp_avg) |
       ((uintptr_t)B_exp_avg_sq)) & 15)
    return set_error(RGP_ERR_ARG,
                     "rgp_sh_backward_adam_lowrank_views: tensors and moments must be 16-byte aligned");
  if (n == 0) return RGP_OK;
  const AdamArgs a = adam_args(step, lr, beta1, beta2, eps, weight_decay);
  const int blocks = (n + 63) / 64;
  if (rank == 1) {
    RGP_K_SWITCH(K, hipLaunchKernelGGL((k_lowrank_adam_q15_views<KK, 1>), dim3(blocks), dim3(64), 0, s, views, n,
                                       cams, pws, dw, dc, dc_exp_avg, dc_exp_avg_sq, A, A_exp_avg, A_exp_avg_sq, B,
                                       B_exp_avg, B_exp_avg_sq, a));
  } else {
    RGP_K_SWITCH(K, hipLaunchKernelGGL((k_lowrank_adam_q15_views<KK, 2>), dim3(blocks), dim3(64), 0, s, views, n,
                                       cams, pws, dw, dc, dc_exp_avg, dc_exp_avg_sq, A, A_exp_avg, A_exp_avg_sq, B,
                                       B_exp_avg, B_exp_avg_sq, a));
  }
  return check_launch("rgp_sh_backward_adam_lowrank_views");
}

extern "C" int rgp_sh_backward_lowrank(int n, int K, int P, int Q, int rank,
                                       const rgp_camera* cam, const float* pws, const float* A,
                                       const float* B, const float* dw, float* ddc, float* dA,
                                       float* dB, hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_sh_backward_lowrank: n < 0");
  int rc = check_lowrank("rgp_sh_backward_lowrank", K, P, Q, rank);
  if (rc) return rc;
  if (n == 0) return RGP_OK;
  // cam is a device pointer: pass the address of its twc field
  const float* twc = (const float*)((const char*)cam + offsetof(rgp_camera, twc));
  rgp::launch_lowrank_backward(n, K, P, Q, rank, A, B, nullptr, pws, twc, dw, ddc, dA, dB, s);
  return check_launch("rgp_sh_backward_lowrank");
}

extern "C" int rgp_sh_backward_lowrank_views(int views, int n, int K, int P, int Q, int rank,
                                             const rgp_camera* cams, const float* pws, const float* A,
                                             const float* B, const float* dw, float* ddc, float* dA, float* dB,
                                             hipStream_t s) {
  if (views < 1 || views > RGP_MAX_VIEWS || n < 0)
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_lowrank_views: bad views or n");
  int rc = check_lowrank("rgp_sh_backward_lowrank_views", K, P, Q, rank);
  if (rc) return rc;
  if (n == 0) return RGP_OK;
  if (Q == 15 && P == K && rank == 1) {
    RGP_K_SWITCH(K, hipLaunchKernelGGL((k_lowrank_backward_q15_views<KK, 1>), dim3(grid1(n)), dim3(256), 0, s,
                                       views, n, A, B, pws, cams, dw, ddc, dA, dB));
  } else if (Q == 15 && P == K && rank == 2) {
    RGP_K_SWITCH(K, hipLaunchKernelGGL((k_lowrank_backward_q15_views<KK, 2>), dim3(grid1(n)), dim3(256), 0, s,
                                       views, n, A, B, pws, cams, dw, ddc, dA, dB));
  } else {
    hipLaunchKernelGGL(k_lowrank_backward_views, dim3(grid1(n)), dim3(256), 0, s, views, n, K, P, Q, rank, A, B,
                       pws, cams, dw, ddc, dA, dB);
  }
  return check_launch("rgp_sh_backward_lowrank_views");
}

extern "C" int rgp_sh_backward_full_views(int views, int n, int K, int sh_dim, const rgp_camera* cams,
                                          const float* pws, const float* dw, float* dshs, hipStream_t s) {
  if (views < 1 || views > RGP_MAX_VIEWS || n < 0 || K < 1 || K > RGP_MAX_DIM)
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_full_views: bad views, n or K");
  if (!(sh_dim == 1 || sh_dim == 4 || sh_dim == 9 || sh_dim == 16))
    return set_error(RGP_ERR_ARG, "rgp_sh_backward_full_views: sh_dim must be 1, 4, 9 or 16");
  if (n == 0) return RGP_OK;
  RGP_K_SWITCH(K, hipLaunchKernelGGL(k_sh_backward_full_views<KK>, dim3(grid1(n)), dim3(256), 0, s, views, n,
                                     sh_dim, cams, pws, dw, dshs));
  return check_launch("rgp_sh_backward_full_views");
}
