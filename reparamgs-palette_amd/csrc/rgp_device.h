// Device-side helpers shared by the gfx950 kernels of the palette-weight
// rasterizer.  All code here is compiled with -ffp-contract=off and correctly
// rounded division and square root, so every FP32 operation is rounded once,
// in source order: the arithmetic of the reference kernels
// (gsplatcu/kernel.cu) written out without implicit contraction; the fused
// multiply-adds that remain are explicit fmaf calls.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rgp.h"

namespace rgp {

constexpr int kTile = 16;               // BLOCK (common.cuh:13)
constexpr int kTilePix = kTile * kTile;  // BLOCK_SIZE
constexpr double kMinDepth = 0.2;       // MIN_DEPTH, a double literal (kernel.cu:10)
constexpr float kBadMarker = -1.f;      // BAD_MARKER (kernel.cu:11)
constexpr int kMaxDim = 8;              // MAX_SPLAT_DIM (kernel.cu:12)

// Real SH constants, degree <= 3 (common.cuh:29-44).  Kept in double: the
// reference multiplies them with floats in double precision.
constexpr double SH_C0_0 = 0.28209479177387814;
constexpr double SH_C1_0 = -0.4886025119029199;
constexpr double SH_C1_1 = 0.4886025119029199;
constexpr double SH_C1_2 = -0.4886025119029199;
constexpr double SH_C2_0 = 1.0925484305920792;
constexpr double SH_C2_1 = -1.0925484305920792;
constexpr double SH_C2_2 = 0.31539156525252005;
constexpr double SH_C2_3 = -1.0925484305920792;
constexpr double SH_C2_4 = 0.5462742152960396;
constexpr double SH_C3_0 = -0.5900435899266435;
constexpr double SH_C3_1 = 2.890611442640554;
constexpr double SH_C3_2 = -0.4570457994644658;
constexpr double SH_C3_3 = 0.3731763325901154;
constexpr double SH_C3_4 = -0.4570457994644658;
constexpr double SH_C3_5 = 1.445305721320277;
constexpr double SH_C3_6 = -0.5900435899266435;

__device__ __forceinline__ float dmul(double c, float x) { return (float)(c * (double)x); }
__device__ __forceinline__ float dmul2(double c, float x, float y) {
  return (float)((c * (double)x) * (double)y);
}

// exp(x): Cody-Waite reduction x = n ln2 + r (n the exact product x log2e
// rounded once to the nearest integer, as kExpRound + n by one fused
// multiply-add; ln2 split in a 16-bit head and a tail, |r| <= ln2/2), e^r by
// a degree-5 polynomial 1 + r + c2 r^2 + ... +
// c5 r^5 whose c2..c5 minimise the largest relative error on that interval
// with the first two coefficients held at 1 (9.1e-8, found by linear
// programming on a dense grid), Horner form with fused multiply-adds, then
// exact scaling by 2^n.  Max error 2 ulp against the correctly rounded exp
// over every float with a normal result (the bound CUDA documents for its
// expf); identical results on any IEEE-754 machine with fma.
constexpr float kExpC2 = 0x1.fffddcp-2f, kExpC3 = 0x1.5557acp-3f, kExpC4 = 0x1.57347cp-5f,
                kExpC5 = 0x1.106baap-7f;
// 1.5 * 2^23: for |y| < 2^22, fl(y + kExpRound) lies in [2^23, 2^24), where
// floats are the integers, so it is kExpRound + (y rounded to nearest even).
constexpr float kExpRound = 0x1.8p23f;
__device__ __forceinline__ float exp_cw(float x) {
  if (!(x > -104.0f)) return (x != x) ? x : 0.0f;
  if (x > 88.72283f) return __uint_as_float(0x7f800000u);
  const float n = __builtin_fmaf(x, 1.44269504088896341f, kExpRound) - kExpRound;
  float r = __builtin_fmaf(n, -0.693145751953125f, x);
  r = __builtin_fmaf(n, -1.42860676533018570e-06f, r);
  float p = kExpC5;
  p = __builtin_fmaf(p, r, kExpC4);
  p = __builtin_fmaf(p, r, kExpC3);
  p = __builtin_fmaf(p, r, kExpC2);
  p = __builtin_fmaf(p, r, 1.0f);
  p = __builtin_fmaf(p, r, 1.0f);
  int ni = (int)n;
  if (ni < -126) { p = p * 5.42101086e-20f; ni += 64; }
  if (ni > 127) { p = p * 2.0f; ni -= 1; }
  return p * __uint_as_float((uint32_t)(ni + 127) << 23);
}

// exp_cw(-0.5f * m) for m in [0, 1e7], formed from m without the scaling
// and without the range handling: n is the same (m (-0.5 log2e) and
// (-0.5 m) log2e are the same real product), the reduced argument is -2r,
// and the Horner partials with the coefficients scaled by (-1/2)^k are the
// original ones scaled by (-1/2)^k.  Scaling by a power of two commutes with
// rounding for these normal values, so the result is bit-identical (checked
// on every float of [0, 1e7], subnormal results included); the final 2^n
// scaling is an ldexp.  m is capped at 1000 (exp(-m/2) is already 0 from
// m = 208 on) so that m log2e / 2 stays within kExpRound's range.
__device__ __forceinline__ float gauss_half(float m) {
  const float mm = fminf(m, 1000.0f);
  const float n = __builtin_fmaf(mm, -0.5f * 1.44269504088896341f, kExpRound) - kExpRound;
  float r = __builtin_fmaf(n, 2.0f * 0.693145751953125f, mm);
  r = __builtin_fmaf(n, 2.0f * 1.42860676533018570e-06f, r);
  float t = kExpC5 * -0.03125f;
  t = __builtin_fmaf(t, r, kExpC4 * 0.0625f);
  t = __builtin_fmaf(t, r, kExpC3 * -0.125f);
  t = __builtin_fmaf(t, r, kExpC2 * 0.25f);
  t = __builtin_fmaf(t, r, -0.5f);
  t = __builtin_fmaf(t, r, 1.0f);
  return __builtin_ldexpf(t, (int)n);
}

// gauss_half(m) for m in [0, 170], where 2^n t is a normal float (n in
// [-123, 0], t in [0.70, 1.42]): the scaling adds n to t's exponent field,
// and n << 23 is the bit pattern of fl(m (-log2e/2) + kExpRound) shifted
// left by 23 (kExpRound's own bits vanish from the word: their low nine bits
// are zero), one integer shift-add in place of the conversion and the ldexp.
// Bit-identical to gauss_half on every float of [0, 170] (checked by the
// oracle).  Records with a larger cutoff take gauss_half.
__device__ __forceinline__ float gauss_half_fast(float m) {
  const float s = __builtin_fmaf(m, -0.5f * 1.44269504088896341f, kExpRound);
  const float n = s - kExpRound;
  float r = __builtin_fmaf(n, 2.0f * 0.693145751953125f, m);
  r = __builtin_fmaf(n, 2.0f * 1.42860676533018570e-06f, r);
  float t = kExpC5 * -0.03125f;
  t = __builtin_fmaf(t, r, kExpC4 * 0.0625f);
  t = __builtin_fmaf(t, r, kExpC3 * -0.125f);
  t = __builtin_fmaf(t, r, kExpC2 * 0.25f);
  t = __builtin_fmaf(t, r, -0.5f);
  t = __builtin_fmaf(t, r, 1.0f);
  return __uint_as_float(__float_as_uint(t) + (__float_as_uint(s) << 23));
}

// log(1 + y) for y >= 0 (an exp output).  u = fl(1 + y) with its exact
// rounding error e (Fast2Sum); log u = n ln2 + log m with m in [sqrt(1/2),
// sqrt(2)) and log m = 2 atanh(s) = 2 (s + s^3/3 + s^5/5 + s^7/7 + s^9/9),
// s = (m - 1)/(m + 1), |s| < 0.172; log(1 + y) = log u + e / u.  Within
// 4 ulp of log1p over [0, e^20].
__device__ __forceinline__ float log1p_pos(float y) {
  if (y != y) return y;
  if (y == __uint_as_float(0x7f800000u)) return y;
  const float u = 1.0f + y;
  const float e = y >= 1.0f ? 1.0f - (u - y) : y - (u - 1.0f);
  const uint32_t b = __float_as_uint(u);
  int n = (int)(b >> 23) - 127;
  float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421356f) { m = 0.5f * m; n += 1; }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  float p = 1.0f / 9.0f;
  p = __builtin_fmaf(p, z, 1.0f / 7.0f);
  p = __builtin_fmaf(p, z, 0.2f);
  p = __builtin_fmaf(p, z, 1.0f / 3.0f);
  const float t = 2.0f * s;
  const float lm = __builtin_fmaf(t * z, p, t);
  const float fn = (float)n;
  const float lu = __builtin_fmaf(fn, 0.693145751953125f, __builtin_fmaf(fn, 1.42860676533018570e-06f, lm));
  return lu + e / u;
}

// exp(-m/2) for a clamped Mahalanobis distance m >= 0.
__device__ __forceinline__ float gauss_of(float m) {
  return m <= 170.0f ? gauss_half(m) : exp_cw(-0.5f * m);
}

// Float -> int conversions that saturate and send NaN to 0.
__device__ __forceinline__ int32_t sat_i32(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x <= -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}
__device__ __forceinline__ uint32_t sat_u32(float x) {
  if (x != x) return 0u;
  if (x >= 4294967296.0f) return UINT32_MAX;
  if (x <= 0.0f) return 0u;
  return (uint32_t)x;
}

// ---------------------------------------------------------------------------
// per-Gaussian geometry (kernel.cu:379-663)
// ---------------------------------------------------------------------------
struct Proj {
  float ux, uy, x, y, z;
  bool ok;
};

// project (kernel.cu:635-654): pc = R pw + t, accumulated from 0 as in
// Matrix::operator* (matrix.cuh:53-70).
__device__ __forceinline__ Proj project_one(const float* R, const float* t, float px, float py,
                                            float pz, float fx, float fy, float cx, float cy) {
  const float pw[3] = {px, py, pz};
  float pc[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) s = s + R[r * 3 + k] * pw[k];
    pc[r] = s + t[r];
  }
  Proj o;
  o.x = pc[0]; o.y = pc[1]; o.z = pc[2];
  o.ok = !((double)o.z < kMinDepth);
  const float z_inv = 1.f / o.z;
  o.ux = (o.x * fx) * z_inv + cx;
  o.uy = (o.y * fy) * z_inv + cy;
  return o;
}

// computeCov3D (kernel.cu:435-461) -> (S00, S01, S02, S11, S12, S22)
__device__ __forceinline__ void cov3d_one(float w, float x, float y, float z, float s0, float s1,
                                          float s2, float* c) {
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
    for (int k = 0; k < 3; ++k) M[r * 3 + k] = R[r * 3 + k] * s[k];
  float S[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 3; ++k) acc = acc + M[r * 3 + k] * M[q * 3 + k];
      S[r * 3 + q] = acc;
    }
  c[0] = S[0]; c[1] = S[1]; c[2] = S[2]; c[3] = S[4]; c[4] = S[5]; c[5] = S[8];
}

// computeCov2D (kernel.cu:526-555) -> (a, b, c) with the 0.3 low-pass
__device__ __forceinline__ void cov2d_one(const float* c3, float x, float y, float z,
                                          const float* R, float fx, float fy, float tfx,
                                          float tfy, float* c2) {
  const float a = c3[0], b = c3[1], c = c3[2], d = c3[3], e = c3[4], f = c3[5];
  const float limx = 1.3f * tfx;
  const float limy = 1.3f * tfy;
  x = fminf(limx, fmaxf(-limx, x / z)) * z;
  y = fminf(limy, fmaxf(-limy, y / z)) * z;
  const float z2 = z * z;
  const float J[6] = {fx / z, 0.0f, -(fx * x) / z2, 0.0f, fy / z, -(fy * y) / z2};
  const float S[9] = {a, b, c, b, d, e, c, e, f};
  float M[6];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 3; ++k) acc = acc + J[r * 3 + k] * R[k * 3 + q];
      M[r * 3 + q] = acc;
    }
  float MS[6];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 3; ++k) acc = acc + M[r * 3 + k] * S[k * 3 + q];
      MS[r * 3 + q] = acc;
    }
  float SP[4];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 3; ++k) acc = acc + MS[r * 3 + k] * M[q * 3 + k];
      SP[r * 2 + q] = acc;
    }
  c2[0] = SP[0] + 0.3f;
  c2[1] = SP[1];
  c2[2] = SP[3] + 0.3f;
}

// inverseCov2D (kernel.cu:391-400).  Returns false when the inverse
// determinant is NaN (the Gaussian is culled).
__device__ __forceinline__ bool inverse_one(float a, float b, float c, float* ci, int* ar) {
  const float det_inv = 1.f / (a * c - b * b);
  if (det_inv != det_inv) return false;
  ci[0] = det_inv * c;
  ci[1] = -det_inv * b;
  ci[2] = det_inv * a;
  ar[0] = sat_i32(ceilf(3.f * sqrtf(fabsf(a))));
  ar[1] = sat_i32(ceilf(3.f * sqrtf(fabsf(c))));
  return true;
}

// SH basis Y[1..15] of the view direction (kernel.cu:718-755, 924-956).
__device__ __forceinline__ void sh_basis(float px, float py, float pz, const float* twc, int sh_dim,
                                         float* Y) {
  const float d0 = px - twc[0], d1 = py - twc[1], d2 = pz - twc[2];
  const float normd_inv = 1.f / sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
  const float x = d0 * normd_inv, y = d1 * normd_inv, z = d2 * normd_inv;
  Y[1] = dmul(SH_C1_0, y);
  Y[2] = dmul(SH_C1_1, z);
  Y[3] = dmul(SH_C1_2, x);
  if (sh_dim > 4) {
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    Y[4] = dmul(SH_C2_0, xy);
    Y[5] = dmul(SH_C2_1, yz);
    Y[6] = dmul(SH_C2_2, 2.f * zz - xx - yy);
    Y[7] = dmul(SH_C2_3, xz);
    Y[8] = dmul(SH_C2_4, xx - yy);
    if (sh_dim > 9) {
      Y[9] = dmul2(SH_C3_0, y, 3.f * xx - yy);
      Y[10] = dmul2(SH_C3_1, xy, z);
      Y[11] = dmul2(SH_C3_2, y, 4.f * zz - xx - yy);
      Y[12] = dmul2(SH_C3_3, z, 2.f * zz - 3.f * xx - 3.f * yy);
      Y[13] = dmul2(SH_C3_4, x, 4.f * zz - xx - yy);
      Y[14] = dmul2(SH_C3_5, z, xx - yy);
      Y[15] = dmul2(SH_C3_6, x, xx - 3.f * yy);
    }
  }
}

// One Gaussian's coefficient row held as two leaves: the K band-0 values
// (registers) and the (sh_dim-1)*K higher-band values (LDS or global).
template <int K>
struct SplitRow {
  const float* lo;
  const float* hi;
  __device__ __forceinline__ float operator[](int j) const { return j < K ? lo[j] : hi[j - K]; }
};

// Full-rank weights (kernel.cu:712-767).  sh: band-major coefficients of one
// Gaussian, sh[lm * K + k], read through an accessor (global memory or a
// staged LDS row); the expression tree is the same for both.
template <int K, class Sh>
__device__ __forceinline__ void sh_weights_full(const Sh& sh, const float* Y, int sh_dim, float* w) {
#pragma unroll
  for (int k = 0; k < K; ++k) w[k] = (float)(0.0 + SH_C0_0 * (double)sh[k]);
  if (sh_dim > 1) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      w[k] = w[k] + ((Y[1] * sh[1 * K + k] + Y[2] * sh[2 * K + k]) + Y[3] * sh[3 * K + k]);
    if (sh_dim > 4) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        w[k] = w[k] + ((((Y[4] * sh[4 * K + k] + Y[5] * sh[5 * K + k]) + Y[6] * sh[6 * K + k]) +
                        Y[7] * sh[7 * K + k]) +
                       Y[8] * sh[8 * K + k]);
      if (sh_dim > 9) {
#pragma unroll
        for (int k = 0; k < K; ++k)
          w[k] = w[k] + ((((((Y[9] * sh[9 * K + k] + Y[10] * sh[10 * K + k]) +
                             Y[11] * sh[11 * K + k]) +
                            Y[12] * sh[12 * K + k]) +
                           Y[13] * sh[13 * K + k]) +
                          Y[14] * sh[14 * K + k]) +
                         Y[15] * sh[15 * K + k]);
      }
    }
  }
}

// Rank-r low-rank weights (kernel.cu:966-981 generalised): a (P, r), b (r, Q).
template <int K>
__device__ __forceinline__ void sh_weights_lowrank(const float* __restrict__ dc,
                                                   const float* __restrict__ a,
                                                   const float* __restrict__ b, const float* Y,
                                                   int L, int Q, int rank, float* w) {
#pragma unroll
  for (int k = 0; k < K; ++k) w[k] = dmul(SH_C0_0, dc[k]);
  if (L <= 0) return;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    float acc = 0.f;
    int g = k;
    int row = g / Q, col = g - row * Q;
    for (int lm = 0; lm < L; ++lm) {
      float c = a[row * rank] * b[col];
      for (int r = 1; r < rank; ++r) c = c + a[row * rank + r] * b[r * Q + col];
      acc = acc + c * Y[lm + 1];
      col += K;  // g += K
      while (col >= Q) { col -= Q; ++row; }
    }
    w[k] = w[k] + acc;
  }
}

// Rank-R low-rank weights for degree-3 SH (L = 15) with Q = 15, so P = K:
// the factors live in registers and every index is a compile-time constant.
// Same operations in the same order as sh_weights_lowrank.
template <int K, int R>
__device__ __forceinline__ void sh_weights_lowrank_q15(const float* __restrict__ dc,
                                                       const float* __restrict__ a,
                                                       const float* __restrict__ b,
                                                       const float* Y, float* w) {
  constexpr int Q = 15, P = K;
  float A[P * R], B[R * Q];
#pragma unroll
  for (int i = 0; i < P * R; ++i) A[i] = a[i];
#pragma unroll
  for (int i = 0; i < R * Q; ++i) B[i] = b[i];
#pragma unroll
  for (int k = 0; k < K; ++k) w[k] = dmul(SH_C0_0, dc[k]);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    float acc = 0.f;
#pragma unroll
    for (int lm = 0; lm < 15; ++lm) {
      const int g = lm * K + k, row = g / Q, col = g % Q;
      float c = A[row * R] * B[col];
#pragma unroll
      for (int r = 1; r < R; ++r) c = c + A[row * R + r] * B[r * Q + col];
      acc = acc + c * Y[lm + 1];
    }
    w[k] = w[k] + acc;
  }
}

// ---------------------------------------------------------------------------
// splat record: what the compositing kernels read per list entry.
//   f4[0] = (u.x, u.y, cinv.x, 2*cinv.y)   f4[1] = (cinv.z, alpha, mcut, ry)
//   f4[2..] = the dim feature channels (weights or colours), zero padded.
// 2*cinv.y is exact, so the Mahalanobis distance below is bit-identical to
// the reference expression.  mcut bounds the distance beyond which
// alpha * exp(-maha/2) < 0.002 with a margin far above rounding, so a skip on
// maha > mcut never changes a result; NaN (alpha NaN) disables the skip.
// ry bounds |dy| inside that region: min over dx of the quadratic form is
// dy^2 / cov_yy, so |dy| > sqrt(mcut * cov_yy) implies maha > mcut.  A
// wavefront whose pixel rows all lie farther than ry from u.y skips the entry.
// ---------------------------------------------------------------------------
// For finite alpha the cutoff is at most 1e7: alpha / 0.002 overflows to inf
// above 6.8e35, and beyond m = 208 exp(-m/2) is already 0 (alpha * 0 < 0.002),
// so the cap skips nothing that contributes.  Every distance that passes
// m <= mcut of a finite alpha therefore lies in [0, 1e7], where gauss_half(m)
// equals gauss_of(m) bit for bit (checked on every float of that interval by
// the oracle); inf and NaN alpha give alpha' = 0.99 whatever exp returns.
__device__ __forceinline__ float mcut_of(float alpha) {
  if (alpha != alpha) return alpha;  // NaN: never skip
  if (alpha < 0.002f) return -1.0f;  // alpha * e <= alpha < 0.002 for every e <= 1
  const float mc = 2.0f * logf(alpha / 0.002f) * 1.00001f + 1e-3f;
  return (alpha <= 3.40282347e38f && !(mc <= 1.0e7f)) ? 1.0e7f : mc;
}

__device__ __forceinline__ float ry_of(float mcut, float cxx, float cxy, float cyy) {
  const float inf = __uint_as_float(0x7f800000u);
  if (!(mcut >= 0.f)) return mcut < 0.f ? -1.f : inf;  // no contribution / NaN
  const double det = (double)cxx * cyy - (double)cxy * cxy;
  if (!(cxx > 0.f) || !(det > 0.0)) return inf;  // not positive definite: no bound
  const double cov_yy = (double)cxx / det;
  const double r = sqrt((double)mcut * cov_yy) * 1.0001 + 1e-3;
  return r < 3.0e38 ? (float)r : inf;
}

template <int DIM>
struct RecStride {
  static constexpr int f4 = 2 + (DIM + 3) / 4;
};

// Non-temporal 16-byte loads and stores for streams touched once (optimizer
// state): the stores go out to HBM during the kernel instead of leaving
// dirty lines in the caches for the next kernel to write back.
typedef float rgp_f4n __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_nt(const float4* a) {
  const rgp_f4n x = __builtin_nontemporal_load(reinterpret_cast<const rgp_f4n*>(a));
  return make_float4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ void st_nt(float4* a, float4 v) {
  const rgp_f4n x = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(x, reinterpret_cast<rgp_f4n*>(a));
}

// Blocks of one launch go round-robin to the 8 XCDs; give each XCD a
// contiguous run of tiles so neighbouring tiles share its L2.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
  const int q = nblocks >> 3, rem = nblocks & 7;
  const int x = b & 7, l = b >> 3;
  return x * q + (x < rem ? x : rem) + l;
}

// Sum over the 64 lanes of a wavefront with DPP row operations; the total
// lands in lane 63.
#define RGP_DPP(v, ctrl, rmask, bmask, bound)                                               \
  __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, rmask, bmask, bound))
__device__ __forceinline__ float wave_sum_to63(float v) {
  v = v + RGP_DPP(v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
  v = v + RGP_DPP(v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
  v = v + RGP_DPP(v, 0x124, 0xF, 0xF, false);  // row_ror:4
  v = v + RGP_DPP(v, 0x128, 0xF, 0xF, false);  // row_ror:8
  v = v + RGP_DPP(v, 0x142, 0xA, 0xF, false);  // row_bcast:15 into rows 1, 3
  v = v + RGP_DPP(v, 0x143, 0xC, 0xF, false);  // row_bcast:31 into rows 2, 3
  return v;
}

// Sum over the 16 lanes of each row; every lane of the row receives it.
__device__ __forceinline__ float row_sum16(float v) {
  v = v + RGP_DPP(v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
  v = v + RGP_DPP(v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
  v = v + RGP_DPP(v, 0x124, 0xF, 0xF, false);  // row_ror:4
  v = v + RGP_DPP(v, 0x128, 0xF, 0xF, false);  // row_ror:8
  return v;
}

// Reduce-scatter of N <= 16 per-lane values over the 16 lanes of each row.
// Level t (xor distance 1, 2, 4, 8) pairs lane l with l ^ d: of each pair of
// values the lane with bit d clear keeps the first and sends the second, its
// partner the reverse, so one level halves the values a lane holds (an odd
// last value is summed in both lanes).  After the four levels every lane
// holds the row sum of one channel, rs_channel(N, l & 15): about 3 operations
// per pair and level instead of 4 DPP adds per channel and a select.
// Exchanges: quad_perm for d = 1, 2; ds_swizzle (bit mode, xor 4) for d = 4;
// row_ror:8, its own inverse, for d = 8.
template <int D>
__device__ __forceinline__ float lane_xor(float v) {
  if constexpr (D == 1) return RGP_DPP(v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
  if constexpr (D == 2) return RGP_DPP(v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
  if constexpr (D == 4) return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x101F));
  if constexpr (D == 8) return RGP_DPP(v, 0x128, 0xF, 0xF, false);  // row_ror:8
}

template <int N, int D>
__device__ __forceinline__ void rs_level(const float* in, float* out, bool bit) {
#pragma unroll
  for (int i = 0; i < N / 2; ++i) {
    const float keep = bit ? in[2 * i + 1] : in[2 * i];
    const float send = bit ? in[2 * i] : in[2 * i + 1];
    out[i] = keep + lane_xor<D>(send);
  }
  if constexpr ((N & 1) != 0) out[N / 2] = in[N - 1] + lane_xor<D>(in[N - 1]);
}

template <int N>
__device__ __forceinline__ float row_reduce_scatter(const float* v, int lane) {
  static_assert(N >= 1 && N <= 16, "one channel per lane of a row");
  constexpr int N1 = (N + 1) / 2, N2 = (N1 + 1) / 2, N3 = (N2 + 1) / 2;
  float a[N1], b[N2], c[N3], d[1];
  rs_level<N, 1>(v, a, lane & 1);
  rs_level<N1, 2>(a, b, lane & 2);
  rs_level<N2, 4>(b, c, lane & 4);
  rs_level<N3, 8>(c, d, lane & 8);
  return d[0];
}

// The channel whose row sum row_reduce_scatter<n> leaves in lane r of a row,
// and whether r is the lowest lane of its row holding that channel.
__host__ __device__ constexpr int rs_channel(int n, int r) {
  int map[16] = {0};
  for (int i = 0; i < 16; ++i) map[i] = i;
  for (int t = 0; t < 4; ++t) {
    const int bit = (r >> t) & 1;
    const int h = (n + 1) / 2;
    for (int i = 0; i < n / 2; ++i) map[i] = map[2 * i + bit];
    if (n & 1) map[n / 2] = map[n - 1];
    n = h;
  }
  return map[0];
}
__host__ __device__ constexpr bool rs_owner(int n, int r) {
  for (int q = 0; q < r; ++q)
    if (rs_channel(n, q) == rs_channel(n, r)) return false;
  return true;
}
// The same, packed at compile time: 4 bits of channel per row lane, one bit
// of ownership per row lane.
template <int N>
struct RsLanes {
  static constexpr uint64_t chan = [] {
    uint64_t t = 0;
    for (int r = 0; r < 16; ++r) t |= (uint64_t)rs_channel(N, r) << (4 * r);
    return t;
  }();
  static constexpr uint32_t owner = [] {
    uint32_t m = 0;
    for (int r = 0; r < 16; ++r) m |= rs_owner(N, r) ? 1u << r : 0u;
    return m;
  }();
};

// Reduce-scatter of N <= 8 per-lane values over the whole wavefront, with no
// LDS exchange.  Three scatter levels over lane bits 5, 4 and 0 halve the
// values a lane holds (of each pair the lanes with the bit clear keep the
// first value's sum, the others the second's; an odd last value is summed in
// both): bit 5 with v_permlane32_swap (which exchanges the upper half of its
// first operand with the lower half of its second, so one swap and one add
// serve a pair), bit 4 with v_permlane16_swap (odd rows of the first operand
// with even rows of the second), bit 0 with a quad_perm exchange and two
// selects.  Bits 1, 2 and 3 are then summed in full with DPP adds (quad_perm
// [2,3,0,1], row_ror:4, row_ror:8).  Lane l ends with the wavefront sum of
// channel ws_channel(N, l).
template <int B>
__device__ __forceinline__ float ws_pair(float x, float y, bool bit) {
  if constexpr (B == 5) {
    auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(p[0]) + __uint_as_float(p[1]);
  } else if constexpr (B == 4) {
    auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(p[0]) + __uint_as_float(p[1]);
  } else {
    const float keep = bit ? y : x, send = bit ? x : y;
    return keep + RGP_DPP(send, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  }
}
template <int B>
__device__ __forceinline__ float ws_self(float x) {
  if constexpr (B == 5) {
    auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(p[0]) + __uint_as_float(p[1]);
  } else if constexpr (B == 4) {
    auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    return __uint_as_float(p[0]) + __uint_as_float(p[1]);
  } else {
    return x + RGP_DPP(x, 0xB1, 0xF, 0xF, false);
  }
}
template <int N, int B>
__device__ __forceinline__ void ws_level(const float* in, float* out, bool bit) {
#pragma unroll
  for (int i = 0; i < N / 2; ++i) out[i] = ws_pair<B>(in[2 * i], in[2 * i + 1], bit);
  if constexpr ((N & 1) != 0) out[N / 2] = ws_self<B>(in[N - 1]);
}
template <int N>
__device__ __forceinline__ float wave_reduce_scatter(const float* v, int lane) {
  static_assert(N >= 1 && N <= 8, "three scatter levels");
  constexpr int N1 = (N + 1) / 2, N2 = (N1 + 1) / 2;
  float a[N1], b[N2], c[1];
  ws_level<N, 5>(v, a, lane & 32);
  ws_level<N1, 4>(a, b, lane & 16);
  ws_level<N2, 0>(b, c, lane & 1);
  float x = c[0];
  x = x + RGP_DPP(x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
  x = x + RGP_DPP(x, 0x124, 0xF, 0xF, false);  // row_ror:4
  x = x + RGP_DPP(x, 0x128, 0xF, 0xF, false);  // row_ror:8
  return x;
}
// The channel wave_reduce_scatter<n> leaves in lane l (it depends on lane
// bits 5, 4 and 0, taken in that order), and whether l is the lowest lane
// holding it.
__host__ __device__ constexpr int ws_channel(int n, int l) {
  int map[8] = {0};
  for (int i = 0; i < 8; ++i) map[i] = i;
  const int bits[3] = {5, 4, 0};
  for (int t = 0; t < 3; ++t) {
    const int bit = (l >> bits[t]) & 1;
    const int h = (n + 1) / 2;
    for (int i = 0; i < n / 2; ++i) map[i] = map[2 * i + bit];
    if (n & 1) map[n / 2] = map[n - 1];
    n = h;
  }
  return map[0];
}
__host__ __device__ constexpr bool ws_owner(int n, int l) {
  for (int q = 0; q < l; ++q)
    if (ws_channel(n, q) == ws_channel(n, l)) return false;
  return true;
}
// Packed at compile time: bit b of `owner` for lane b.
template <int N>
struct WsLanes {
  static constexpr uint64_t owner = [] {
    uint64_t m = 0;
    for (int l = 0; l < 64; ++l) m |= ws_owner(N, l) ? 1ull << l : 0ull;
    return m;
  }();
  static constexpr uint32_t chan8 = [] {  // channel of lane (bit0 | bit4 << 1 | bit5 << 2)
    uint32_t t = 0;
    for (int i = 0; i < 8; ++i) t |= (uint32_t)ws_channel(N, (i & 1) | ((i >> 1) & 1) << 4 | ((i >> 2) & 1) << 5) << (4 * i);
    return t;
  }();
};
__device__ __forceinline__ int ws_lane_index(int lane) { return (lane & 1) | ((lane >> 3) & 6); }

// Sum of lanes l, l^16, l^32 and l^48 into every one of them: the gfx950
// v_permlane16_swap exchanges rows 0<->1 and 2<->3 between its two copies,
// v_permlane32_swap the two halves of the wavefront.
__device__ __forceinline__ float across_rows(float v) {
  auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  v = __uint_as_float(p[0]) + __uint_as_float(p[1]);
  auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(q[0]) + __uint_as_float(q[1]);
}

}  // namespace rgp
