// CPU oracle for the palette-weight Gaussian rasterizer.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU, the arithmetic of
// the reference `gsplatcu` extension (gsplatcu/kernel.cu, gsplatcu/gausplat.cu,
// gsplatcu/common.cuh, gsplatcu/matrix.cuh) so that the HIP kernels can be
// checked against it.  Only tests, the smoke check and the CPU-baseline leg of
// the benchmark may load it; the product path never does.
//
// Every function is a template over the arithmetic type T:
//   T = float  : the reference's FP32 arithmetic, operation by operation, in
//                source order, without implicit contraction (built with
//                -ffp-contract=off).  The compositing recurrence is the one
//                place with fused multiply-adds, written out as std::fma:
//                mahaSqDist's two multiply-adds and the weight update
//                W += (T alpha') w (see maha_sq and draw), the operations nvcc
//                contracts by default and the HIP kernels perform as fmaf.
//                Where the reference multiplies a float by
//                a `double` SH constant (common.cuh:29-44 are double literals),
//                the product is formed in double and rounded to float, exactly
//                as the C++ usual arithmetic conversions prescribe.
//   T = double : the same formulas in FP64, used as the accuracy reference.
// exp(): the reference calls the float overload of exp (kernel.cu:352, 1367).
// The FP32 oracle substitutes rgp_expf below (Cody-Waite reduction + degree-5
// minimax polynomial evaluated with fmaf), which the HIP kernels implement with
// the same operations, so FP32 CPU and GPU agree bit for bit.  Its accuracy
// against libm is pinned by the tests.
//
// Parity status per function is recorded in the header of oracle/oracle.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// constants (common.cuh:13-15, 29-44; kernel.cu:10-12)
// ---------------------------------------------------------------------------
static const int BLOCK = 16;
static const int BLOCK_SIZE = BLOCK * BLOCK;
static const double MIN_DEPTH = 0.2;  // a double literal: comparisons promote
static const float BAD_MARKER = -1.f;
static const int MAX_SPLAT_DIM = 8;

static const double SH_C0_0 = 0.28209479177387814;
static const double SH_C1_0 = -0.4886025119029199;
static const double SH_C1_1 = 0.4886025119029199;
static const double SH_C1_2 = -0.4886025119029199;
static const double SH_C2_0 = 1.0925484305920792;
static const double SH_C2_1 = -1.0925484305920792;
static const double SH_C2_2 = 0.31539156525252005;
static const double SH_C2_3 = -1.0925484305920792;
static const double SH_C2_4 = 0.5462742152960396;
static const double SH_C3_0 = -0.5900435899266435;
static const double SH_C3_1 = 2.890611442640554;
static const double SH_C3_2 = -0.4570457994644658;
static const double SH_C3_3 = 0.3731763325901154;
static const double SH_C3_4 = -0.4570457994644658;
static const double SH_C3_5 = 1.445305721320277;
static const double SH_C3_6 = -0.5900435899266435;

// (T)(c * x) with the product formed in double: `SH_Cx * y` in the reference.
template <typename T> static inline T dmul(double c, T x) { return (T)(c * (double)x); }
// (T)((c * x) * y) in double: `SH_Cx * y * (expr)` in the reference.
template <typename T> static inline T dmul2(double c, T x, T y) {
  return (T)((c * (double)x) * (double)y);
}

// ---------------------------------------------------------------------------
// exp
// ---------------------------------------------------------------------------
// Cody & Waite (1980) argument reduction x = n ln2 + r, |r| <= ln2/2, with ln2
// split into a 16-bit head and a tail, then e^r ~ 1 + r + c2 r^2 + ... + c5 r^5
// with c2..c5 the minimax choice for relative error on |r| <= 0.3466 under
// the constraint that the first two coefficients are 1 (found by linear
// programming on a dense grid; 9.1e-8 relative), Horner form with fmaf, and
// an exact scaling by 2^n assembled from the exponent bits.  Within 2 ulp of
// the correctly rounded exp on every float with a normal result (the bound
// CUDA documents for the expf the reference calls).
static inline float bits2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

const float kC2 = 0x1.fffddcp-2f, kC3 = 0x1.5557acp-3f, kC4 = 0x1.57347cp-5f, kC5 = 0x1.106baap-7f;
const float kRound = 0x1.8p23f;  // 1.5 * 2^23

float rgp_expf(float x) {
  if (!(x > -104.0f)) return (x != x) ? x : 0.0f;
  if (x > 88.72283f) return bits2f(0x7f800000u);
  const float LOG2E = 1.44269504088896341f;
  const float LN2_HI = 0.693145751953125f;        // 0x3f317200
  const float LN2_LO = 1.42860676533018570e-06f;  // ln2 - LN2_HI
  // n: the exact product x log2e rounded once to the nearest integer, read off
  // kRound + n (floats in [2^23, 2^24) are the integers)
  const float n = fmaf(x, LOG2E, kRound) - kRound;
  float r = fmaf(n, -LN2_HI, x);
  r = fmaf(n, -LN2_LO, r);
  float p = kC5;
  p = fmaf(p, r, kC4);
  p = fmaf(p, r, kC3);
  p = fmaf(p, r, kC2);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  int ni = (int)n;
  if (ni < -126) { p = p * 5.42101086e-20f; ni += 64; }  // 2^-64
  if (ni > 127) { p = p * 2.0f; ni -= 1; }
  return p * bits2f((uint32_t)(ni + 127) << 23);
}

// exp(-m/2) as the compositing kernels form it from the Mahalanobis distance
// m >= 0 without a range test: m capped at 1000 (exp(-m/2) is 0 from m = 208
// on), n = m (-log2e/2) rounded once to the nearest integer, the reduction
// against 2 ln2 (head and tail), the Horner partials of rgp_expf with the
// coefficients scaled by (-1/2)^k, and ldexp by n.  This is rgp_expf(-0.5 m)
// with every intermediate scaled by a power of two; orc_gauss_half_check
// compares the two on every float of an interval (test infrastructure: it
// pins the kernels' use of this form for every distance below the largest
// cutoff a finite opacity can have).
float gauss_half_form(float m) {
  const float mm = fminf(m, 1000.0f);
  const float n = fmaf(mm, -0.5f * 1.44269504088896341f, kRound) - kRound;
  float r = fmaf(n, 2.0f * 0.693145751953125f, mm);
  r = fmaf(n, 2.0f * 1.42860676533018570e-06f, r);
  float t = kC5 * -0.03125f;
  t = fmaf(t, r, kC4 * 0.0625f);
  t = fmaf(t, r, kC3 * -0.125f);
  t = fmaf(t, r, kC2 * 0.25f);
  t = fmaf(t, r, -0.5f);
  t = fmaf(t, r, 1.0f);
  return std::ldexp(t, (int)n);
}

// The kernels' form of gauss_half_form for entries whose cutoff is at most 170:
// the same n and polynomial t, scaled by adding the low bits of
// fl(m (-log2e/2) + kRound), shifted left by 23, to t's bit pattern (an
// exponent-field add, exact while 2^n t is normal).  orc_gauss_half_fast_check
// compares it with gauss_half_form on every float of an interval.
float gauss_half_fast_form(float m) {
  const float s = fmaf(m, -0.5f * 1.44269504088896341f, kRound);
  const float n = s - kRound;
  float r = fmaf(n, 2.0f * 0.693145751953125f, m);
  r = fmaf(n, 2.0f * 1.42860676533018570e-06f, r);
  float t = kC5 * -0.03125f;
  t = fmaf(t, r, kC4 * 0.0625f);
  t = fmaf(t, r, kC3 * -0.125f);
  t = fmaf(t, r, kC2 * 0.25f);
  t = fmaf(t, r, -0.5f);
  t = fmaf(t, r, 1.0f);
  uint32_t tb, sb;
  std::memcpy(&tb, &t, 4);
  std::memcpy(&sb, &s, 4);
  return bits2f(tb + (sb << 23));
}

template <typename T> static inline T ex(T x);
template <> inline float ex<float>(float x) { return rgp_expf(x); }
template <> inline double ex<double>(double x) { return std::exp(x); }

// Float -> integer conversions that saturate and map NaN to 0, the GPU
// convention the reference relies on (e.g. kernel.cu:119, 147-150, 400).
template <typename T> static inline int32_t sat_i32(T x) {
  if (x != x) return 0;
  if (x >= (T)2147483648.0) return INT32_MAX;
  if (x <= (T)-2147483648.0) return INT32_MIN;
  return (int32_t)x;
}
template <typename T> static inline uint32_t sat_u32(T x) {
  if (x != x) return 0;
  if (x >= (T)4294967296.0) return UINT32_MAX;
  if (x <= (T)0) return 0;
  return (uint32_t)x;
}

// fminf / fmaxf semantics (a NaN operand yields the other operand).
template <typename T> static inline T fmn(T a, T b) { return std::fmin(a, b); }
template <typename T> static inline T fmx(T a, T b) { return std::fmax(a, b); }

// ---------------------------------------------------------------------------
// A1 project (kernel.cu:618-663; matrix.cuh:53-70 accumulation order)
// ---------------------------------------------------------------------------
template <typename T>
void project(int n, const T* pws, const T* R, const T* t, T fx, T fy, T cx, T cy,
             T* us, T* pcs, T* depths) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    T pc[3];
    for (int r = 0; r < 3; ++r) {
      T s = 0;
      for (int k = 0; k < 3; ++k) s += R[r * 3 + k] * pws[i * 3 + k];
      pc[r] = s + t[r];
    }
    const T x = pc[0], y = pc[1], z = pc[2];
    if ((double)z < MIN_DEPTH) { depths[i] = (T)BAD_MARKER; continue; }
    const T z_inv = (T)1 / z;
    const T x_fx = x * fx;
    const T y_fy = y * fy;
    us[i * 2 + 0] = x_fx * z_inv + cx;
    us[i * 2 + 1] = y_fy * z_inv + cy;
    pcs[i * 3 + 0] = x; pcs[i * 3 + 1] = y; pcs[i * 3 + 2] = z;
    depths[i] = z;
  }
}

// ---------------------------------------------------------------------------
// A2 computeCov3D (kernel.cu:422-462)
// ---------------------------------------------------------------------------
template <typename T>
void cov3d(int n, const T* rots, const T* scales, const T* depths, T* cov3ds) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    if ((double)depths[i] < MIN_DEPTH) continue;
    const T w = rots[i * 4 + 0], x = rots[i * 4 + 1], y = rots[i * 4 + 2], z = rots[i * 4 + 3];
    const T s0 = scales[i * 3 + 0], s1 = scales[i * 3 + 1], s2 = scales[i * 3 + 2];
    const T xx = x * x, yy = y * y, zz = z * z;
    const T xy = x * y, xz = x * z, yz = y * z;
    const T xw = x * w, yw = y * w, zw = z * w;
    const T one = 1, two = 2;
    T R[9] = {one - two * (yy + zz), two * (xy - zw), two * (xz + yw),
              two * (xy + zw), one - two * (xx + zz), two * (yz - xw),
              two * (xz - yw), two * (yz + xw), one - two * (xx + yy)};
    T M[9];
    const T s[3] = {s0, s1, s2};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) M[r * 3 + c] = R[r * 3 + c] * s[c];
    T S[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        T acc = 0;
        for (int k = 0; k < 3; ++k) acc += M[r * 3 + k] * M[c * 3 + k];
        S[r * 3 + c] = acc;
      }
    T* o = cov3ds + i * 6;
    o[0] = S[0]; o[1] = S[1]; o[2] = S[2]; o[3] = S[4]; o[4] = S[5]; o[5] = S[8];
  }
}

// ---------------------------------------------------------------------------
// A3 computeCov2D (kernel.cu:508-555; tan_fov from gausplat.cu:357-358)
// ---------------------------------------------------------------------------
template <typename T>
void cov2d(int n, const T* cov3ds, const T* pcs, const T* Rcw, const T* depths,
           T fx, T fy, T tan_fovx, T tan_fovy, T* cov2ds) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    if ((double)depths[i] < MIN_DEPTH) continue;
    T x = pcs[i * 3 + 0], y = pcs[i * 3 + 1];
    const T z = pcs[i * 3 + 2];
    const T* c3 = cov3ds + i * 6;
    const T a = c3[0], b = c3[1], c = c3[2], d = c3[3], e = c3[4], f = c3[5];
    const T limx = (T)1.3f * tan_fovx;
    const T limy = (T)1.3f * tan_fovy;
    x = fmn(limx, fmx(-limx, x / z)) * z;
    y = fmn(limy, fmx(-limy, y / z)) * z;
    const T z2 = z * z;
    const T J[6] = {fx / z, (T)0, -(fx * x) / z2, (T)0, fy / z, -(fy * y) / z2};
    const T S[9] = {a, b, c, b, d, e, c, e, f};
    T M[6];
    for (int r = 0; r < 2; ++r)
      for (int cc = 0; cc < 3; ++cc) {
        T acc = 0;
        for (int k = 0; k < 3; ++k) acc += J[r * 3 + k] * Rcw[k * 3 + cc];
        M[r * 3 + cc] = acc;
      }
    T MS[6];
    for (int r = 0; r < 2; ++r)
      for (int cc = 0; cc < 3; ++cc) {
        T acc = 0;
        for (int k = 0; k < 3; ++k) acc += M[r * 3 + k] * S[k * 3 + cc];
        MS[r * 3 + cc] = acc;
      }
    T SP[4];
    for (int r = 0; r < 2; ++r)
      for (int cc = 0; cc < 2; ++cc) {
        T acc = 0;
        for (int k = 0; k < 3; ++k) acc += MS[r * 3 + k] * M[cc * 3 + k];
        SP[r * 2 + cc] = acc;
      }
    cov2ds[i * 3 + 0] = SP[0] + (T)0.3f;
    cov2ds[i * 3 + 1] = SP[1];
    cov2ds[i * 3 + 2] = SP[3] + (T)0.3f;
  }
}

// ---------------------------------------------------------------------------
// A4 inverseCov2D (kernel.cu:379-400)
// ---------------------------------------------------------------------------
template <typename T>
void inverse_cov2d(int n, const T* cov2ds, T* depths, T* cinv2ds, int32_t* areas) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    if ((double)depths[i] < MIN_DEPTH) continue;
    const T a = cov2ds[i * 3 + 0], b = cov2ds[i * 3 + 1], c = cov2ds[i * 3 + 2];
    const T det_inv = (T)1 / (a * c - b * b);
    if (std::isnan(det_inv)) { depths[i] = (T)BAD_MARKER; continue; }
    cinv2ds[i * 3 + 0] = det_inv * c;
    cinv2ds[i * 3 + 1] = -det_inv * b;
    cinv2ds[i * 3 + 2] = det_inv * a;
    areas[i * 2 + 0] = sat_i32(std::ceil((T)3 * std::sqrt(std::fabs(a))));
    areas[i * 2 + 1] = sat_i32(std::ceil((T)3 * std::sqrt(std::fabs(c))));
  }
}

// ---------------------------------------------------------------------------
// SH basis of the view direction (kernel.cu:716-760 / 930-957)
// Y[0] = SH_C0_0 is not produced here; Y[1..15] are the 15 higher terms.
// ---------------------------------------------------------------------------
template <typename T>
static void sh_dir_basis(const T* pw, const T* twc, int sh_dim, T* Y /*[16]*/) {
  const T d0 = pw[0] - twc[0], d1 = pw[1] - twc[1], d2 = pw[2] - twc[2];
  const T normd_inv = (T)1 / std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  const T x = d0 * normd_inv, y = d1 * normd_inv, z = d2 * normd_inv;
  Y[1] = dmul(SH_C1_0, y);
  Y[2] = dmul(SH_C1_1, z);
  Y[3] = dmul(SH_C1_2, x);
  if (sh_dim > 4) {
    const T xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    Y[4] = dmul(SH_C2_0, xy);
    Y[5] = dmul(SH_C2_1, yz);
    Y[6] = dmul(SH_C2_2, (T)2 * zz - xx - yy);
    Y[7] = dmul(SH_C2_3, xz);
    Y[8] = dmul(SH_C2_4, xx - yy);
    if (sh_dim > 9) {
      Y[9] = dmul2(SH_C3_0, y, (T)3 * xx - yy);
      Y[10] = dmul2(SH_C3_1, xy, z);
      Y[11] = dmul2(SH_C3_2, y, (T)4 * zz - xx - yy);
      Y[12] = dmul2(SH_C3_3, z, (T)2 * zz - (T)3 * xx - (T)3 * yy);
      Y[13] = dmul2(SH_C3_4, x, (T)4 * zz - xx - yy);
      Y[14] = dmul2(SH_C3_5, z, xx - yy);
      Y[15] = dmul2(SH_C3_6, x, xx - (T)3 * yy);
    }
  }
}

// ---------------------------------------------------------------------------
// A5 sh2Weight, full rank (kernel.cu:670-790).  basis (N,16): dweight_dshs.
// ---------------------------------------------------------------------------
template <typename T>
void sh2weight(int n, int K, int sh_dim, const T* shs, const T* pws, const T* twc,
               T* weights, T* basis /*nullable, (N, sh_dim)*/) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const T* sh = shs + (size_t)i * sh_dim * K;
    T w[MAX_SPLAT_DIM];
    for (int k = 0; k < K; ++k) w[k] = (T)((double)(T)0 + SH_C0_0 * (double)sh[k]);
    T Y[16] = {0};
    if (sh_dim > 1) {
      sh_dir_basis(pws + i * 3, twc, sh_dim, Y);
      for (int k = 0; k < K; ++k)
        w[k] += Y[1] * sh[1 * K + k] + Y[2] * sh[2 * K + k] + Y[3] * sh[3 * K + k];
      if (sh_dim > 4) {
        for (int k = 0; k < K; ++k)
          w[k] += Y[4] * sh[4 * K + k] + Y[5] * sh[5 * K + k] + Y[6] * sh[6 * K + k] +
                  Y[7] * sh[7 * K + k] + Y[8] * sh[8 * K + k];
        if (sh_dim > 9) {
          for (int k = 0; k < K; ++k)
            w[k] += Y[9] * sh[9 * K + k] + Y[10] * sh[10 * K + k] + Y[11] * sh[11 * K + k] +
                    Y[12] * sh[12 * K + k] + Y[13] * sh[13 * K + k] + Y[14] * sh[14 * K + k] +
                    Y[15] * sh[15 * K + k];
        }
      }
    }
    for (int k = 0; k < K; ++k) weights[(size_t)i * K + k] = w[k];
    if (basis) {
      T* b = basis + (size_t)i * sh_dim;
      b[0] = (T)SH_C0_0;
      for (int l = 1; l < sh_dim; ++l) b[l] = Y[l];
    }
  }
}

// ---------------------------------------------------------------------------
// A6 sh2Weight_fast generalised to rank r (kernel.cu:884-987).
// A: (N, P, r) row-major, B: (N, r, Q).  C[row][col] = sum_r A[row][r] B[r][col],
// accumulated over r in ascending order; for r = 1 this is A[row]*B[col], the
// reference expression, bit for bit.
// ---------------------------------------------------------------------------
template <typename T>
void sh2weight_fast(int n, int K, int P, int Q, int rank, const T* dc, const T* A,
                    const T* B, const T* pws, const T* twc, T* weights, T* ylm /*nullable (N,L)*/) {
  const int L = P * Q / K;
  const int sh_dim = L + 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    T w[MAX_SPLAT_DIM];
    for (int k = 0; k < K; ++k) w[k] = dmul(SH_C0_0, dc[(size_t)i * K + k]);
    if (sh_dim <= 1) {
      for (int k = 0; k < K; ++k) weights[(size_t)i * K + k] = w[k];
      continue;
    }
    T Y[16] = {0};
    sh_dir_basis(pws + i * 3, twc, sh_dim, Y);
    if (ylm)
      for (int l = 0; l < L; ++l) ylm[(size_t)i * L + l] = Y[l + 1];
    const T* a = A + (size_t)i * P * rank;
    const T* b = B + (size_t)i * rank * Q;
    for (int k = 0; k < K; ++k) {
      T acc = 0;
      for (int lm = 0; lm < L; ++lm) {
        const int g = lm * K + k, row = g / Q, col = g % Q;
        T c = a[row * rank + 0] * b[0 * Q + col];
        for (int r = 1; r < rank; ++r) c += a[row * rank + r] * b[r * Q + col];
        acc += c * Y[lm + 1];
      }
      w[k] += acc;
    }
    for (int k = 0; k < K; ++k) weights[(size_t)i * K + k] = w[k];
  }
}

// ---------------------------------------------------------------------------
// A16 sh2Weight_fast_B generalised to rank r (kernel.cu:994-1032).  The
// reference's atomics hit thread-private rows, so its sums run in loop order
// (k outer, lm inner), which is the order used here.
// ---------------------------------------------------------------------------
template <typename T>
void sh2weight_fast_backward(int n, int K, int P, int Q, int rank, const T* A, const T* B,
                             const T* ylm, const T* dw, T* dA, T* dB) {
  const int L = P * Q / K;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const T* a = A + (size_t)i * P * rank;
    const T* b = B + (size_t)i * rank * Q;
    const T* y = ylm + (size_t)i * L;
    T* da = dA + (size_t)i * P * rank;
    T* db = dB + (size_t)i * rank * Q;
    for (int j = 0; j < P * rank; ++j) da[j] = 0;
    for (int j = 0; j < rank * Q; ++j) db[j] = 0;
    for (int k = 0; k < K; ++k) {
      const T dwk = dw[(size_t)i * K + k];
      for (int lm = 0; lm < L; ++lm) {
        const int g = lm * K + k, row = g / Q, col = g % Q;
        for (int r = 0; r < rank; ++r) {
          da[row * rank + r] += dwk * b[r * Q + col] * y[lm];
          db[r * Q + col] += dwk * a[row * rank + r] * y[lm];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// sh2Color (kernel.cu:1040-1110).  `SH_C0_0 * sh0` binds the double constant
// to the `const float&` of operator*(float, float3), so it is rounded to float
// first; the other band terms use the double-formed basis values.
// ---------------------------------------------------------------------------
template <typename T>
void sh2color(int n, int sh_dim, const T* shs, const T* pws, const T* twc, T* colors) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const T* sh = shs + (size_t)i * sh_dim * 3;
    const T c0 = (T)SH_C0_0;
    T col[3];
    for (int c = 0; c < 3; ++c) col[c] = c0 * sh[c];
    if (sh_dim > 1) {
      T Y[16] = {0};
      sh_dir_basis(pws + i * 3, twc, sh_dim, Y);
      for (int c = 0; c < 3; ++c)
        col[c] += Y[1] * sh[3 + c] + Y[2] * sh[6 + c] + Y[3] * sh[9 + c];
      if (sh_dim > 4) {
        for (int c = 0; c < 3; ++c)
          col[c] += Y[4] * sh[12 + c] + Y[5] * sh[15 + c] + Y[6] * sh[18 + c] +
                    Y[7] * sh[21 + c] + Y[8] * sh[24 + c];
        if (sh_dim > 9) {
          for (int c = 0; c < 3; ++c)
            col[c] += Y[9] * sh[27 + c] + Y[10] * sh[30 + c] + Y[11] * sh[33 + c] +
                      Y[12] * sh[36 + c] + Y[13] * sh[39 + c] + Y[14] * sh[42 + c] +
                      Y[15] * sh[45 + c];
        }
      }
    }
    for (int c = 0; c < 3; ++c) colors[(size_t)i * 3 + c] = col[c];
  }
}

// sh2Color's calc_J outputs (kernel.cu:1112-1167): dcolor_dshs (N, sh_dim) =
// (SH_C0_0, Y_1, ...), and dcolor_dpws (N,3,3) = dc_dr dr_dpw accumulated from
// 0.  dc_dr is common.cuh's float3 arithmetic per colour channel: a double
// constant times a float is a double product rounded to float where it meets
// the float3, then float products and sums left to right (F below).  The Y_9
// term of dc_dr1 is C30 (3x^2 - 2y^2) (kernel.cu:1152), as in sh2Weight.
template <typename T>
void sh2color_jac(int n, int sh_dim, const T* shs, const T* pws, const T* twc, T* dshs, T* dpws) {
  auto F = [](double v) { return (T)v; };
  for (int i = 0; i < n; ++i) {
    const T* sh = shs + (size_t)i * sh_dim * 3;
    T Y[16] = {0};
    if (sh_dim > 1) sh_dir_basis(pws + i * 3, twc, sh_dim, Y);
    dshs[(size_t)i * sh_dim] = (T)SH_C0_0;
    for (int l = 1; l < sh_dim; ++l) dshs[(size_t)i * sh_dim + l] = Y[l];
    T out[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (sh_dim > 1) {
      const T d0 = pws[3 * i] - twc[0], d1 = pws[3 * i + 1] - twc[1], d2 = pws[3 * i + 2] - twc[2];
      const T ni = (T)1 / std::sqrt((d0 * d0 + d1 * d1) + d2 * d2);
      const T x = d0 * ni, y = d1 * ni, z = d2 * ni;
      const T xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
      const T n3 = (ni * ni) * ni;
      const T r00 = ((-d0) * d0) * n3 + ni, r11 = ((-d1) * d1) * n3 + ni, r22 = ((-d2) * d2) * n3 + ni;
      const T r01 = ((-d0) * d1) * n3, r02 = ((-d0) * d2) * n3, r12 = ((-d1) * d2) * n3;
      const T R[9] = {r00, r01, r02, r01, r11, r12, r02, r12, r22};
      for (int c = 0; c < 3; ++c) {
        auto h = [&](int l) { return sh[3 * l + c]; };
        T D[3] = {0, 0, 0};
        D[0] = D[0] + h(3) * F(SH_C1_2);
        D[1] = D[1] + h(1) * F(SH_C1_0);
        D[2] = D[2] + h(2) * F(SH_C1_1);
        if (sh_dim > 4) {
          const double X = (double)x, Yd = (double)y, Z = (double)z;
          D[0] = D[0] + (((h(4) * F(SH_C2_0 * Yd) - h(6) * F((SH_C2_2 * 2) * X)) + h(7) * F(SH_C2_3 * Z)) +
                         h(8) * F((SH_C2_4 * 2) * X));
          D[1] = D[1] + (((h(4) * F(SH_C2_0 * X) + h(5) * F(SH_C2_1 * Z)) - h(6) * F((SH_C2_2 * 2.0) * Yd)) -
                         h(8) * F((SH_C2_4 * 2) * Yd));
          D[2] = D[2] + ((h(5) * F(SH_C2_1 * Yd) + h(6) * F((SH_C2_2 * 4.0) * Z)) + h(7) * F(SH_C2_3 * X));
          if (sh_dim > 9) {
            const T a[7] = {((h(9) * F(6.0 * SH_C3_0)) * x) * y, (h(10) * F(SH_C3_1)) * yz,
                            (h(11) * F(2 * SH_C3_2)) * xy,         (h(12) * F(6.0 * SH_C3_3)) * xz,
                            (h(13) * F(SH_C3_4)) * (((T)4 * zz - (T)3 * xx) - yy),
                            (h(14) * F(2 * SH_C3_5)) * xz,         (h(15) * F(SH_C3_6)) * ((T)3 * xx - (T)3 * yy)};
            D[0] = D[0] + ((((((a[0] + a[1]) - a[2]) - a[3]) + a[4]) + a[5]) + a[6]);
            const T b[7] = {(h(9) * F(SH_C3_0)) * ((T)3 * xx - (T)2 * yy), (h(10) * F(SH_C3_1)) * xz,
                            (h(11) * F(SH_C3_2)) * (((T)4 * zz - xx) - (T)3 * yy),
                            (h(12) * F(6.0 * SH_C3_3)) * yz,     (h(13) * F(SH_C3_4)) * ((T)-2 * xy),
                            (h(14) * F(2 * SH_C3_5)) * yz,       (h(15) * F(6.0 * SH_C3_6)) * xy};
            D[1] = D[1] + ((((((b[0] + b[1]) + b[2]) - b[3]) + b[4]) - b[5]) - b[6]);
            const T cc[5] = {(h(10) * F(SH_C3_1)) * xy, (h(11) * F(8.0 * SH_C3_2)) * yz,
                             (h(12) * F(SH_C3_3)) * (((T)6 * zz - (T)3 * xx) - (T)3 * yy),
                             (h(13) * F(8.0 * SH_C3_4)) * xz, (h(14) * F(SH_C3_5)) * (xx - yy)};
            D[2] = D[2] + ((((cc[0] + cc[1]) + cc[2]) + cc[3]) + cc[4]);
          }
        }
        for (int j = 0; j < 3; ++j) {
          T acc = 0;
          for (int k = 0; k < 3; ++k) acc = acc + D[k] * R[3 * k + j];
          out[3 * c + j] = acc;
        }
      }
    }
    for (int e = 0; e < 9; ++e) dpws[(size_t)i * 9 + e] = out[e];
  }
}

// ---------------------------------------------------------------------------
// A7 getRects (kernel.cu:127-162): rect = (x0, y0, x1, y1) and tile count,
// culling empty rects in place.  DIV_ROUND_UP on a float argument expands to
// ((v + 16) - 1) / 16 in float.
// ---------------------------------------------------------------------------
template <typename T>
void get_rects(int n, const T* us, int32_t* areas, T* depths, int gx, int gy,
               uint32_t* rects /*(N,4)*/, uint32_t* counts) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    counts[i] = 0;
    if ((double)depths[i] < MIN_DEPTH) continue;
    const T ux = us[2 * i], uy = us[2 * i + 1];
    const T xs = (T)areas[2 * i], ys = (T)areas[2 * i + 1];
    const T B = (T)BLOCK;
    auto clampu = [](int32_t v, int g) -> uint32_t {
      int32_t m = std::max((int32_t)0, v);
      return std::min((uint32_t)g, (uint32_t)m);
    };
    uint32_t r0 = clampu(sat_i32((ux - xs) / B), gx);
    uint32_t r1 = clampu(sat_i32((uy - ys) / B), gy);
    uint32_t r2 = clampu(sat_i32((((ux + xs) + B) - (T)1) / B), gx);
    uint32_t r3 = clampu(sat_i32((((uy + ys) + B) - (T)1) / B), gy);
    uint32_t cnt = (r3 - r1) * (r2 - r0);
    if (cnt == 0) {
      depths[i] = (T)BAD_MARKER;
      areas[2 * i] = 0; areas[2 * i + 1] = 0;
      continue;
    }
    rects[4 * i + 0] = r0; rects[4 * i + 1] = r1; rects[4 * i + 2] = r2; rects[4 * i + 3] = r3;
    counts[i] = cnt;
  }
}

// ---------------------------------------------------------------------------
// A8-A11: inclusive scan, createKeys (kernel.cu:96-124), stable sort of the
// u64 keys (gausplat.cu:216, a stable radix sort), getRanges (kernel.cu:165-187)
// including its M == 1 behaviour.  Returns M.
// ---------------------------------------------------------------------------
// The stable order is the order of the pairs' positions o (Gaussian order), so
// the sort orders (key, o), unique, with a parallel comparison sort; secs
// (nullable) receives the seconds of scan, key duplication, sort and ranges.
struct KeyPos {
  uint64_t key;
  int64_t o;
  bool operator<(const KeyPos& b) const { return key < b.key || (key == b.key && o < b.o); }
};

template <typename T>
int64_t bin_sort(int n, const T* depths, const uint32_t* rects, const uint32_t* counts,
                 int gx, int gy, uint64_t* keys_out /*cap*/, int32_t* ids_out /*cap*/,
                 int64_t cap, int32_t* ranges /*(T,2)*/, double* secs) {
  double t0 = omp_get_wtime();
  std::vector<uint64_t> off(n > 0 ? n : 1);
  uint64_t run = 0;
  for (int i = 0; i < n; ++i) { run += counts[i]; off[i] = run; }
  const int64_t M = (int64_t)run;
  const int tiles = gx * gy;
  for (int t = 0; t < 2 * tiles; ++t) ranges[t] = 0;
  if (M > cap) return M;
  double t1 = omp_get_wtime();
  std::vector<uint64_t> keys(M);
  std::vector<int32_t> ids(M);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const T depth = depths[i];
    if ((double)depth < MIN_DEPTH) continue;
    uint64_t o = (i == 0) ? 0 : off[i - 1];
    if (counts[i] == 0) continue;  // getRects leaves such Gaussians culled
    const uint32_t* r = rects + 4 * i;
    for (uint32_t y = r[1]; y < r[3]; ++y)
      for (uint32_t x = r[0]; x < r[2]; ++x) {
        uint64_t key = (uint64_t)(y * (uint32_t)gx + x);
        key <<= 32;
        key |= (uint64_t)sat_u32(depth * (T)1000);
        keys[o] = key; ids[o] = i; ++o;
      }
  }
  double t2 = omp_get_wtime();
  std::vector<KeyPos> kp(M);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < M; ++p) kp[p] = KeyPos{keys[p], p};
  __gnu_parallel::sort(kp.begin(), kp.end());
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < M; ++p) { keys_out[p] = kp[p].key; ids_out[p] = ids[kp[p].o]; }
  double t3 = omp_get_wtime();
  for (int64_t p = 0; p < M; ++p) {
    const int64_t prv = p == 0 ? 0 : p - 1;
    const uint32_t cur_t = (uint32_t)(keys_out[p] >> 32);
    const uint32_t prv_t = (uint32_t)(keys_out[prv] >> 32);
    if (p == 0) ranges[2 * cur_t + 0] = 0;
    else if (p == M - 1) ranges[2 * cur_t + 1] = (int32_t)M;
    if (prv_t != cur_t) { ranges[2 * prv_t + 1] = (int32_t)p; ranges[2 * cur_t + 0] = (int32_t)p; }
  }
  if (secs) {
    const double t4 = omp_get_wtime();
    secs[0] = t1 - t0;
    secs[1] = t2 - t1;
    secs[2] = t3 - t2;
    secs[3] = t4 - t3;
  }
  return M;
}

// ---------------------------------------------------------------------------
// mahaSqDist (common.cuh:86-89): max(0, cx dx^2 + cz dy^2 + 2 cy dx dy) with
// the two multiply-adds contracted, as nvcc contracts them by default:
// fma((2 cy) dx, dy, fma(cz dy, dy, (cx dx) dx)).  The HIP kernels form the
// same two fused operations.
// ---------------------------------------------------------------------------
template <typename T>
static inline T maha_sq(T cx, T cy, T cz, T dx, T dy) {
  return fmx((T)0, std::fma((T)2 * cy * dx, dy, std::fma(cz * dy, dy, cx * dx * dx)));
}

// ---------------------------------------------------------------------------
// A12 draw_weight (kernel.cu:283-372) and draw_color (kernel.cu:194-276): one
// pixel at a time over its tile list.  `dim` channels of per-Gaussian feature
// (K weights, or 3 colours).  Outputs are written for every pixel of a tile
// with a non-empty list; the caller initialises the rest (final_tau 0 for the
// weight path, gausplat.cu:178; 1 for the colour path, gausplat.cu:39).
// Also returns per-pixel evaluation and contribution counts (for the flop
// model of the benchmark) and, when strips is not null, for every list
// position the mask of the tile's 16x4 row strips (bit (py mod 16) / 4) in
// which some pixel took a contribution from that entry (zero-filled by the
// caller; the compositing kernels' record of which strips the replay backward
// must evaluate).
// ---------------------------------------------------------------------------
template <typename T>
void draw(int width, int height, int dim, const int32_t* ranges, const int32_t* ids,
          const T* us, const T* cinv, const T* alphas, const T* feat, T* layer,
          int32_t* contrib, T* final_tau, int64_t* n_eval, int64_t* n_contrib, uint8_t* strips) {
  const int gx = (width + BLOCK - 1) / BLOCK, gy = (height + BLOCK - 1) / BLOCK;
  const size_t HW = (size_t)width * height;
  int64_t tot_e = 0, tot_c = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : tot_e, tot_c)
  for (int tile = 0; tile < gx * gy; ++tile) {
    const int tx = tile % gx, ty = tile / gx;
    const int r0 = ranges[2 * tile], r1 = ranges[2 * tile + 1];
    const int gs_num = r1 - r0;
    if (gs_num == 0) continue;
    for (int py = ty * BLOCK; py < std::min(height, ty * BLOCK + BLOCK); ++py)
      for (int px = tx * BLOCK; px < std::min(width, tx * BLOCK + BLOCK); ++px) {
        T acc[MAX_SPLAT_DIM];
        for (int k = 0; k < dim; ++k) acc[k] = 0;
        T tau = 1;
        int cont = 0, cont_tmp = 0;
        for (int j = r0; j < r1; ++j) {
          const int g = ids[j];
          const T dx = us[2 * g] - (T)px, dy = us[2 * g + 1] - (T)py;
          const T cx = cinv[3 * g], cy = cinv[3 * g + 1], cz = cinv[3 * g + 2];
          ++cont_tmp;
          ++tot_e;
          const T maha = maha_sq(cx, cy, cz, dx, dy);
          const T ap = fmn((T)0.99f, alphas[g] * ex((T)-0.5f * maha));
          if (ap < (T)0.002f) continue;
          ++tot_c;
          // W_k += T alpha' w_k, the multiply-add contracted (one rounding)
          const T ta = tau * ap;
          for (int k = 0; k < dim; ++k) acc[k] = std::fma(ta, feat[(size_t)g * dim + k], acc[k]);
          if (strips) strips[j] |= (uint8_t)(1u << ((py - ty * BLOCK) / 4));
          cont = cont_tmp;
          tau *= ((T)1 - ap);
          if (tau < (T)0.0001f) break;
        }
        const size_t pix = (size_t)py * width + px;
        for (int k = 0; k < dim; ++k) layer[k * HW + pix] = acc[k];
        contrib[pix] = cont;
        final_tau[pix] = tau;
      }
  }
  if (n_eval) *n_eval = tot_e;
  if (n_contrib) *n_contrib = tot_c;
}

// ---------------------------------------------------------------------------
// A14 draw_weight_B (kernel.cu:1291-1379).  Per (pixel, contributing entry)
// the term dL/dW_k * alpha' * tau is formed exactly as in the reference, in
// type T.  mode 0: terms are summed in double into grad (N,K) and their
// absolute values into gabs (N,K), which bounds the reordering error of any
// atomic summation (OpenMP over tiles; the double sums' own order changes
// them by ~1e-16 relative).  mode 1: OpenMP over tiles, atomic adds in T.
// ---------------------------------------------------------------------------
// replay: the same terms with tau taken from a front-to-back replay of the
// forward recurrence (tau before the entry, the value the forward multiplied
// the weights by) instead of the reference's division from final_tau; the
// replay stops where the forward stopped, so the contributing set is the same.
template <typename T>
void draw_backward_replay(int width, int height, int K, const int32_t* ranges, const int32_t* ids,
                          const T* us, const T* cinv, const T* alphas, const T* dlayer,
                          double* grad, double* gabs) {
  const int gx = (width + BLOCK - 1) / BLOCK, gy = (height + BLOCK - 1) / BLOCK;
  const size_t HW = (size_t)width * height;
#pragma omp parallel for schedule(dynamic, 4)
  for (int tile = 0; tile < gx * gy; ++tile) {
    const int tx = tile % gx, ty = tile / gx;
    const int r0 = ranges[2 * tile], r1 = ranges[2 * tile + 1];
    if (r1 - r0 == 0) continue;
    for (int py = ty * BLOCK; py < std::min(height, ty * BLOCK + BLOCK); ++py)
      for (int px = tx * BLOCK; px < std::min(width, tx * BLOCK + BLOCK); ++px) {
        const size_t pix = (size_t)py * width + px;
        T tau = 1;
        for (int j = r0; j < r1; ++j) {
          const int g = ids[j];
          const T dx = us[2 * g] - (T)px, dy = us[2 * g + 1] - (T)py;
          const T cx = cinv[3 * g], cy = cinv[3 * g + 1], cz = cinv[3 * g + 2];
          const T maha = maha_sq(cx, cy, cz, dx, dy);
          const T ap = fmn((T)0.99f, alphas[g] * ex((T)-0.5f * maha));
          if (ap < (T)0.002f) continue;
          const T ta = tau * ap;
          for (int k = 0; k < K; ++k) {
            const T term = dlayer[k * HW + pix] * ta;
#pragma omp atomic
            grad[(size_t)g * K + k] += (double)term;
#pragma omp atomic
            gabs[(size_t)g * K + k] += std::fabs((double)term);
          }
          tau *= ((T)1 - ap);
          if (tau < (T)0.0001f) break;
        }
      }
  }
}

template <typename T>
void draw_backward(int width, int height, int K, const int32_t* ranges, const int32_t* ids,
                   const T* us, const T* cinv, const T* alphas, const int32_t* contrib,
                   const T* final_tau, const T* dlayer, int mode, double* grad, double* gabs,
                   T* grad_t) {
  if (mode == 2) {
    draw_backward_replay(width, height, K, ranges, ids, us, cinv, alphas, dlayer, grad, gabs);
    return;
  }
  const int gx = (width + BLOCK - 1) / BLOCK, gy = (height + BLOCK - 1) / BLOCK;
  const size_t HW = (size_t)width * height;
#pragma omp parallel for schedule(dynamic, 4)
  for (int tile = 0; tile < gx * gy; ++tile) {
    const int tx = tile % gx, ty = tile / gx;
    const int r0 = ranges[2 * tile], r1 = ranges[2 * tile + 1];
    const int gs_num = r1 - r0;
    if (gs_num == 0) continue;
    for (int py = ty * BLOCK; py < std::min(height, ty * BLOCK + BLOCK); ++py)
      for (int px = tx * BLOCK; px < std::min(width, tx * BLOCK + BLOCK); ++px) {
        const size_t pix = (size_t)py * width + px;
        T dl[MAX_SPLAT_DIM];
        for (int k = 0; k < K; ++k) dl[k] = dlayer[k * HW + pix];
        T tau = final_tau[pix];
        const int cont = contrib[pix];
        for (int i = 0; i < gs_num; ++i) {
          if (i < gs_num - cont) continue;
          const int g = ids[r1 - i - 1];
          const T dx = us[2 * g] - (T)px, dy = us[2 * g + 1] - (T)py;
          const T cx = cinv[3 * g], cy = cinv[3 * g + 1], cz = cinv[3 * g + 2];
          const T maha = maha_sq(cx, cy, cz, dx, dy);
          const T gg = ex((T)-0.5f * maha);
          const T ap = fmn((T)0.99f, alphas[g] * gg);
          if (ap < (T)0.002f) continue;
          tau /= ((T)1 - ap);
          const T dgdw = ap * tau;
          for (int k = 0; k < K; ++k) {
            const T term = dl[k] * dgdw;
            if (mode == 0) {
#pragma omp atomic
              grad[(size_t)g * K + k] += (double)term;
#pragma omp atomic
              gabs[(size_t)g * K + k] += std::fabs((double)term);
            } else {
#pragma omp atomic
              grad_t[(size_t)g * K + k] += term;
            }
          }
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Colour backward with geometry gradients (vanilla_3dgs/gsplatcu/kernel.cu:
// 811-971, drawB): per pixel, the list in reverse from its contrib, tau
// recovered by division from final_tau, the colour behind accumulated as
// gamma = a' c + (1 - a') gamma.  For each contributing (pixel, entry):
//   dL/da' = dot(dL/dgamma, tau (c - gamma)) + dL/dfinal_tau (-final_tau / (1 - a'))
//   dL/dalpha = dL/da' g            dL/dc = dL/dgamma (a' tau)
//   dL/du = dL/da' ((-cx dx - cy dy) a', (-cy dx - cz dy) a')
//   dL/dcinv = dL/da' (-0.5 a' dx^2, -1 a' dx dy, -0.5 a' dy^2)
// (the derivatives of the 0.99 clamp and of the max(0, .) are not taken,
// as in the reference).  grad/gabs (N, 9) = [du (2), dcinv (3), dalpha,
// dcolor (3)], each term summed in double.  dtau may be null (zero).
// ---------------------------------------------------------------------------
template <typename T>
void draw_color_backward(int width, int height, const int32_t* ranges, const int32_t* ids, const T* us,
                         const T* cinv, const T* alphas, const T* colors, const int32_t* contrib,
                         const T* final_tau, const T* dimg, const T* dtau, double* grad, double* gabs) {
  const int gx = (width + BLOCK - 1) / BLOCK, gy = (height + BLOCK - 1) / BLOCK;
  const size_t HW = (size_t)width * height;
  for (int tile = 0; tile < gx * gy; ++tile) {
    const int tx = tile % gx, ty = tile / gx;
    const int r0 = ranges[2 * tile], r1 = ranges[2 * tile + 1];
    const int gs_num = r1 - r0;
    if (gs_num == 0) continue;
    for (int py = ty * BLOCK; py < std::min(height, ty * BLOCK + BLOCK); ++py)
      for (int px = tx * BLOCK; px < std::min(width, tx * BLOCK + BLOCK); ++px) {
        const size_t pix = (size_t)py * width + px;
        const T dl[3] = {dimg[pix], dimg[HW + pix], dimg[2 * HW + pix]};
        const T dlt = dtau ? dtau[pix] : (T)0;
        const T ft = final_tau[pix];
        T tau = ft;
        T gam[3] = {0, 0, 0};
        const int cont = contrib[pix];
        for (int i = 0; i < gs_num; ++i) {
          if (i < gs_num - cont) continue;
          const int g = ids[r1 - i - 1];
          const T dx = us[2 * g] - (T)px, dy = us[2 * g + 1] - (T)py;
          const T cx = cinv[3 * g], cy = cinv[3 * g + 1], cz = cinv[3 * g + 2];
          const T gg = ex((T)-0.5f * maha_sq(cx, cy, cz, dx, dy));
          const T ap = fmn((T)0.99f, alphas[g] * gg);
          if (ap < (T)0.002f) continue;
          tau = tau / ((T)1 - ap);
          const T* c = colors + 3 * g;
          const T dg[3] = {tau * (c[0] - gam[0]), tau * (c[1] - gam[1]), tau * (c[2] - gam[2])};
          const T dap_c = (dl[0] * dg[0] + dl[1] * dg[1]) + dl[2] * dg[2];
          const T dap_t = dlt * (-ft / ((T)1 - ap));
          const T dap = dap_c + dap_t;
          const T at = ap * tau;
          T term[9];
          term[0] = dap * ((-cx * dx - cy * dy) * ap);
          term[1] = dap * ((-cy * dx - cz * dy) * ap);
          term[2] = dap * (((T)-0.5f * ap) * (dx * dx));
          term[3] = dap * (((T)-1.0f * ap) * (dx * dy));
          term[4] = dap * (((T)-0.5f * ap) * (dy * dy));
          term[5] = dap * gg;
          for (int k = 0; k < 3; ++k) term[6 + k] = dl[k] * at;
          for (int e = 0; e < 9; ++e) {
            grad[(size_t)g * 9 + e] += (double)term[e];
            gabs[(size_t)g * 9 + e] += std::fabs((double)term[e]);
          }
          for (int k = 0; k < 3; ++k) gam[k] = ap * c[k] + ((T)1 - ap) * gam[k];
        }
      }
  }
}

// ---------------------------------------------------------------------------
// A13 palette mix: RGB[c, p] = sum_k P[k, c] W[k, p], k ascending, no
// contraction.  Equal in exact arithmetic to baking the palette into the
// per-Gaussian colours and re-rendering (gsmodel.py:488-509 + 13-26).
// ---------------------------------------------------------------------------
template <typename T>
void mix(int K, int64_t HW, const T* pal /*(K,3)*/, const T* W, T* rgb) {
  for (int c = 0; c < 3; ++c)
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < HW; ++p) {
      T acc = pal[0 * 3 + c] * W[0 * HW + p];
      for (int k = 1; k < K; ++k) acc = acc + pal[k * 3 + c] * W[k * HW + p];
      rgb[c * HW + p] = acc;
    }
}

// ---------------------------------------------------------------------------
// calc_J outputs of A1-A5 (the Jacobians gsplatcu returns with calc_J=True).
// Products of the small matrices accumulate from 0 over k in ascending order
// (Matrix::operator*, matrix.cuh:53-70); the other expressions are evaluated
// left to right as written in the reference, every operation rounded on its
// own.  Entries the reference leaves unwritten stay as the caller filled them
// (zeros, gausplat.cu:312-316, 350-354, 392-393, 442-446).
// ---------------------------------------------------------------------------
template <typename T>
static void matmul(const T* A, const T* B, T* C, int I, int K, int J) {
  for (int i = 0; i < I; ++i)
    for (int j = 0; j < J; ++j) {
      T acc = 0;
      for (int k = 0; k < K; ++k) acc += A[i * K + k] * B[k * J + j];
      C[i * J + j] = acc;
    }
}

// du/dpc (N,2,3) of project (kernel.cu:648-660): entries 0, 2, 4, 5.
template <typename T>
void project_jac(int n, const T* pws, const T* R, const T* t, T fx, T fy, T* du) {
  for (int i = 0; i < n; ++i) {
    T pc[3];
    for (int r = 0; r < 3; ++r) {
      T s = 0;
      for (int k = 0; k < 3; ++k) s += R[r * 3 + k] * pws[i * 3 + k];
      pc[r] = s + t[r];
    }
    if ((double)pc[2] < MIN_DEPTH) continue;
    const T z_inv = (T)1 / pc[2];
    const T z2_inv = z_inv * z_inv;
    T* o = du + (size_t)i * 6;
    o[0] = fx * z_inv;
    o[2] = -(pc[0] * fx) * z2_inv;
    o[4] = fy * z_inv;
    o[5] = -(pc[1] * fy) * z2_inv;
  }
}

// dcov3d/drot (N,6,4) and dcov3d/dscale (N,6,3) of computeCov3D
// (kernel.cu:463-500): dSigma/dM (6x9) times dM/dq (9x4) and dM/ds (9x3).
template <typename T>
void cov3d_jac(int n, const T* rots, const T* scales, const T* depths, T* drot, T* dscale) {
  for (int i = 0; i < n; ++i) {
    if ((double)depths[i] < MIN_DEPTH) continue;
    const T w = rots[i * 4 + 0], x = rots[i * 4 + 1], y = rots[i * 4 + 2], z = rots[i * 4 + 3];
    const T s0 = scales[i * 3 + 0], s1 = scales[i * 3 + 1], s2 = scales[i * 3 + 2];
    const T xx = x * x, yy = y * y, zz = z * z;
    const T xy = x * y, xz = x * z, yz = y * z;
    const T xw = x * w, yw = y * w, zw = z * w;
    const T one = 1, two = 2;
    const T R[9] = {one - two * (yy + zz), two * (xy - zw), two * (xz + yw),
                    two * (xy + zw), one - two * (xx + zz), two * (yz - xw),
                    two * (xz - yw), two * (yz + xw), one - two * (xx + yy)};
    const T s[3] = {s0, s1, s2};
    T M[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) M[r * 3 + c] = R[r * 3 + c] * s[c];
    // d(S00, S01, S02, S11, S12, S22) / d(M row-major)
    T dS[54] = {0};
    for (int c = 0; c < 3; ++c) {
      dS[0 * 9 + c] = two * M[c];                 // S00 = M0.M0
      dS[1 * 9 + c] = M[3 + c];                   // S01 = M0.M1
      dS[1 * 9 + 3 + c] = M[c];
      dS[2 * 9 + c] = M[6 + c];                   // S02 = M0.M2
      dS[2 * 9 + 6 + c] = M[c];
      dS[3 * 9 + 3 + c] = two * M[3 + c];         // S11
      dS[4 * 9 + 3 + c] = M[6 + c];               // S12 = M1.M2
      dS[4 * 9 + 6 + c] = M[3 + c];
      dS[5 * 9 + 6 + c] = two * M[6 + c];         // S22
    }
    // dM[r][c] / d(w, x, y, z): M[r][c] = R[r][c] s_c
    const T m4 = -4, m2 = -2;
    const T dMq[36] = {
        0, 0, m4 * s0 * y, m4 * s0 * z,                              // M00
        m2 * s1 * z, two * s1 * y, two * s1 * x, m2 * s1 * w,        // M01
        two * s2 * y, two * s2 * z, two * s2 * w, two * s2 * x,      // M02
        two * s0 * z, two * s0 * y, two * s0 * x, two * s0 * w,      // M10
        0, m4 * s1 * x, 0, m4 * s1 * z,                              // M11
        m2 * s2 * x, m2 * s2 * w, two * s2 * z, two * s2 * y,        // M12
        m2 * s0 * y, two * s0 * z, m2 * s0 * w, two * s0 * x,        // M20
        two * s1 * x, two * s1 * w, two * s1 * z, two * s1 * y,      // M21
        0, m4 * s2 * x, m4 * s2 * y, 0};                             // M22
    T dMs[27] = {0};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) dMs[(r * 3 + c) * 3 + c] = R[r * 3 + c];
    matmul(dS, dMq, drot + (size_t)i * 24, 6, 9, 4);
    matmul(dS, dMs, dscale + (size_t)i * 18, 6, 9, 3);
  }
}

// dcov2d/dcov3d (N,3,6) and dcov2d/dpc (N,3,3) of computeCov2D
// (kernel.cu:557-607).  The clamped x, y enter J and dM/dpc; the clamp
// itself is not differentiated (as in the reference).
template <typename T>
void cov2d_jac(int n, const T* cov3ds, const T* pcs, const T* Rcw, const T* depths, T fx, T fy,
               T tan_fovx, T tan_fovy, T* dcov3d, T* dpc) {
  for (int i = 0; i < n; ++i) {
    if ((double)depths[i] < MIN_DEPTH) continue;
    T x = pcs[i * 3 + 0], y = pcs[i * 3 + 1];
    const T z = pcs[i * 3 + 2];
    const T* c3 = cov3ds + i * 6;
    const T a = c3[0], b = c3[1], c = c3[2], d = c3[3], e = c3[4], f = c3[5];
    const T limx = (T)1.3f * tan_fovx;
    const T limy = (T)1.3f * tan_fovy;
    x = fmn(limx, fmx(-limx, x / z)) * z;
    y = fmn(limy, fmx(-limy, y / z)) * z;
    const T z2 = z * z;
    const T J[6] = {fx / z, (T)0, -(fx * x) / z2, (T)0, fy / z, -(fy * y) / z2};
    T M[6];
    matmul(J, Rcw, M, 2, 3, 3);
    const T two = 2;
    T* o = dcov3d + (size_t)i * 18;
    // row r of the 2x2 result pairs rows (p, q) of M: (0,0), (0,1), (1,1)
    const int pr[3][2] = {{0, 0}, {0, 1}, {1, 1}};
    for (int r = 0; r < 3; ++r) {
      const T* P = M + 3 * pr[r][0];
      const T* Q = M + 3 * pr[r][1];
      if (pr[r][0] == pr[r][1]) {
        o[r * 6 + 0] = P[0] * P[0];
        o[r * 6 + 1] = two * P[0] * P[1];
        o[r * 6 + 2] = two * P[0] * P[2];
        o[r * 6 + 3] = P[1] * P[1];
        o[r * 6 + 4] = two * P[1] * P[2];
        o[r * 6 + 5] = P[2] * P[2];
      } else {
        o[r * 6 + 0] = P[0] * Q[0];
        o[r * 6 + 1] = P[0] * Q[1] + P[1] * Q[0];
        o[r * 6 + 2] = P[0] * Q[2] + P[2] * Q[0];
        o[r * 6 + 3] = P[1] * Q[1];
        o[r * 6 + 4] = P[1] * Q[2] + P[2] * Q[1];
        o[r * 6 + 5] = P[2] * Q[2];
      }
    }
    // dcov2d / dM (3x6): rows of Sigma times rows of M
    const T S[9] = {a, b, c, b, d, e, c, e, f};
    T dM[18] = {0};
    for (int cc = 0; cc < 3; ++cc) {
      const T* Sc = S + 3 * cc;
      dM[0 * 6 + cc] = two * Sc[0] * M[0] + two * Sc[1] * M[1] + two * Sc[2] * M[2];
      dM[1 * 6 + cc] = Sc[0] * M[3] + Sc[1] * M[4] + Sc[2] * M[5];
      dM[1 * 6 + 3 + cc] = Sc[0] * M[0] + Sc[1] * M[1] + Sc[2] * M[2];
      dM[2 * 6 + 3 + cc] = two * Sc[0] * M[3] + two * Sc[1] * M[4] + two * Sc[2] * M[5];
    }
    const T z2_inv = (T)1 / (z * z);
    const T z3_inv = z2_inv / z;
    T dMp[18] = {0};
    for (int cc = 0; cc < 3; ++cc) {
      const T R0 = Rcw[0 * 3 + cc], R1 = Rcw[1 * 3 + cc], R2 = Rcw[2 * 3 + cc];
      dMp[cc * 3 + 0] = -fx * R2 * z2_inv;
      dMp[cc * 3 + 2] = -fx * R0 * z2_inv + two * fx * R2 * x * z3_inv;
      dMp[(3 + cc) * 3 + 1] = -fy * R2 * z2_inv;
      dMp[(3 + cc) * 3 + 2] = -fy * R1 * z2_inv + two * fy * R2 * y * z3_inv;
    }
    matmul(dM, dMp, dpc + (size_t)i * 9, 3, 6, 3);
  }
}

// dcinv/dcov2d (N,3,3) of inverseCov2D (kernel.cu:402-414).
template <typename T>
void inverse_jac(int n, const T* cov2ds, const T* depths, T* dcinv) {
  for (int i = 0; i < n; ++i) {
    if ((double)depths[i] < MIN_DEPTH) continue;
    const T a = cov2ds[i * 3 + 0], b = cov2ds[i * 3 + 1], c = cov2ds[i * 3 + 2];
    const T di = (T)1 / (a * c - b * b);
    if (std::isnan(di)) continue;
    const T d2 = di * di, two = 2;
    T* o = dcinv + (size_t)i * 9;
    o[0] = -c * c * d2;
    o[1] = two * b * c * d2;
    o[2] = -a * c * d2 + di;
    o[3] = b * c * d2;
    o[4] = -two * b * b * d2 - di;
    o[5] = a * b * d2;
    o[6] = -a * c * d2 + di;
    o[7] = two * a * b * d2;
    o[8] = -a * a * d2;
  }
}

// dW/dpw (N,K,3) of sh2Weight (kernel.cu:769-869): dW/ddir (from the SH
// polynomials, SH constants in double as in the reference, each sum added
// to its float accumulator) times ddir/dpw = (I - dir dir^T) / |d|.
template <typename T>
void sh2weight_jac(int n, int K, int sh_dim, const T* shs, const T* pws, const T* twc, T* dwdpw) {
  typedef double D;
  for (int i = 0; i < n; ++i) {
    const T* sh = shs + (size_t)i * sh_dim * K;
    T* out = dwdpw + (size_t)i * K * 3;
    if (sh_dim <= 1) {
      for (int q = 0; q < K * 3; ++q) out[q] = 0;
      continue;
    }
    const T d0 = pws[i * 3 + 0] - twc[0], d1 = pws[i * 3 + 1] - twc[1], d2 = pws[i * 3 + 2] - twc[2];
    const T ni = (T)1 / std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const T x = d0 * ni, y = d1 * ni, z = d2 * ni;
    const T ni3 = ni * ni * ni;
    const T r00 = -d0 * d0 * ni3 + ni, r11 = -d1 * d1 * ni3 + ni, r22 = -d2 * d2 * ni3 + ni;
    const T r01 = -d0 * d1 * ni3, r02 = -d0 * d2 * ni3, r12 = -d1 * d2 * ni3;
    const T dr[9] = {r00, r01, r02, r01, r11, r12, r02, r12, r22};
    const T xx = x * x, yy = y * y, zz = z * z;
    for (int k = 0; k < K; ++k) {
      auto S = [&](int l) { return (D)sh[l * K + k]; };
      T g0 = (T)((D)(T)0 + SH_C1_2 * S(3));
      T g1 = (T)((D)(T)0 + SH_C1_0 * S(1));
      T g2 = (T)((D)(T)0 + SH_C1_1 * S(2));
      if (sh_dim > 4) {
        g0 = (T)((D)g0 + ((((SH_C2_0 * (D)y) * S(4) - ((SH_C2_2 * 2) * (D)x) * S(6)) +
                          (SH_C2_3 * (D)z) * S(7)) + ((SH_C2_4 * 2) * (D)x) * S(8)));
        g1 = (T)((D)g1 + ((((SH_C2_0 * (D)x) * S(4) + (SH_C2_1 * (D)z) * S(5)) -
                          ((SH_C2_2 * 2) * (D)y) * S(6)) - ((SH_C2_4 * 2) * (D)y) * S(8)));
        g2 = (T)((D)g2 + (((SH_C2_1 * (D)y) * S(5) + ((SH_C2_2 * 4) * (D)z) * S(6)) +
                          (SH_C2_3 * (D)x) * S(7)));
        if (sh_dim > 9) {
          const T p0 = (T)4 * zz - (T)3 * xx - yy, p1 = (T)3 * xx - (T)3 * yy;
          const T q0 = (T)3 * xx - (T)2 * yy, q1 = (T)4 * zz - xx - (T)3 * yy;
          const T u0 = (T)6 * zz - (T)3 * xx - (T)3 * yy, u1 = xx - yy;
          D a0 = (((6.0 * SH_C3_0) * (D)x) * (D)y) * S(9);
          a0 = a0 + ((SH_C3_1 * (D)y) * (D)z) * S(10);
          a0 = a0 - (((2.0 * SH_C3_2) * (D)x) * (D)y) * S(11);
          a0 = a0 - (((6.0 * SH_C3_3) * (D)x) * (D)z) * S(12);
          a0 = a0 + (SH_C3_4 * S(13)) * (D)p0;
          a0 = a0 + (((2.0 * SH_C3_5) * (D)x) * (D)z) * S(14);
          a0 = a0 + (SH_C3_6 * S(15)) * (D)p1;
          D a1 = (SH_C3_0 * S(9)) * (D)q0;
          a1 = a1 + ((SH_C3_1 * (D)x) * (D)z) * S(10);
          a1 = a1 + (SH_C3_2 * S(11)) * (D)q1;
          a1 = a1 - (((6.0 * SH_C3_3) * (D)y) * (D)z) * S(12);
          a1 = a1 - (((2.0 * SH_C3_4) * (D)x) * (D)y) * S(13);
          a1 = a1 - (((2.0 * SH_C3_5) * (D)y) * (D)z) * S(14);
          a1 = a1 - (((6.0 * SH_C3_6) * (D)x) * (D)y) * S(15);
          D a2 = ((SH_C3_1 * (D)x) * (D)y) * S(10);
          a2 = a2 + (((8.0 * SH_C3_2) * (D)y) * (D)z) * S(11);
          a2 = a2 + (SH_C3_3 * S(12)) * (D)u0;
          a2 = a2 + (((8.0 * SH_C3_4) * (D)x) * (D)z) * S(13);
          a2 = a2 + (SH_C3_5 * S(14)) * (D)u1;
          g0 = (T)((D)g0 + a0);
          g1 = (T)((D)g1 + a1);
          g2 = (T)((D)g2 + a2);
        }
      }
      for (int j = 0; j < 3; ++j) out[k * 3 + j] = g0 * dr[0 * 3 + j] + g1 * dr[1 * 3 + j] + g2 * dr[2 * 3 + j];
    }
  }
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C entry points (ctypes); suffix f = FP32, d = FP64.
// ---------------------------------------------------------------------------
#define ORC_BOTH(NAME, T, S)                                                                      \
  void orc_project_##S(int n, const T* pws, const T* R, const T* t, T fx, T fy, T cx, T cy,       \
                       T* us, T* pcs, T* depths) {                                                \
    orc::project<T>(n, pws, R, t, fx, fy, cx, cy, us, pcs, depths);                               \
  }                                                                                               \
  void orc_cov3d_##S(int n, const T* rots, const T* scales, const T* depths, T* cov3ds) {         \
    orc::cov3d<T>(n, rots, scales, depths, cov3ds);                                               \
  }                                                                                               \
  void orc_cov2d_##S(int n, const T* cov3ds, const T* pcs, const T* Rcw, const T* depths, T fx,   \
                     T fy, T tfx, T tfy, T* cov2ds) {                                             \
    orc::cov2d<T>(n, cov3ds, pcs, Rcw, depths, fx, fy, tfx, tfy, cov2ds);                         \
  }                                                                                               \
  void orc_inverse_cov2d_##S(int n, const T* cov2ds, T* depths, T* cinv2ds, int32_t* areas) {     \
    orc::inverse_cov2d<T>(n, cov2ds, depths, cinv2ds, areas);                                     \
  }                                                                                               \
  void orc_sh2weight_##S(int n, int K, int sh_dim, const T* shs, const T* pws, const T* twc,      \
                         T* weights, T* basis) {                                                  \
    orc::sh2weight<T>(n, K, sh_dim, shs, pws, twc, weights, basis);                               \
  }                                                                                               \
  void orc_sh2weight_fast_##S(int n, int K, int P, int Q, int rank, const T* dc, const T* A,      \
                              const T* B, const T* pws, const T* twc, T* weights, T* ylm) {       \
    orc::sh2weight_fast<T>(n, K, P, Q, rank, dc, A, B, pws, twc, weights, ylm);                   \
  }                                                                                               \
  void orc_sh2weight_fast_backward_##S(int n, int K, int P, int Q, int rank, const T* A,          \
                                       const T* B, const T* ylm, const T* dw, T* dA, T* dB) {     \
    orc::sh2weight_fast_backward<T>(n, K, P, Q, rank, A, B, ylm, dw, dA, dB);                     \
  }                                                                                               \
  void orc_sh2color_##S(int n, int sh_dim, const T* shs, const T* pws, const T* twc, T* colors) { \
    orc::sh2color<T>(n, sh_dim, shs, pws, twc, colors);                                           \
  }                                                                                               \
  void orc_sh2color_jac_##S(int n, int sh_dim, const T* shs, const T* pws, const T* twc, T* dshs,  \
                            T* dpws) {                                                            \
    orc::sh2color_jac<T>(n, sh_dim, shs, pws, twc, dshs, dpws);                                   \
  }                                                                                               \
  void orc_get_rects_##S(int n, const T* us, int32_t* areas, T* depths, int gx, int gy,           \
                         uint32_t* rects, uint32_t* counts) {                                     \
    orc::get_rects<T>(n, us, areas, depths, gx, gy, rects, counts);                               \
  }                                                                                               \
  int64_t orc_bin_sort_##S(int n, const T* depths, const uint32_t* rects, const uint32_t* counts, \
                           int gx, int gy, uint64_t* keys, int32_t* ids, int64_t cap,             \
                           int32_t* ranges, double* secs) {                                       \
    return orc::bin_sort<T>(n, depths, rects, counts, gx, gy, keys, ids, cap, ranges, secs);      \
  }                                                                                               \
  void orc_draw_##S(int w, int h, int dim, const int32_t* ranges, const int32_t* ids,             \
                    const T* us, const T* cinv, const T* alphas, const T* feat, T* layer,         \
                    int32_t* contrib, T* final_tau, int64_t* ne, int64_t* nc, uint8_t* strips) {  \
    orc::draw<T>(w, h, dim, ranges, ids, us, cinv, alphas, feat, layer, contrib, final_tau, ne,  \
                 nc, strips);                                                                     \
  }                                                                                               \
  void orc_draw_backward_##S(int w, int h, int K, const int32_t* ranges, const int32_t* ids,      \
                             const T* us, const T* cinv, const T* alphas, const int32_t* contrib, \
                             const T* final_tau, const T* dlayer, int mode, double* grad,         \
                             double* gabs, T* grad_t) {                                           \
    orc::draw_backward<T>(w, h, K, ranges, ids, us, cinv, alphas, contrib, final_tau, dlayer,    \
                          mode, grad, gabs, grad_t);                                              \
  }                                                                                               \
  void orc_draw_color_backward_##S(int w, int h, const int32_t* ranges, const int32_t* ids,         \
                                   const T* us, const T* cinv, const T* alphas, const T* colors,  \
                                   const int32_t* contrib, const T* final_tau, const T* dimg,     \
                                   const T* dtau, double* grad, double* gabs) {                   \
    orc::draw_color_backward<T>(w, h, ranges, ids, us, cinv, alphas, colors, contrib, final_tau,  \
                                dimg, dtau, grad, gabs);                                          \
  }                                                                                               \
  void orc_mix_##S(int K, int64_t HW, const T* pal, const T* W, T* rgb) {                         \
    orc::mix<T>(K, HW, pal, W, rgb);                                                              \
  }                                                                                               \
  void orc_project_jac_##S(int n, const T* pws, const T* R, const T* t, T fx, T fy, T* du) {      \
    orc::project_jac<T>(n, pws, R, t, fx, fy, du);                                                \
  }                                                                                               \
  void orc_cov3d_jac_##S(int n, const T* rots, const T* scales, const T* depths, T* drot,         \
                         T* dscale) {                                                             \
    orc::cov3d_jac<T>(n, rots, scales, depths, drot, dscale);                                     \
  }                                                                                               \
  void orc_cov2d_jac_##S(int n, const T* cov3ds, const T* pcs, const T* Rcw, const T* depths,     \
                         T fx, T fy, T tfx, T tfy, T* dcov3d, T* dpc) {                           \
    orc::cov2d_jac<T>(n, cov3ds, pcs, Rcw, depths, fx, fy, tfx, tfy, dcov3d, dpc);                \
  }                                                                                               \
  void orc_inverse_jac_##S(int n, const T* cov2ds, const T* depths, T* dcinv) {                   \
    orc::inverse_jac<T>(n, cov2ds, depths, dcinv);                                                \
  }                                                                                               \
  void orc_sh2weight_jac_##S(int n, int K, int sh_dim, const T* shs, const T* pws, const T* twc,  \
                             T* dwdpw) {                                                          \
    orc::sh2weight_jac<T>(n, K, sh_dim, shs, pws, twc, dwdpw);                                    \
  }

namespace orc {
// ---------------------------------------------------------------------------
// Stages of the finetune step that only the CPU timing sample needs, in FP32.
//
// L1 + SSIM (gsplat/pytorch_ssim.py:17-66 with zero padding, 11x11 window as
// two separable passes of the float32 taps; (1 - lam) L1 + lam (1 - SSIM),
// vanilla_3dgs/gsplat/pytorch_ssim.py:64-67): the loss and dL/dimg for
// dL/dloss = 1, as
//   dL/dx = (1-lam)/n sign(x-y) - lam/n (w*A + 2x (w*B) + y (w*C)),
//   A = dS/dmu_x, B = dS/dE[x^2], C = dS/dE[xy] per pixel.
// OpenMP over rows.
// ---------------------------------------------------------------------------
static void blur_plane(int H, int W, const float* w11, const float* in, float* tmp, float* out) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      float a = 0.f;
      for (int t = 0; t < 11; ++t) {
        const int xx = x + t - 5;
        if (xx >= 0 && xx < W) a = a + w11[t] * in[(size_t)y * W + xx];
      }
      tmp[(size_t)y * W + x] = a;
    }
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      float a = 0.f;
      for (int t = 0; t < 11; ++t) {
        const int yy = y + t - 5;
        if (yy >= 0 && yy < H) a = a + w11[t] * tmp[(size_t)yy * W + x];
      }
      out[(size_t)y * W + x] = a;
    }
}

double l1_ssim(int C, int H, int W, const float* img, const float* tgt, const float* w11, float C1, float C2,
               float lam, float* dimg) {
  const size_t HW = (size_t)H * W;
  std::vector<float> xx(HW), yy(HW), xy(HW), tmp(HW), m1(HW), m2(HW), e11(HW), e22(HW), e12(HW);
  std::vector<float> A(HW), B(HW), Cc(HW), bA(HW), bB(HW), bC(HW);
  const double n = (double)C * HW;
  double l1 = 0.0, ss = 0.0;
  for (int c = 0; c < C; ++c) {
    const float* x = img + c * HW;
    const float* y = tgt + c * HW;
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < HW; ++p) {
      xx[p] = x[p] * x[p];
      yy[p] = y[p] * y[p];
      xy[p] = x[p] * y[p];
    }
    blur_plane(H, W, w11, x, tmp.data(), m1.data());
    blur_plane(H, W, w11, y, tmp.data(), m2.data());
    blur_plane(H, W, w11, xx.data(), tmp.data(), e11.data());
    blur_plane(H, W, w11, yy.data(), tmp.data(), e22.data());
    blur_plane(H, W, w11, xy.data(), tmp.data(), e12.data());
    double l1c = 0.0, ssc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : l1c, ssc)
    for (size_t p = 0; p < HW; ++p) {
      const float mx = m1[p], my = m2[p];
      const float s1 = e11[p] - mx * mx, s2 = e22[p] - my * my, s12 = e12[p] - mx * my;
      const float N1 = 2.f * mx * my + C1, N2 = 2.f * s12 + C2;
      const float D1 = mx * mx + my * my + C1, D2 = s1 + s2 + C2;
      const float DD = D1 * D2;
      ssc += (double)((N1 * N2) / DD);
      l1c += (double)std::fabs(x[p] - y[p]);
      const float S = (N1 * N2) / DD;
      A[p] = 2.f * my * N2 / DD - 2.f * mx * S / D1 + 2.f * mx * S / D2 - 2.f * my * N1 / DD;
      B[p] = -S / D2;
      Cc[p] = 2.f * N1 / DD;
    }
    l1 += l1c;
    ss += ssc;
    if (dimg) {
      blur_plane(H, W, w11, A.data(), tmp.data(), bA.data());
      blur_plane(H, W, w11, B.data(), tmp.data(), bB.data());
      blur_plane(H, W, w11, Cc.data(), tmp.data(), bC.data());
      const float cl = (float)((1.0 - lam) / n), cs = (float)(lam / n);
      float* d = dimg + c * HW;
#pragma omp parallel for schedule(static)
      for (size_t p = 0; p < HW; ++p) {
        const float sg = x[p] > y[p] ? 1.f : (x[p] < y[p] ? -1.f : 0.f);
        d[p] = cl * sg - cs * ((bA[p] + 2.f * x[p] * bB[p]) + y[p] * bC[p]);
      }
    }
  }
  return (1.0 - lam) * l1 / n + lam * (1.0 - ss / n);
}

// dL/dshs[n, l*K + k] = dL/dw[n, k] * Y_l(n) (gsmodel.py:140-143), in FP32.
void sh_backward_full(int64_t n, int K, int L, const float* dw, const float* basis, float* dshs) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    for (int l = 0; l < L; ++l)
      for (int k = 0; k < K; ++k) dshs[(i * L + l) * K + k] = dw[i * K + k] * basis[i * L + l];
}

// torch.optim.Adam's element update (no weight decay), scalars in double.
void adam(int64_t n, float* p, const float* g, float* m, float* v, int64_t step, double lr, double b1,
          double b2, double eps) {
  const double bc1 = 1.0 - std::pow(b1, (double)step), bc2 = 1.0 - std::pow(b2, (double)step);
  const float a = (float)(-lr / bc1), rs = (float)(1.0 / std::sqrt(bc2)), e = (float)eps;
  const float c1 = (float)(1.0 - b1), c2 = (float)(1.0 - b2), fb2 = (float)b2;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const float gi = g[i];
    m[i] = m[i] + c1 * (gi - m[i]);
    v[i] = v[i] * fb2 + (c2 * gi) * gi;
    p[i] = p[i] + a * (m[i] / (std::sqrt(v[i]) * rs + e));
  }
}
}  // namespace orc

extern "C" {
ORC_BOTH(f, float, f)
ORC_BOTH(d, double, d)
float orc_expf(float x) { return orc::rgp_expf(x); }
void orc_expf_array(int64_t n, const float* x, float* y) {
  for (int64_t i = 0; i < n; ++i) y[i] = orc::rgp_expf(x[i]);
}
// Every float m in [lo, hi] (lo >= 0): counts those where gauss_half_form(m)
// and rgp_expf(-0.5f * m) differ in any bit; *first gets the smallest such m.
void orc_gauss_half_check(float lo, float hi, int64_t* checked, int64_t* differ, float* first) {
  uint32_t a, b;
  std::memcpy(&a, &lo, 4);
  std::memcpy(&b, &hi, 4);
  int64_t nd = 0;
  uint32_t fb = 0xffffffffu;
#pragma omp parallel for reduction(+ : nd) reduction(min : fb) schedule(static)
  for (int64_t u = a; u <= (int64_t)b; ++u) {
    const float m = orc::bits2f((uint32_t)u);
    const float g1 = orc::gauss_half_form(m), g2 = orc::rgp_expf(-0.5f * m);
    uint32_t x1, x2;
    std::memcpy(&x1, &g1, 4);
    std::memcpy(&x2, &g2, 4);
    if (x1 != x2) {
      ++nd;
      if ((uint32_t)u < fb) fb = (uint32_t)u;
    }
  }
  *checked = (int64_t)b - (int64_t)a + 1;
  *differ = nd;
  *first = nd ? orc::bits2f(fb) : 0.f;
}
// The same over every float m in [lo, hi] for gauss_half_fast_form(m) against
// gauss_half_form(m).
void orc_gauss_half_fast_check(float lo, float hi, int64_t* checked, int64_t* differ, float* first) {
  uint32_t a, b;
  std::memcpy(&a, &lo, 4);
  std::memcpy(&b, &hi, 4);
  int64_t nd = 0;
  uint32_t fb = 0xffffffffu;
#pragma omp parallel for reduction(+ : nd) reduction(min : fb) schedule(static)
  for (int64_t u = a; u <= (int64_t)b; ++u) {
    const float m = orc::bits2f((uint32_t)u);
    const float g1 = orc::gauss_half_fast_form(m), g2 = orc::gauss_half_form(m);
    uint32_t x1, x2;
    std::memcpy(&x1, &g1, 4);
    std::memcpy(&x2, &g2, 4);
    if (x1 != x2) {
      ++nd;
      if ((uint32_t)u < fb) fb = (uint32_t)u;
    }
  }
  *checked = (int64_t)b - (int64_t)a + 1;
  *differ = nd;
  *first = nd ? orc::bits2f(fb) : 0.f;
}
double orc_l1_ssim_f(int C, int H, int W, const float* img, const float* tgt, const float* w11, float C1, float C2,
                     float lam, float* dimg) {
  return orc::l1_ssim(C, H, W, img, tgt, w11, C1, C2, lam, dimg);
}
void orc_sh_backward_full_f(int64_t n, int K, int L, const float* dw, const float* basis, float* dshs) {
  orc::sh_backward_full(n, K, L, dw, basis, dshs);
}
void orc_adam_f(int64_t n, float* p, const float* g, float* m, float* v, int64_t step, double lr, double b1,
                double b2, double eps) {
  orc::adam(n, p, g, m, v, step, lr, b1, b2, eps);
}
int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_num_threads(int t) {
#ifdef _OPENMP
  omp_set_num_threads(t);
#else
  (void)t;
#endif
}
}
