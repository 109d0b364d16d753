// Barycentric palette loss of the finetune step for gfx950, forward and
// backward in one pass over the image.
//
// The reference computes it with torch operations on (HW, K_c, K_c+2)
// temporaries: barycentric_coords_3d (gsplat/gsmodel.py:219-285) and the loss
// terms of finetune.py:353-392.  Per pixel, for each of the K_c tetrahedra
// (black, white, P_k, P_{k+1}): coordinates from the facet's adjugate,
// softplus projection, reconstruction error; a softmax over facets (tau) of
// the errors weights the coordinates into W_bary (K_c+2 channels).  The loss
// is L_geo = mean((W_tilde - W_bary)^2) + lam_pos mean(relu(-W_tilde)), with
// the optional clamp/renormalisation of W_bary, and, given a ground-truth
// alpha a: C / max(a, 1e-3) clamped to [0,1], W_bary * a, and lam_alpha
// mean((sum W_tilde - a)^2).
//
// Every loss term is a per-pixel mean, so dL/dW_tilde is known at the pixel
// and dL/dW_bary is pulled back through the barycentric solve right away.
// The palette gradient is the sum over pixels of 16 terms per facet (d adj,
// d dinv, the direct dP_i, dP_j of the reconstruction): each thread sums its
// pixels in registers, the block reduces them with DPP and adds them to
// float64 accumulators; a one-block kernel maps them to dL/dpalette in double.
//
// FP32 arithmetic is that of the CPU oracle (oracle/rgp_oracle_bary.cpp),
// operation by operation: exp_cw, log1p_pos, correctly rounded division.  The
// sums over zero rows of the reference's matmuls and einsums are skipped:
// every partial sum stays a non-negative-zero-free finite value, so adding a
// signed zero product leaves it unchanged.
#include <mutex>

#include "rgp_internal.h"

namespace rgp {

constexpr int kBaryAcc = 16;  // per facet: d adj (9), d dinv, dP_i (3), dP_j (3)
constexpr int kBaryThreads = 256;

struct BaryFacet {
  float a[9], adj[9], dinv;
};

// gsmodel.py:225-252: columns white - black, P_i - black, P_j - black.
__device__ __forceinline__ void bary_facet(const float* __restrict__ P, int i, int j, BaryFacet* f) {
  float* a = f->a;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    a[3 * r + 0] = 1.0f - 0.0f;
    a[3 * r + 1] = P[3 * i + r] - 0.0f;
    a[3 * r + 2] = P[3 * j + r] - 0.0f;
  }
  const float a00 = a[0], a01 = a[1], a02 = a[2], a10 = a[3], a11 = a[4], a12 = a[5], a20 = a[6],
              a21 = a[7], a22 = a[8];
  const float det = (a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20)) +
                    a02 * (a10 * a21 - a11 * a20);
  f->dinv = 1.0f / (det + 1e-12f);
  f->adj[0] = a11 * a22 - a12 * a21;
  f->adj[1] = -(a01 * a22 - a02 * a21);
  f->adj[2] = a01 * a12 - a02 * a11;
  f->adj[3] = -(a10 * a22 - a12 * a20);
  f->adj[4] = a00 * a22 - a02 * a20;
  f->adj[5] = -(a00 * a12 - a02 * a10);
  f->adj[6] = a10 * a21 - a11 * a20;
  f->adj[7] = -(a00 * a21 - a01 * a20);
  f->adj[8] = a00 * a11 - a01 * a10;
}

__device__ __forceinline__ float clamp_min0(float x, float lo) { return x < lo ? lo : x; }

// softplus (beta 1, threshold 20) and z / (z + 1)
__device__ __forceinline__ void softplus_d(float x, float* v, float* d) {
  if (x > 20.0f) { *v = x; *d = 1.0f; return; }
  const float z = exp_cw(x);
  *v = log1p_pos(z);
  *d = z / (z + 1.0f);
}

// channel of coordinate slot s (0 alpha, 1 beta, 2 gamma, 3 delta) of facet k;
// with one colour delta's assignment takes gamma's slot (gsmodel.py:264-267)
template <int KC>
__device__ __forceinline__ constexpr int bary_channel(int k, int s) {
  return s == 0 ? 0 : s == 1 ? 1 : s == 2 ? (KC == 1 ? -1 : k + 2) : (k + 1) % KC + 2;
}

struct BaryPix {
  float t[3], c[4], sp[4], dsp[4], den, q[4], rec[3], x, p;
};

// MODE 0: fused loss (W_tilde given; writes dW_tilde, optional W_bary)
// MODE 1: W_bary only
// MODE 2: palette gradient from a given dL/dW_bary (the backward of MODE 1)
template <int KC, int MODE, bool CLAMP, bool ALPHA>
__global__ __launch_bounds__(kBaryThreads) void k_bary(
    int64_t HW, const float* __restrict__ Cimg, const float* __restrict__ Wt, const float* __restrict__ Aimg,
    const float* __restrict__ Pg, float tau, float cg, float cp, float ca, float* __restrict__ wbary,
    float* __restrict__ dWt, const float* __restrict__ dWb_in, double* __restrict__ acc) {
  constexpr int KT = KC + 2;
  constexpr int NACC = KC * kBaryAcc + 3;
  __shared__ BaryFacet sF[KC];
  __shared__ float sP[3 * KC];
  __shared__ float sRed[kBaryThreads / 64][NACC];
  if (threadIdx.x < 3 * KC) sP[threadIdx.x] = Pg[threadIdx.x];
  if (threadIdx.x < KC) bary_facet(Pg, threadIdx.x, (threadIdx.x + 1) % KC, &sF[threadIdx.x]);
  __syncthreads();
  float P[3 * KC];
#pragma unroll
  for (int e = 0; e < 3 * KC; ++e) P[e] = sP[e];

  float gacc[KC * kBaryAcc];
#pragma unroll
  for (int e = 0; e < KC * kBaryAcc; ++e) gacc[e] = 0.f;
  float lgeo = 0.f, lpos = 0.f, lal = 0.f;

  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < HW; n += stride) {
    float rhs[3];
    const float a = ALPHA ? Aimg[n] : 1.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rhs[c] = Cimg[c * HW + n];
      if (ALPHA) {
        const float q = rhs[c] / clamp_min0(a, 1e-3f);
        rhs[c] = q > 1.0f ? 1.0f : clamp_min0(q, 0.0f);
      }
    }
    // ---- forward: facets, softmax, W_bary
    BaryPix px[KC];
    float xmax = 0.f;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const BaryFacet& f = sF[k];
      BaryPix& s = px[k];
#pragma unroll
      for (int r = 0; r < 3; ++r)
        s.t[r] = (f.adj[3 * r] * rhs[0] + f.adj[3 * r + 1] * rhs[1]) + f.adj[3 * r + 2] * rhs[2];
      const float dinv = f.dinv;
      s.c[1] = s.t[0] * dinv;
      s.c[2] = s.t[1] * dinv;
      s.c[3] = s.t[2] * dinv;
      s.c[0] = ((1.0f - s.c[1]) - s.c[2]) - s.c[3];
#pragma unroll
      for (int q = 0; q < 4; ++q) softplus_d(s.c[q], &s.sp[q], &s.dsp[q]);
      s.den = (((s.sp[0] + s.sp[1]) + s.sp[2]) + s.sp[3]) + 1e-12f;
#pragma unroll
      for (int q = 0; q < 4; ++q) s.q[q] = s.sp[q] / s.den;
      const int i = k, j = (k + 1) % KC;
      float err = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        // proj @ [black; white; P]: rows in order, zero rows skipped
        float r = 0.f + s.q[1];
        if (KC == 1) {
          r = r + s.q[3] * P[3 * j + c];
        } else if (i < j) {
          r = r + s.q[2] * P[3 * i + c];
          r = r + s.q[3] * P[3 * j + c];
        } else {
          r = r + s.q[3] * P[3 * j + c];
          r = r + s.q[2] * P[3 * i + c];
        }
        s.rec[c] = r;
        const float d = r - rhs[c];
        err = err + d * d;
      }
      s.x = (-err) / tau;
      xmax = k == 0 ? s.x : (s.x > xmax ? s.x : xmax);
    }
    float psum = 0.f;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      px[k].p = exp_cw(px[k].x - xmax);
      psum = psum + px[k].p;
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) px[k].p = px[k].p / psum;
    float Wb[KT];
#pragma unroll
    for (int m = 0; m < KT; ++m) {
      float s = 0.f;
#pragma unroll
      for (int k = 0; k < KC; ++k)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (bary_channel<KC>(k, q) == m) s = s + px[k].p * px[k].c[q];
      Wb[m] = s;
    }
    if (MODE == 1) {
#pragma unroll
      for (int m = 0; m < KT; ++m) wbary[m * HW + n] = Wb[m];
      continue;
    }
    // ---- loss terms and dL/dW_bary
    float dWb[KT];
    if (MODE == 2) {
#pragma unroll
      for (int m = 0; m < KT; ++m) dWb[m] = dWb_in[m * HW + n];
    } else {
      float Wc[KT], Wf[KT];
      float S = 0.f, Sc = 1.0f;
      if (CLAMP) {
#pragma unroll
        for (int m = 0; m < KT; ++m) {
          Wc[m] = clamp_min0(Wb[m], 0.0f);
          S = S + Wc[m];
        }
        Sc = clamp_min0(S, 1e-8f);
#pragma unroll
        for (int m = 0; m < KT; ++m) Wf[m] = Wc[m] / Sc;
      } else {
#pragma unroll
        for (int m = 0; m < KT; ++m) Wf[m] = Wb[m];
      }
      if (ALPHA) {
#pragma unroll
        for (int m = 0; m < KT; ++m) Wf[m] = Wf[m] * a;
      }
      if (wbary) {
#pragma unroll
        for (int m = 0; m < KT; ++m) wbary[m * HW + n] = Wf[m];
      }
      float g[KT], wt[KT];
      float sg = 0.f, sp = 0.f, wa = 0.f;
#pragma unroll
      for (int m = 0; m < KT; ++m) {
        wt[m] = Wt[m * HW + n];
        g[m] = wt[m] - Wf[m];
        sg = sg + g[m] * g[m];
        sp = sp + clamp_min0(-wt[m], 0.0f);
        wa = wa + wt[m];
      }
      const float da = wa - a;
      lgeo += sg;
      lpos += sp;
      if (ALPHA) lal += da * da;
      float dWf[KT];
#pragma unroll
      for (int m = 0; m < KT; ++m) {
        float d = cg * g[m] - (wt[m] < 0.0f ? cp : 0.0f);
        if (ALPHA) d = d + ca * da;
        dWt[m * HW + n] = d;
        dWf[m] = -(cg * g[m]);
        if (ALPHA) dWf[m] = dWf[m] * a;
      }
      if (CLAMP) {
        float dot = 0.f;
#pragma unroll
        for (int m = 0; m < KT; ++m) dot = dot + dWf[m] * Wc[m];
        const float dS = S >= 1e-8f ? -(dot / (Sc * Sc)) : 0.0f;
#pragma unroll
        for (int m = 0; m < KT; ++m) dWb[m] = Wb[m] >= 0.0f ? dWf[m] / Sc + dS : 0.0f;
      } else {
#pragma unroll
        for (int m = 0; m < KT; ++m) dWb[m] = dWf[m];
      }
    }
    // ---- backward of the barycentric weights
    float dp[KC];
    float sdp = 0.f;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      float s = 0.f;
#pragma unroll
      for (int m = 0; m < KT; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (bary_channel<KC>(k, q) == m) s = s + dWb[m] * px[k].c[q];
      dp[k] = s;
      sdp = sdp + px[k].p * dp[k];
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const BaryPix& s = px[k];
      const int i = k, j = (k + 1) % KC;
      float* gk = gacc + k * kBaryAcc;
      const float dx = s.p * (dp[k] - sdp);
      const float derr = -(dx / tau);
      float drec[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) drec[c] = (2.0f * (s.rec[c] - rhs[c])) * derr;
      constexpr bool gslot = KC != 1;
      float dq[4];
      dq[0] = 0.f;
      dq[1] = (drec[0] + drec[1]) + drec[2];
      dq[2] = gslot ? (drec[0] * P[3 * i] + drec[1] * P[3 * i + 1]) + drec[2] * P[3 * i + 2] : 0.f;
      dq[3] = (drec[0] * P[3 * j] + drec[1] * P[3 * j + 1]) + drec[2] * P[3 * j + 2];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        gk[10 + c] += gslot ? s.q[2] * drec[c] : 0.f;
        gk[13 + c] += s.q[3] * drec[c];
      }
      const float sq = ((dq[0] * s.sp[0] + dq[1] * s.sp[1]) + dq[2] * s.sp[2]) + dq[3] * s.sp[3];
      const float dden = -(sq / (s.den * s.den));
      float dc[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ch = bary_channel<KC>(k, q);
        const float draw = ch >= 0 ? s.p * dWb[ch < 0 ? 0 : ch] : 0.f;
        const float dsp = dq[q] / s.den + dden;
        dc[q] = draw + dsp * s.dsp[q];
      }
      const float db = dc[1] - dc[0], dg = dc[2] - dc[0], dd = dc[3] - dc[0];
      const float dinv = sF[k].dinv;
      const float dt[3] = {db * dinv, dg * dinv, dd * dinv};
      gk[9] += (db * s.t[0] + dg * s.t[1]) + dd * s.t[2];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) gk[3 * r + c] += dt[r] * rhs[c];
    }
  }
  if (MODE == 1) return;
  // ---- block reduction into the float64 accumulators
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  auto wave_total = [&](float v) { return across_rows(row_sum16(v)); };
#pragma unroll
  for (int e = 0; e < KC * kBaryAcc; ++e) {
    const float v = wave_total(gacc[e]);
    if (lane == 0) sRed[w][e] = v;
  }
  {
    const float v0 = wave_total(lgeo), v1 = wave_total(lpos), v2 = wave_total(lal);
    if (lane == 0) {
      sRed[w][KC * kBaryAcc] = v0;
      sRed[w][KC * kBaryAcc + 1] = v1;
      sRed[w][KC * kBaryAcc + 2] = v2;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NACC; e += blockDim.x) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kBaryThreads / 64; ++q) s += (double)sRed[q][e];
    if (s != 0.0) atomicAdd(&acc[e], s);
  }
}

__global__ void k_fill_f64(double* __restrict__ p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// d adj_m / d a_var = sign * a_coef (adj_m = +-(a_p a_q - a_r a_s))
struct BaryAdjTerm {
  signed char m, var, coef, sign;
};
__constant__ BaryAdjTerm kBaryAdj[36] = {
    {0, 4, 8, 1},  {0, 8, 4, 1},  {0, 5, 7, -1}, {0, 7, 5, -1}, {1, 1, 8, -1}, {1, 8, 1, -1},
    {1, 2, 7, 1},  {1, 7, 2, 1},  {2, 1, 5, 1},  {2, 5, 1, 1},  {2, 2, 4, -1}, {2, 4, 2, -1},
    {3, 3, 8, -1}, {3, 8, 3, -1}, {3, 5, 6, 1},  {3, 6, 5, 1},  {4, 0, 8, 1},  {4, 8, 0, 1},
    {4, 2, 6, -1}, {4, 6, 2, -1}, {5, 0, 5, -1}, {5, 5, 0, -1}, {5, 2, 3, 1},  {5, 3, 2, 1},
    {6, 3, 7, 1},  {6, 7, 3, 1},  {6, 4, 6, -1}, {6, 6, 4, -1}, {7, 0, 7, -1}, {7, 7, 0, -1},
    {7, 1, 6, 1},  {7, 6, 1, 1},  {8, 0, 4, 1},  {8, 4, 0, 1},  {8, 1, 3, -1}, {8, 3, 1, -1}};

// Per-facet sums -> dL/dpalette, in double (one thread): d det = -dinv^2
// d dinv, d det / d a_rc = adj[c][r], the adjugate minors, a_r1 = P_i[r],
// a_r2 = P_j[r].  Also the loss: L_geo + lam_pos L_pos + lam_alpha L_alpha.
__global__ void k_bary_finish(int Kc, const float* __restrict__ Pg, const double* __restrict__ acc,
                              double inv_nk, double inv_hw, double lam_pos, double lam_alpha, int alpha,
                              float* __restrict__ dpal, float* __restrict__ loss) {
  if (threadIdx.x != 0) return;
  double dP[18];
  for (int e = 0; e < 3 * Kc; ++e) dP[e] = 0.0;
  for (int k = 0; k < Kc; ++k) {
    BaryFacet f;
    const int i = k, j = (k + 1) % Kc;
    bary_facet(Pg, i, j, &f);
    const double* g = acc + k * kBaryAcc;
    const double di = (double)f.dinv;
    const double ddet = -(di * di) * g[9];
    double da[9];
    for (int v = 0; v < 9; ++v) da[v] = ddet * (double)f.adj[3 * (v % 3) + v / 3];
    for (int t = 0; t < 36; ++t) {
      const BaryAdjTerm& b = kBaryAdj[t];
      da[b.var] += g[b.m] * ((double)b.sign * (double)f.a[b.coef]);
    }
    for (int r = 0; r < 3; ++r) {
      dP[3 * i + r] += da[3 * r + 1] + g[10 + r];
      dP[3 * j + r] += da[3 * r + 2] + g[13 + r];
    }
  }
  for (int e = 0; e < 3 * Kc; ++e) dpal[e] = (float)dP[e];
  if (loss) {
    const double* L = acc + Kc * kBaryAcc;
    const double geo = L[0] * inv_nk, pos = L[1] * inv_nk, al = alpha ? L[2] * inv_hw : 0.0;
    loss[0] = (float)(geo + lam_pos * pos + lam_alpha * al);
    loss[1] = (float)geo;
    loss[2] = (float)pos;
    loss[3] = (float)al;
  }
}

}  // namespace rgp

using namespace rgp;

extern "C" size_t rgp_bary_workspace_bytes(int Kc) {
  return (size_t)(Kc * kBaryAcc + 3) * sizeof(double);
}

static int bary_grid(int64_t hw) {
  const int64_t per = (int64_t)kBaryThreads * 8;  // about 8 pixels per thread
  int64_t b = (hw + per - 1) / per;
  return (int)(b < 1 ? 1 : b > 4096 ? 4096 : b);
}

#define RGP_KC_SWITCH(KCV, CALL)                    \
  switch (KCV) {                                    \
    case 1: { constexpr int KC = 1; CALL; } break;  \
    case 2: { constexpr int KC = 2; CALL; } break;  \
    case 3: { constexpr int KC = 3; CALL; } break;  \
    case 4: { constexpr int KC = 4; CALL; } break;  \
    case 5: { constexpr int KC = 5; CALL; } break;  \
    default: { constexpr int KC = 6; CALL; } break; \
  }

static int bary_args(const char* who, int width, int height, int Kc, float tau) {
  if (width <= 0 || height <= 0 || Kc < 1 || Kc > RGP_MAX_DIM - 2)
    return set_error(RGP_ERR_ARG, "%s: need width, height > 0 and 1 <= Kc <= %d", who, RGP_MAX_DIM - 2);
  if (!(tau > 0.0f)) return set_error(RGP_ERR_ARG, "%s: tau must be positive", who);
  return RGP_OK;
}

// The loss pass over the image with its accumulators at acc (Kc * kBaryAcc + 3
// doubles, cleared here by a fill kernel, not a memset node).
static int bary_loss_launch(const char* who, int width, int height, int Kc, const float* image,
                            const float* w_tilde, const float* gt_alpha, const float* palette, float tau,
                            int clamp, float lambda_pos, float lambda_alpha, float* w_bary, float* dw_tilde,
                            float* dpalette, float* loss, double* acc, hipStream_t s) {
  const int64_t hw = (int64_t)width * height, nk = hw * (Kc + 2);
  const float cg = (float)(2.0 / (double)nk), cp = (float)((double)lambda_pos / (double)nk);
  const float ca = (float)(2.0 * (double)lambda_alpha / (double)hw);
  const int nacc = Kc * kBaryAcc + 3;
  hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, s, acc, nacc);
  const int grid = bary_grid(hw);
#define RGP_BARY0(CL, AL)                                                                           \
  RGP_KC_SWITCH(Kc, hipLaunchKernelGGL((k_bary<KC, 0, CL, AL>), dim3(grid), dim3(kBaryThreads), 0, s, \
                                       hw, image, w_tilde, gt_alpha, palette, tau, cg, cp, ca, w_bary, \
                                       dw_tilde, (const float*)nullptr, acc))
  if (clamp) {
    if (gt_alpha) RGP_BARY0(true, true) else RGP_BARY0(true, false)
  } else {
    if (gt_alpha) RGP_BARY0(false, true) else RGP_BARY0(false, false)
  }
#undef RGP_BARY0
  hipLaunchKernelGGL(k_bary_finish, dim3(1), dim3(64), 0, s, Kc, palette, acc, 1.0 / (double)nk,
                     1.0 / (double)hw, (double)lambda_pos, (double)lambda_alpha, gt_alpha ? 1 : 0,
                     dpalette, loss);
  return check_launch(who);
}

extern "C" int rgp_bary_loss(int width, int height, int Kc, const float* image, const float* w_tilde,
                             const float* gt_alpha, const float* palette, float tau, int clamp,
                             float lambda_pos, float lambda_alpha, float* w_bary, float* dw_tilde,
                             float* dpalette, float* loss, void* workspace, hipStream_t s) {
  if (int rc = bary_args("rgp_bary_loss", width, height, Kc, tau)) return rc;
  if (!image || !w_tilde || !palette || !dw_tilde || !dpalette || !workspace)
    return set_error(RGP_ERR_ARG, "rgp_bary_loss: null buffer");
  return bary_loss_launch("rgp_bary_loss", width, height, Kc, image, w_tilde, gt_alpha, palette, tau, clamp,
                          lambda_pos, lambda_alpha, w_bary, dw_tilde, dpalette, loss, (double*)workspace, s);
}

// ---------------------------------------------------------------------------
// rgp_bary_loss without a workspace argument, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Its float64 accumulators live in `scratch`, a
// file-scope device array of the size the largest palette needs
// ((RGP_MAX_DIM - 2) * kBaryAcc + 3 doubles, 792 bytes); every GPU the
// library runs on holds its own copy until the process ends.  The host-side
// name `scratch` is not a device address: the array's address on the current
// device is looked up with hipGetSymbolAddress on the first call on that
// device and kept per device, and only that pointer reaches kernels.  The
// accumulators are cleared by a fill kernel at the start of every call, and
// calls that overlap on one GPU race on them.
// ---------------------------------------------------------------------------
constexpr int kBaryScratchDoubles = (RGP_MAX_DIM - 2) * kBaryAcc + 3;
static __device__ __align__(256) double scratch[kBaryScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The device address of `scratch` on the current device (which must own s).
static int bary_scratch(const char* who, hipStream_t s, double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return set_error(RGP_ERR_ARG, "%s: no current device (or device index above %d)", who, kMaxDevices - 1);
  if (s != nullptr) {
    hipDevice_t sdev = -1;
    if (hipStreamGetDevice(s, &sdev) != hipSuccess)
      return set_error(RGP_ERR_ARG, "%s: cannot query the stream's device", who);
    if (sdev != dev)
      return set_error(RGP_ERR_ARG, "%s: the stream belongs to device %d, the current device is %d", who,
                       (int)sdev, dev);
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return set_error(RGP_ERR_LAUNCH, "%s: hipGetSymbolAddress(scratch) on device %d: %s", who, dev,
                       hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return RGP_OK;
}

extern "C" int rgp_bary_loss_noalloc(int width, int height, int Kc, const float* image, const float* w_tilde,
                                     const float* gt_alpha, const float* palette, float tau, int clamp,
                                     float lambda_pos, float lambda_alpha, float* w_bary, float* dw_tilde,
                                     float* dpalette, float* loss, hipStream_t s) {
  static const char* who = "rgp_bary_loss_noalloc";
  if (int rc = bary_args(who, width, height, Kc, tau)) return rc;
  if (Kc * kBaryAcc + 3 > kBaryScratchDoubles)
    return set_error(RGP_ERR_ARG, "%s: Kc = %d needs %d accumulators, the scratch holds %d", who, Kc,
                     Kc * kBaryAcc + 3, kBaryScratchDoubles);
  if (!image || !w_tilde || !palette || !dw_tilde || !dpalette)
    return set_error(RGP_ERR_ARG, "%s: null buffer", who);
  double* acc = nullptr;
  if (int rc = bary_scratch(who, s, &acc)) return rc;
  return bary_loss_launch(who, width, height, Kc, image, w_tilde, gt_alpha, palette, tau, clamp, lambda_pos,
                          lambda_alpha, w_bary, dw_tilde, dpalette, loss, acc, s);
}

extern "C" int rgp_bary_coords(int width, int height, int Kc, const float* image, const float* palette,
                               float tau, float* w_bary, hipStream_t s) {
  if (int rc = bary_args("rgp_bary_coords", width, height, Kc, tau)) return rc;
  if (!image || !palette || !w_bary) return set_error(RGP_ERR_ARG, "rgp_bary_coords: null buffer");
  const int64_t hw = (int64_t)width * height;
  RGP_KC_SWITCH(Kc, hipLaunchKernelGGL((k_bary<KC, 1, false, false>), dim3(bary_grid(hw)), dim3(kBaryThreads),
                                       0, s, hw, image, (const float*)nullptr, (const float*)nullptr, palette,
                                       tau, 0.f, 0.f, 0.f, w_bary, (float*)nullptr, (const float*)nullptr,
                                       (double*)nullptr));
  return check_launch("rgp_bary_coords");
}

extern "C" int rgp_bary_coords_backward(int width, int height, int Kc, const float* image,
                                        const float* palette, float tau, const float* dw_bary,
                                        float* dpalette, void* workspace, hipStream_t s) {
  if (int rc = bary_args("rgp_bary_coords_backward", width, height, Kc, tau)) return rc;
  if (!image || !palette || !dw_bary || !dpalette || !workspace)
    return set_error(RGP_ERR_ARG, "rgp_bary_coords_backward: null buffer");
  const int64_t hw = (int64_t)width * height;
  double* acc = (double*)workspace;
  hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, s, acc, Kc * kBaryAcc + 3);
  RGP_KC_SWITCH(Kc, hipLaunchKernelGGL((k_bary<KC, 2, false, false>), dim3(bary_grid(hw)), dim3(kBaryThreads),
                                       0, s, hw, image, (const float*)nullptr, (const float*)nullptr, palette,
                                       tau, 0.f, 0.f, 0.f, (float*)nullptr, (float*)nullptr, dw_bary, acc));
  hipLaunchKernelGGL(k_bary_finish, dim3(1), dim3(64), 0, s, Kc, palette, acc, 0.0, 0.0, 0.0, 0.0, 0,
                     dpalette, (float*)nullptr);
  return check_launch("rgp_bary_coords_backward");
}
