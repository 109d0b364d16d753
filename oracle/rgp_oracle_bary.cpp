// CPU oracle of the barycentric palette loss of the finetune step.
//
// TEST INFRASTRUCTURE ONLY (like rgp_oracle.cpp, whose library it joins): only
// tests, the smoke check and the benchmark's baseline leg load it.
//
// It restates barycentric_coords_3d (gsplat/gsmodel.py:219-285) and the
// per-pixel loss terms the finetune step builds on it (finetune.py:353-392):
//   * with a ground-truth alpha image a: C <- clamp(C / max(a, 1e-3), 0, 1);
//   * W_bary = barycentric_coords_3d(C, palette), K_c chromatic colours plus
//     black and white, one tetrahedron (black, white, P_k, P_{k+1}) per k;
//   * optionally W_bary <- max(W_bary, 0) / max(sum_m max(W_bary, 0), 1e-8);
//   * with a: W_bary <- W_bary * a, and L_alpha = mean((sum_m W_tilde - a)^2);
//   * L_geo = mean((W_tilde - W_bary)^2), L_pos = mean(relu(-W_tilde)).
// The gradients of L_geo + lam_pos L_pos + lam_alpha L_alpha go to W_tilde
// (per pixel) and, through W_bary, to the palette (summed over pixels).
//
// T = float: FP32 in source order (-ffp-contract=off).  exp is rgp_expf of
// rgp_oracle.cpp, log1p is rgp_log1pf below; the HIP kernel performs the same
// operations, so FP32 results agree bit for bit.  Where torch leaves an order
// open (the einsum, the matmul, the sums), the order is the one written here.
// T = double: std::exp and std::log1p, the reference for accuracy and for the
// finite-difference checks of the backward.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {
float rgp_expf(float x);
}

namespace orcb {

static inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// log(1 + y) for y >= 0 (an exp output).  u = fl(1 + y) with its exact
// rounding error e (Fast2Sum); log u = n ln2 + log m, m in [sqrt(1/2),
// sqrt(2)), log m = 2 atanh(s) = 2 (s + s^3/3 + s^5/5 + s^7/7 + s^9/9) with
// s = (m - 1)/(m + 1), |s| < 0.172; then log(1 + y) = log u + e / u.
float rgp_log1pf(float y) {
  if (y != y) return y;
  if (y == u2f(0x7f800000u)) return y;
  const float u = 1.0f + y;
  const float e = y >= 1.0f ? 1.0f - (u - y) : y - (u - 1.0f);
  const uint32_t b = f2u(u);
  int n = (int)(b >> 23) - 127;
  float m = u2f((b & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421356f) { m = 0.5f * m; n += 1; }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  float p = 1.0f / 9.0f;
  p = std::fma(p, z, 1.0f / 7.0f);
  p = std::fma(p, z, 0.2f);
  p = std::fma(p, z, 1.0f / 3.0f);
  const float t = 2.0f * s;
  const float lm = std::fma(t * z, p, t);
  const float fn = (float)n;
  const float lu = std::fma(fn, 0.693145751953125f, std::fma(fn, 1.42860676533018570e-06f, lm));
  return lu + e / u;
}

template <typename T> static inline T ex(T x);
template <> inline float ex<float>(float x) { return orc::rgp_expf(x); }
template <> inline double ex<double>(double x) { return std::exp(x); }
template <typename T> static inline T l1p(T x);
template <> inline float l1p<float>(float x) { return rgp_log1pf(x); }
template <> inline double l1p<double>(double x) { return std::log1p(x); }

// torch.nn.functional.softplus (beta 1, threshold 20) and its derivative
// z / (z + 1), z = exp(x), as torch's backward forms it.
template <typename T> static inline void softplus(T x, T* v, T* d) {
  if (x > T(20)) { *v = x; *d = T(1); return; }
  const T z = ex<T>(x);
  *v = l1p<T>(z);
  *d = z / (z + T(1));
}

// torch.clamp(min=lo): NaN passes through.
template <typename T> static inline T clamp_min(T x, T lo) { return x < lo ? lo : x; }

template <typename T> struct Facet {
  T a[9], adj[9], det, dinv;
  int i, j;
};

// gsmodel.py:225-252: A_k = [white - black, P_i - black, P_j - black] as
// columns, j = (i + 1) % K_c; det by the first row, adjugate as listed.
template <typename T> static void make_facets(int Kc, const T* P, T eps, Facet<T>* F) {
  for (int k = 0; k < Kc; ++k) {
    Facet<T>& f = F[k];
    f.i = k;
    f.j = (k + 1) % Kc;
    T* a = f.a;
    for (int r = 0; r < 3; ++r) {
      a[3 * r + 0] = T(1) - T(0);
      a[3 * r + 1] = P[3 * f.i + r] - T(0);
      a[3 * r + 2] = P[3 * f.j + r] - T(0);
    }
    const T a00 = a[0], a01 = a[1], a02 = a[2], a10 = a[3], a11 = a[4], a12 = a[5], a20 = a[6],
            a21 = a[7], a22 = a[8];
    f.det = (a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20)) +
            a02 * (a10 * a21 - a11 * a20);
    f.dinv = T(1) / (f.det + eps);
    f.adj[0] = a11 * a22 - a12 * a21;
    f.adj[1] = -(a01 * a22 - a02 * a21);
    f.adj[2] = a01 * a12 - a02 * a11;
    f.adj[3] = -(a10 * a22 - a12 * a20);
    f.adj[4] = a00 * a22 - a02 * a20;
    f.adj[5] = -(a00 * a12 - a02 * a10);
    f.adj[6] = a10 * a21 - a11 * a20;
    f.adj[7] = -(a00 * a21 - a01 * a20);
    f.adj[8] = a00 * a11 - a01 * a10;
  }
}

constexpr int kMaxKc = 6;  // K_c + 2 <= 8 palette channels
constexpr int kAcc = 16;   // per facet: d adj (9), d dinv, dP_i (3), dP_j (3)

// One facet's state at one pixel.
template <typename T> struct Pix {
  T t[3];        // adj * rhs
  T c[4];        // alpha, beta, gamma, delta
  T sp[4], dsp[4], den, q[4];
  T rec[3], x, p;
};

// raw / proj index of coordinate slot s (0 alpha, 1 beta, 2 gamma, 3 delta);
// gamma loses its slot to delta when i == j (gsmodel.py:264-267).
static inline int slot_channel(int s, int i, int j) {
  return s == 0 ? 0 : s == 1 ? 1 : s == 2 ? (i == j ? -1 : i + 2) : j + 2;
}

// gsmodel.py:231-283 for one pixel: coordinates, softplus projection,
// reconstruction error, softmax over facets, W = sum_k p_k raw_k.
template <typename T>
static void pixel_forward(int Kc, const Facet<T>* F, const T* P, const T* rhs, T tau, T eps, Pix<T>* px,
                          T* Wb) {
  const int Kt = Kc + 2;
  T xmax = T(0);
  for (int k = 0; k < Kc; ++k) {
    const Facet<T>& f = F[k];
    Pix<T>& s = px[k];
    for (int r = 0; r < 3; ++r)
      s.t[r] = (f.adj[3 * r] * rhs[0] + f.adj[3 * r + 1] * rhs[1]) + f.adj[3 * r + 2] * rhs[2];
    s.c[1] = s.t[0] * f.dinv;
    s.c[2] = s.t[1] * f.dinv;
    s.c[3] = s.t[2] * f.dinv;
    s.c[0] = ((T(1) - s.c[1]) - s.c[2]) - s.c[3];
    for (int q = 0; q < 4; ++q) softplus<T>(s.c[q], &s.sp[q], &s.dsp[q]);
    s.den = (((s.sp[0] + s.sp[1]) + s.sp[2]) + s.sp[3]) + eps;
    for (int q = 0; q < 4; ++q) s.q[q] = s.sp[q] / s.den;
    // recon = proj @ P_aug over the Kt rows in order (zero rows included)
    T proj[8];
    for (int m = 0; m < Kt; ++m) proj[m] = T(0);
    for (int q = 0; q < 4; ++q) {
      const int ch = slot_channel(q, f.i, f.j);
      if (ch >= 0) proj[ch] = s.q[q];
    }
    T err = T(0);
    for (int cc = 0; cc < 3; ++cc) {
      T acc = T(0);
      for (int m = 0; m < Kt; ++m) {
        const T pa = m == 0 ? T(0) : m == 1 ? T(1) : P[3 * (m - 2) + cc];
        acc = acc + proj[m] * pa;
      }
      s.rec[cc] = acc;
      const T d = acc - rhs[cc];
      err = err + d * d;
    }
    s.x = (-err) / tau;
    xmax = k == 0 ? s.x : (s.x > xmax ? s.x : xmax);
  }
  T sum = T(0);
  for (int k = 0; k < Kc; ++k) {
    px[k].p = ex<T>(px[k].x - xmax);
    sum = sum + px[k].p;
  }
  for (int k = 0; k < Kc; ++k) px[k].p = px[k].p / sum;
  for (int m = 0; m < Kt; ++m) {
    T acc = T(0);
    for (int k = 0; k < Kc; ++k) {
      T raw = T(0);
      for (int q = 0; q < 4; ++q)
        if (slot_channel(q, F[k].i, F[k].j) == m) raw = px[k].c[q];
      acc = acc + px[k].p * raw;
    }
    Wb[m] = acc;
  }
}

// Backward of pixel_forward for dL/dW (Kt): adds this pixel's terms to the
// per-facet accumulators g[k*kAcc + ...] (d adj, d dinv, dP_i, dP_j).
template <typename T>
static void pixel_backward(int Kc, const Facet<T>* F, const T* P, const T* rhs, T tau, const Pix<T>* px,
                           const T* dW, T* g) {
  T dp[kMaxKc];
  T sdp = T(0);
  for (int k = 0; k < Kc; ++k) {
    T acc = T(0);
    for (int m = 0; m < Kc + 2; ++m) {
      T raw = T(0);
      for (int q = 0; q < 4; ++q)
        if (slot_channel(q, F[k].i, F[k].j) == m) raw = px[k].c[q];
      acc = acc + dW[m] * raw;
    }
    dp[k] = acc;
    sdp = sdp + px[k].p * dp[k];
  }
  for (int k = 0; k < Kc; ++k) {
    const Facet<T>& f = F[k];
    const Pix<T>& s = px[k];
    T* gk = g + k * kAcc;
    const T dx = s.p * (dp[k] - sdp);
    const T derr = -(dx / tau);
    T drec[3];
    for (int cc = 0; cc < 3; ++cc) drec[cc] = (T(2) * (s.rec[cc] - rhs[cc])) * derr;
    const bool gslot = f.i != f.j;
    T dq[4];
    dq[0] = T(0);
    dq[1] = (drec[0] + drec[1]) + drec[2];
    dq[2] = gslot ? (drec[0] * P[3 * f.i] + drec[1] * P[3 * f.i + 1]) + drec[2] * P[3 * f.i + 2] : T(0);
    dq[3] = (drec[0] * P[3 * f.j] + drec[1] * P[3 * f.j + 1]) + drec[2] * P[3 * f.j + 2];
    for (int cc = 0; cc < 3; ++cc) {
      gk[10 + cc] = gk[10 + cc] + (gslot ? s.q[2] * drec[cc] : T(0));
      gk[13 + cc] = gk[13 + cc] + s.q[3] * drec[cc];
    }
    const T sq = ((dq[0] * s.sp[0] + dq[1] * s.sp[1]) + dq[2] * s.sp[2]) + dq[3] * s.sp[3];
    const T dden = -(sq / (s.den * s.den));
    T dc[4];
    for (int q = 0; q < 4; ++q) {
      const int ch = slot_channel(q, f.i, f.j);
      const T draw = ch >= 0 ? s.p * dW[ch] : T(0);
      const T dsp = dq[q] / s.den + dden;
      dc[q] = draw + dsp * s.dsp[q];
    }
    const T db = dc[1] - dc[0], dg = dc[2] - dc[0], dd = dc[3] - dc[0];
    const T dt[3] = {db * f.dinv, dg * f.dinv, dd * f.dinv};
    gk[9] = gk[9] + ((db * s.t[0] + dg * s.t[1]) + dd * s.t[2]);
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) gk[3 * r + cc] = gk[3 * r + cc] + dt[r] * rhs[cc];
  }
}

// d adj_m / d a_rc for the adjugate of make_facets, as (m, rc, coefficient
// index into a, sign): adj_m = +-(a_p a_q - a_r a_s).
struct AdjTerm { int m, var, coef, sign; };
static const AdjTerm kAdjTerms[36] = {
    {0, 4, 8, 1},  {0, 8, 4, 1},  {0, 5, 7, -1}, {0, 7, 5, -1},   // a11 a22 - a12 a21
    {1, 1, 8, -1}, {1, 8, 1, -1}, {1, 2, 7, 1},  {1, 7, 2, 1},    // -(a01 a22 - a02 a21)
    {2, 1, 5, 1},  {2, 5, 1, 1},  {2, 2, 4, -1}, {2, 4, 2, -1},   // a01 a12 - a02 a11
    {3, 3, 8, -1}, {3, 8, 3, -1}, {3, 5, 6, 1},  {3, 6, 5, 1},    // -(a10 a22 - a12 a20)
    {4, 0, 8, 1},  {4, 8, 0, 1},  {4, 2, 6, -1}, {4, 6, 2, -1},   // a00 a22 - a02 a20
    {5, 0, 5, -1}, {5, 5, 0, -1}, {5, 2, 3, 1},  {5, 3, 2, 1},    // -(a00 a12 - a02 a10)
    {6, 3, 7, 1},  {6, 7, 3, 1},  {6, 4, 6, -1}, {6, 6, 4, -1},   // a10 a21 - a11 a20
    {7, 0, 7, -1}, {7, 7, 0, -1}, {7, 1, 6, 1},  {7, 6, 1, 1},    // -(a00 a21 - a01 a20)
    {8, 0, 4, 1},  {8, 4, 0, 1},  {8, 1, 3, -1}, {8, 3, 1, -1}};  // a00 a11 - a01 a10

// Per-facet sums -> dL/dpalette (Kc,3), in double: through dinv = 1/(det+eps)
// (d det = -dinv^2 d dinv; d det / d a_rc = adj[c][r]), the adjugate minors,
// and the entries a_r1 = P_i[r], a_r2 = P_j[r].  With abs = true the map
// takes absolute values (a bound on the sum of |terms|).
template <typename T>
static void finalize(int Kc, const Facet<T>* F, const double* acc, bool abs_, double* dP) {
  for (int e = 0; e < 3 * Kc; ++e) dP[e] = 0.0;
  for (int k = 0; k < Kc; ++k) {
    const Facet<T>& f = F[k];
    const double* g = acc + k * kAcc;
    const double di = (double)f.dinv;
    double ddet = -(di * di) * g[9];
    if (abs_) ddet = std::fabs(ddet);
    double da[9];
    for (int v = 0; v < 9; ++v) {
      const int r = v / 3, c = v % 3;
      const double cof = (double)f.adj[3 * c + r];
      da[v] = ddet * (abs_ ? std::fabs(cof) : cof);
    }
    for (const AdjTerm& t : kAdjTerms) {
      const double c = t.sign * (double)f.a[t.coef];
      da[t.var] += g[t.m] * (abs_ ? std::fabs(c) : c);
    }
    for (int r = 0; r < 3; ++r) {
      dP[3 * f.i + r] += da[3 * r + 1] + g[10 + r];
      dP[3 * f.j + r] += da[3 * r + 2] + g[13 + r];
    }
  }
}

template <typename T>
static void bary_loss(int H, int W, int Kc, const T* C, const T* Wt, const T* A, const T* P, T tau, int clamp,
                      T lam_pos, T lam_alpha, T* wbary, T* dWt, double* parts, double* dpal, double* dpal_abs,
                      double* facet_acc) {
  const int Kt = Kc + 2;
  const int64_t HW = (int64_t)H * W, NK = HW * Kt;
  const T eps = T(1e-12);
  Facet<T> F[kMaxKc];
  make_facets<T>(Kc, P, eps, F);
  const T cg = (T)(2.0 / (double)NK);
  const T cp = (T)((double)lam_pos / (double)NK);
  const T ca = (T)(2.0 * (double)lam_alpha / (double)HW);
  std::vector<double> acc(Kc * kAcc, 0.0), aab(Kc * kAcc, 0.0);
  double geo = 0.0, pos = 0.0, al = 0.0;
  for (int64_t n = 0; n < HW; ++n) {
    T rhs[3];
    const T a = A ? A[n] : T(1);
    for (int c = 0; c < 3; ++c) {
      rhs[c] = C[c * HW + n];
      if (A) {
        const T q = rhs[c] / clamp_min(a, T(1e-3));
        rhs[c] = q > T(1) ? T(1) : clamp_min(q, T(0));
      }
    }
    Pix<T> px[kMaxKc];
    T Wb[8], Wc[8], Wf[8], dWf[8], dWb[8];
    pixel_forward<T>(Kc, F, P, rhs, tau, eps, px, Wb);
    T S = T(0), Sc = T(1);
    if (clamp) {
      for (int m = 0; m < Kt; ++m) {
        Wc[m] = clamp_min(Wb[m], T(0));
        S = S + Wc[m];
      }
      Sc = clamp_min(S, T(1e-8));
      for (int m = 0; m < Kt; ++m) Wf[m] = Wc[m] / Sc;
    } else {
      for (int m = 0; m < Kt; ++m) Wf[m] = Wb[m];
    }
    if (A)
      for (int m = 0; m < Kt; ++m) Wf[m] = Wf[m] * a;
    if (wbary)
      for (int m = 0; m < Kt; ++m) wbary[m * HW + n] = Wf[m];
    T sg = T(0), sp = T(0), wa = T(0);
    T g[8];
    for (int m = 0; m < Kt; ++m) {
      const T wt = Wt[m * HW + n];
      g[m] = wt - Wf[m];
      sg = sg + g[m] * g[m];
      sp = sp + clamp_min(-wt, T(0));
      wa = wa + wt;
    }
    const T da = wa - a;
    geo += (double)sg;
    pos += (double)sp;
    if (A) al += (double)(da * da);
    for (int m = 0; m < Kt; ++m) {
      const T wt = Wt[m * HW + n];
      T d = cg * g[m] - (wt < T(0) ? cp : T(0));
      if (A) d = d + ca * da;
      if (dWt) dWt[m * HW + n] = d;
      dWf[m] = -(cg * g[m]);
      if (A) dWf[m] = dWf[m] * a;
    }
    if (clamp) {
      T dot = T(0);
      for (int m = 0; m < Kt; ++m) dot = dot + dWf[m] * Wc[m];
      const T dS = S >= T(1e-8) ? -(dot / (Sc * Sc)) : T(0);
      for (int m = 0; m < Kt; ++m) dWb[m] = Wb[m] >= T(0) ? dWf[m] / Sc + dS : T(0);
    } else {
      for (int m = 0; m < Kt; ++m) dWb[m] = dWf[m];
    }
    T gp[kMaxKc * kAcc];
    for (int e = 0; e < Kc * kAcc; ++e) gp[e] = T(0);
    pixel_backward<T>(Kc, F, P, rhs, tau, px, dWb, gp);
    for (int e = 0; e < Kc * kAcc; ++e) {
      acc[e] += (double)gp[e];
      aab[e] += std::fabs((double)gp[e]);
    }
  }
  parts[0] = geo;
  parts[1] = pos;
  parts[2] = al;
  finalize<T>(Kc, F, acc.data(), false, dpal);
  finalize<T>(Kc, F, aab.data(), true, dpal_abs);
  if (facet_acc)
    for (int e = 0; e < Kc * kAcc; ++e) facet_acc[e] = acc[e];
}

// The backward of barycentric_coords_3d alone: dL/dpalette for a given
// dL/dW_bary (Kc+2, H, W), summed over pixels in double.
template <typename T>
static void bary_coords_backward(int H, int W, int Kc, const T* C, const T* P, T tau, const T* dWb,
                                 double* dpal, double* dpal_abs) {
  const int Kt = Kc + 2;
  const int64_t HW = (int64_t)H * W;
  const T eps = T(1e-12);
  Facet<T> F[kMaxKc];
  make_facets<T>(Kc, P, eps, F);
  std::vector<double> acc(Kc * kAcc, 0.0), aab(Kc * kAcc, 0.0);
  for (int64_t n = 0; n < HW; ++n) {
    T rhs[3], Wb[8], dW[8];
    for (int c = 0; c < 3; ++c) rhs[c] = C[c * HW + n];
    for (int m = 0; m < Kt; ++m) dW[m] = dWb[m * HW + n];
    Pix<T> px[kMaxKc];
    pixel_forward<T>(Kc, F, P, rhs, tau, eps, px, Wb);
    T gp[kMaxKc * kAcc];
    for (int e = 0; e < Kc * kAcc; ++e) gp[e] = T(0);
    pixel_backward<T>(Kc, F, P, rhs, tau, px, dW, gp);
    for (int e = 0; e < Kc * kAcc; ++e) {
      acc[e] += (double)gp[e];
      aab[e] += std::fabs((double)gp[e]);
    }
  }
  finalize<T>(Kc, F, acc.data(), false, dpal);
  finalize<T>(Kc, F, aab.data(), true, dpal_abs);
}

}  // namespace orcb

extern "C" {
float orc_log1pf(float y) { return orcb::rgp_log1pf(y); }
void orc_log1pf_array(int64_t n, const float* y, float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orcb::rgp_log1pf(y[i]);
}
#define ORC_BARY(T, S)                                                                             \
  void orc_bary_loss_##S(int h, int w, int Kc, const T* C, const T* Wt, const T* A, const T* P, T tau, \
                         int clamp, T lam_pos, T lam_alpha, T* wbary, T* dWt, double* parts,       \
                         double* dpal, double* dpal_abs, double* facet_acc) {                      \
    orcb::bary_loss<T>(h, w, Kc, C, Wt, A, P, tau, clamp, lam_pos, lam_alpha, wbary, dWt, parts, dpal, \
                       dpal_abs, facet_acc);                                                       \
  }
ORC_BARY(float, f)
ORC_BARY(double, d)
void orc_bary_coords_backward_f(int h, int w, int Kc, const float* C, const float* P, float tau, const float* dWb,
                                double* dpal, double* dpal_abs) {
  orcb::bary_coords_backward<float>(h, w, Kc, C, P, tau, dWb, dpal, dpal_abs);
}
void orc_bary_coords_backward_d(int h, int w, int Kc, const double* C, const double* P, double tau,
                                const double* dWb, double* dpal, double* dpal_abs) {
  orcb::bary_coords_backward<double>(h, w, Kc, C, P, tau, dWb, dpal, dpal_abs);
}
}
