// Tile compositing kernels for gfx950: the forward of draw_weight/draw_color
// (gsplatcu/kernel.cu:194-372), the reverse-order backward of draw_weight_B
// (kernel.cu:1291-1379) and the palette mix.
//
// A 256-thread workgroup owns one 16x16 tile; wavefront w covers rows
// 4w..4w+3 (64 pixels).  The tile's list is staged in LDS 256 entries at a
// time as packed splat records; each wavefront then walks the batch on its
// own, without the per-entry workgroup barrier of the reference.  A pixel
// whose transmittance drops below 1e-4 stops; a wavefront whose pixels have
// all stopped leaves the batch loop, and the workgroup leaves when all its
// pixels have stopped (one barrier count per batch).  Per entry and pixel the
// Mahalanobis distance is compared with the record's cutoff first, so exp()
// runs only where alpha' can reach 0.002.  Per pixel the arithmetic is the
// reference recurrence, operation for operation, so results do not depend on
// any of this scheduling.
#include "rgp_internal.h"

namespace rgp {

// COUNT: also accumulate, over the launch, the list entries each pixel
// evaluates before it stops (counts[0]) and its contributions (counts[1]).
template <int DIM, bool COUNT = false>
__global__ __launch_bounds__(256) void k_draw_fwd(int W, int H, int gx, int tiles,
                                                  const int2* __restrict__ ranges,
                                                  const int32_t* __restrict__ ids,
                                                  const float4* __restrict__ recs, float empty_tau,
                                                  float* __restrict__ layer,
                                                  int32_t* __restrict__ contrib,
                                                  float* __restrict__ final_tau,
                                                  unsigned long long* __restrict__ counts = nullptr) {
  constexpr int F4 = RecStride<DIM>::f4;
  __shared__ float4 s_rec[256 * F4];
  const int tile = xcd_tile(blockIdx.x, tiles);
  const int ty = tile / gx, tx = tile - ty * gx;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int px = tx * kTile + (lane & 15), py = ty * kTile + w * 4 + (lane >> 4);
  const bool inside = px < W && py < H;
  const int2 rg = ranges[tile];
  const float fpx = (float)px, fpy = (float)py;
  const float ycen = (float)(ty * kTile + w * 4) + 1.5f;  // centre of this wavefront's rows

  float acc[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) acc[k] = 0.f;
  float tau = 1.f;
  int cont = 0;
  bool done = !inside;
  int n_eval = 0, n_con = 0;  // COUNT only
  if (rg.y - rg.x <= 0) {
    tau = empty_tau;  // the reference kernel returns before writing
  } else {
    for (int b0 = rg.x; b0 < rg.y; b0 += 256) {
      if (__syncthreads_count(done) == 256) break;
      const int e = b0 + t;
      if (e < rg.y) {
        const float4* src = recs + (size_t)ids[e] * F4;
#pragma unroll
        for (int q = 0; q < F4; ++q) s_rec[t * F4 + q] = src[q];
      }
      __syncthreads();
      const int cnt = min(256, rg.y - b0);
      for (int j = 0; j < cnt; ++j) {
        if (__all(done)) break;
        const float4 g0 = s_rec[j * F4];
        const float4 g1 = s_rec[j * F4 + 1];
        // wavefront-uniform: the entry cannot reach any of this strip's rows
        if ((g0.y - ycen > g1.w + 1.5f) || (ycen - g0.y > g1.w + 1.5f)) continue;
        if (done) continue;
        const float dx = g0.x - fpx, dy = g0.y - fpy;
        const float m = maha_dist(g0, g1.x, dx, dy);
        if (m > g1.z) continue;
        const float ap = fminf(0.99f, g1.y * exp_cw(-0.5f * m));
        if (ap < 0.002f) continue;
        const float ta = tau * ap;
#pragma unroll
        for (int q = 0; q < F4 - 2; ++q) {
          const float4 f = s_rec[j * F4 + 2 + q];
          if (4 * q + 0 < DIM) acc[4 * q + 0] = acc[4 * q + 0] + ta * f.x;
          if (4 * q + 1 < DIM) acc[4 * q + 1] = acc[4 * q + 1] + ta * f.y;
          if (4 * q + 2 < DIM) acc[4 * q + 2] = acc[4 * q + 2] + ta * f.z;
          if (4 * q + 3 < DIM) acc[4 * q + 3] = acc[4 * q + 3] + ta * f.w;
        }
        cont = b0 + j - rg.x + 1;
        tau = tau * (1.f - ap);
        if (COUNT) ++n_con;
        if (tau < 0.0001f) {
          done = true;
          if (COUNT) n_eval = cont;
        }
      }
    }
    if (COUNT && inside && !done) n_eval = rg.y - rg.x;
  }
  if (inside) {
    const size_t HW = (size_t)W * H, pix = (size_t)py * W + px;
#pragma unroll
    for (int k = 0; k < DIM; ++k) layer[k * HW + pix] = acc[k];
    contrib[pix] = cont;
    final_tau[pix] = tau;
  }
  if (COUNT) {
    unsigned long long e = (unsigned long long)n_eval, c = (unsigned long long)n_con;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      e += __shfl_xor(e, o);
      c += __shfl_xor(c, o);
    }
    if (lane == 0) {
      atomicAdd(&counts[0], e);
      atomicAdd(&counts[1], c);
    }
  }
}

// Backward.  Entries are visited in reverse list order from the largest
// contrib of the workgroup (each wavefront starts at its own largest contrib);
// per pixel the transmittance is recovered by division as in the reference.
// The K gradient terms of an entry are summed over the wavefront with DPP,
// the four wavefront sums go to LDS, and after each batch one thread per
// entry adds their total to the Gaussian's gradient: one atomic per
// (workgroup, entry, channel) instead of one per (pixel, entry, channel).
template <int DIM>
__global__ __launch_bounds__(256) void k_draw_bwd(int W, int H, int gx, int tiles,
                                                  const int2* __restrict__ ranges,
                                                  const int32_t* __restrict__ ids,
                                                  const float4* __restrict__ recs,
                                                  const int32_t* __restrict__ contrib,
                                                  const float* __restrict__ final_tau,
                                                  const float* __restrict__ dlayer,
                                                  float* __restrict__ dfeat) {
  constexpr int F4 = RecStride<DIM>::f4;
  __shared__ float4 s_rec[256 * F4];
  __shared__ float s_part[4][DIM][256];
  __shared__ int32_t s_id[256];
  __shared__ int32_t s_max[4];
  const int tile = xcd_tile(blockIdx.x, tiles);
  const int ty = tile / gx, tx = tile - ty * gx;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int px = tx * kTile + (lane & 15), py = ty * kTile + w * 4 + (lane >> 4);
  const bool inside = px < W && py < H;
  const int2 rg = ranges[tile];
  if (rg.y - rg.x <= 0) return;
  const float fpx = (float)px, fpy = (float)py;
  const float ycen = (float)(ty * kTile + w * 4) + 1.5f;

  float dl[DIM];
  float tau = 0.f;
  int cont = 0;
  const size_t HW = (size_t)W * H, pix = (size_t)py * W + px;
#pragma unroll
  for (int k = 0; k < DIM; ++k) dl[k] = inside ? dlayer[k * HW + pix] : 0.f;
  if (inside) {
    tau = final_tau[pix];
    cont = contrib[pix];
  }
  int wmax = cont;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) wmax = max(wmax, __shfl_xor(wmax, o));
  if (lane == 0) s_max[w] = wmax;
  __syncthreads();
  const int maxc = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
  for (int top = rg.x + maxc; top > rg.x; top -= 256) {
    const int lo = max(rg.x, top - 256);
    const int cnt = top - lo;
    const int e = top - 1 - t;
    if (e >= lo) {
      const int g = ids[e];
      s_id[t] = g;
      const float4* src = recs + (size_t)g * F4;
#pragma unroll
      for (int q = 0; q < F4; ++q) s_rec[t * F4 + q] = src[q];
    }
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      s_part[0][k][t] = 0.f; s_part[1][k][t] = 0.f; s_part[2][k][t] = 0.f; s_part[3][k][t] = 0.f;
    }
    __syncthreads();
    // entries with forward index >= wmax hold no contributor of this wavefront
    const int jstart = max(0, top - rg.x - wmax);
    for (int j = jstart; j < cnt; ++j) {
      const int fi = top - 1 - j - rg.x;  // index of the entry in forward order
      const float4 g0 = s_rec[j * F4];
      const float4 g1 = s_rec[j * F4 + 1];
      if ((g0.y - ycen > g1.w + 1.5f) || (ycen - g0.y > g1.w + 1.5f)) continue;
      bool on = false;
      float gv = 0.f;
      if (fi < cont) {
        const float dx = g0.x - fpx, dy = g0.y - fpy;
        const float m = maha_dist(g0, g1.x, dx, dy);
        if (!(m > g1.z)) {
          const float ap = fminf(0.99f, g1.y * exp_cw(-0.5f * m));
          if (!(ap < 0.002f)) {
            tau = tau / (1.f - ap);
            gv = ap * tau;
            on = true;
          }
        }
      }
      if (__any(on)) {
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
          const float sum = wave_sum_to63(on ? dl[k] * gv : 0.f);
          if (lane == 63) s_part[w][k][j] = sum;
        }
      }
    }
    __syncthreads();
    if (t < cnt) {
      const int g = s_id[t];
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        const float v = (s_part[0][k][t] + s_part[1][k][t]) + (s_part[2][k][t] + s_part[3][k][t]);
        if (v != 0.f) unsafeAtomicAdd(&dfeat[(size_t)g * DIM + k], v);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// palette mix
// ---------------------------------------------------------------------------
struct MixArgs {
  const float* const* W;
  float* const* rgb;
  const int64_t* hw;
  const float* pal;
};

template <int K>
__global__ __launch_bounds__(256) void k_mix(MixArgs a) {
  const int img = blockIdx.y;
  const int64_t hw = a.hw[img];
  const float* __restrict__ Wp = a.W[img];
  float* __restrict__ out = a.rgb[img];
  float P[K][3];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) P[k][c] = a.pal[((size_t)img * K + k) * 3 + c];
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

template <int K>
__global__ __launch_bounds__(256) void k_mix_bwd(int64_t hw, const float* __restrict__ Wp,
                                                 const float* __restrict__ pal,
                                                 const float* __restrict__ drgb,
                                                 float* __restrict__ dW,
                                                 float* __restrict__ dpal) {
  __shared__ float s_part[4][K * 3];
  float P[K][3];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) P[k][c] = pal[k * 3 + c];
  float gp[K][3];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) gp[k][c] = 0.f;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < hw; p += stride) {
    const float d0 = drgb[p], d1 = drgb[hw + p], d2 = drgb[2 * hw + p];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const float wk = Wp[k * hw + p];
      dW[k * hw + p] = (P[k][0] * d0 + P[k][1] * d1) + P[k][2] * d2;
      gp[k][0] = gp[k][0] + wk * d0;
      gp[k][1] = gp[k][1] + wk * d1;
      gp[k][2] = gp[k][2] + wk * d2;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float s = wave_sum_to63(gp[k][c]);
      if (lane == 63) s_part[w][k * 3 + c] = s;
    }
  __syncthreads();
  if (threadIdx.x < K * 3) {
    const float s = (s_part[0][threadIdx.x] + s_part[1][threadIdx.x]) +
                    (s_part[2][threadIdx.x] + s_part[3][threadIdx.x]);
    unsafeAtomicAdd(&dpal[threadIdx.x], s);
  }
}

}  // namespace rgp

using namespace rgp;

#define RGP_DIM_SWITCH(D, CALL)                    \
  switch (D) {                                     \
    case 1: { constexpr int DD = 1; CALL; } break; \
    case 2: { constexpr int DD = 2; CALL; } break; \
    case 3: { constexpr int DD = 3; CALL; } break; \
    case 4: { constexpr int DD = 4; CALL; } break; \
    case 5: { constexpr int DD = 5; CALL; } break; \
    case 6: { constexpr int DD = 6; CALL; } break; \
    case 7: { constexpr int DD = 7; CALL; } break; \
    case 8: { constexpr int DD = 8; CALL; } break; \
    default: break;                                \
  }

extern "C" int rgp_draw_forward(int width, int height, int dim, const int32_t* ranges,
                                const int32_t* ids, const float* records, float empty_tau,
                                float* layer, int32_t* contrib, float* final_tau,
                                hipStream_t s) {
  if (width <= 0 || height <= 0 || dim < 1 || dim > RGP_MAX_DIM)
    return set_error(RGP_ERR_ARG, "rgp_draw_forward: bad size or dim");
  const int gx = (width + kTile - 1) / kTile, gy = (height + kTile - 1) / kTile;
  const int tiles = gx * gy;
  RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd<DD, false>), dim3(tiles), dim3(256), 0, s,
                                         width, height, gx, tiles, (const int2*)ranges, ids,
                                         (const float4*)records, empty_tau, layer, contrib,
                                         final_tau, (unsigned long long*)nullptr));
  return check_launch("rgp_draw_forward");
}

extern "C" int rgp_draw_forward_counts(int width, int height, int dim, const int32_t* ranges,
                                       const int32_t* ids, const float* records, float empty_tau,
                                       float* layer, int32_t* contrib, float* final_tau,
                                       int64_t* counts, hipStream_t s) {
  if (width <= 0 || height <= 0 || dim < 1 || dim > RGP_MAX_DIM)
    return set_error(RGP_ERR_ARG, "rgp_draw_forward_counts: bad size or dim");
  const int gx = (width + kTile - 1) / kTile, gy = (height + kTile - 1) / kTile;
  const int tiles = gx * gy;
  RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd<DD, true>), dim3(tiles), dim3(256), 0, s, width,
                                         height, gx, tiles, (const int2*)ranges, ids,
                                         (const float4*)records, empty_tau, layer, contrib,
                                         final_tau, (unsigned long long*)counts));
  return check_launch("rgp_draw_forward_counts");
}

extern "C" int rgp_draw_backward(int width, int height, int dim, const int32_t* ranges,
                                 const int32_t* ids, const float* records, const int32_t* contrib,
                                 const float* final_tau, const float* dlayer, float* dfeat,
                                 hipStream_t s) {
  if (width <= 0 || height <= 0 || dim < 1 || dim > RGP_MAX_DIM)
    return set_error(RGP_ERR_ARG, "rgp_draw_backward: bad size or dim");
  const int gx = (width + kTile - 1) / kTile, gy = (height + kTile - 1) / kTile;
  const int tiles = gx * gy;
  RGP_DIM_SWITCH(dim, hipLaunchKernelGGL(k_draw_bwd<DD>, dim3(tiles), dim3(256), 0, s, width,
                                         height, gx, tiles, (const int2*)ranges, ids,
                                         (const float4*)records, contrib, final_tau, dlayer,
                                         dfeat));
  return check_launch("rgp_draw_backward");
}

extern "C" int rgp_mix_batched(int count, int K, const float* const* W, float* const* rgb,
                               const int64_t* hw, const float* palettes, int64_t max_hw,
                               hipStream_t s) {
  if (count < 0 || K < 1 || K > RGP_MAX_DIM || max_hw < 0)
    return set_error(RGP_ERR_ARG, "rgp_mix_batched: bad count, K or size");
  if (count == 0 || max_hw == 0) return RGP_OK;
  if (count > 65535) return set_error(RGP_ERR_ARG, "rgp_mix_batched: at most 65535 images");
  MixArgs a{W, rgb, hw, palettes};
  const unsigned gxm = (unsigned)((max_hw + 1023) / 1024);
  RGP_DIM_SWITCH(K, hipLaunchKernelGGL(k_mix<DD>, dim3(gxm, count), dim3(256), 0, s, a));
  return check_launch("rgp_mix_batched");
}

extern "C" int rgp_mix_backward(int K, int64_t hw, const float* W, const float* palette,
                                const float* drgb, float* dW, float* dpalette, hipStream_t s) {
  if (K < 1 || K > RGP_MAX_DIM || hw < 0)
    return set_error(RGP_ERR_ARG, "rgp_mix_backward: bad K or size");
  if (hw == 0) return RGP_OK;
  int64_t blocks = (hw + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  RGP_DIM_SWITCH(K, hipLaunchKernelGGL(k_mix_bwd<DD>, dim3((unsigned)blocks), dim3(256), 0, s, hw,
                                       W, palette, drgb, dW, dpalette));
  return check_launch("rgp_mix_backward");
}
