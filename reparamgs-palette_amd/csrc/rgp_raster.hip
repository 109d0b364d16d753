// Tile compositing kernels for gfx950: the forward of draw_weight/draw_color
// (gsplatcu/kernel.cu:194-372), the reverse-order backward of draw_weight_B
// (kernel.cu:1291-1379), a front-to-back replay backward, and the palette mix.
//
// Layout.  One wavefront (a 64-thread workgroup) owns one 16x16 tile.  Lane l
// holds the column px = x0 + (l & 15) in four row slots: slot j covers row
// y0 + 4j + (l >> 4), so slot j across the wavefront is a 16x4 strip.  The
// tile's list is staged in LDS 64 entries at a time (one per lane; the
// backward prefetches the next batch into registers while the current one is
// composited, the forward only its ids); no barrier beyond the wavefront is
// needed.
//
// Work per entry.  The terms of the Mahalanobis distance that depend only on
// dx, (cinv.x dx) dx and (2 cinv.y) dx, are formed once per lane and reused by
// its four pixels; per pixel the distance is
// fma((2 cinv.y) dx, dy, fma(cinv.z dy, dy, (cinv.x dx) dx)) and the weight
// update W_k = fma(T alpha', w_k, W_k): mahaSqDist and the accumulation of
// kernel.cu:351-361 with their multiply-adds contracted (nvcc's default),
// the same operations the CPU oracle performs, so every forward value is
// bit-identical to it.
// While staging, each lane also works out which of the four strips its entry
// can reach (|dy| within the record's cutoff radius ry); the wavefront reads
// that 4-bit mask into scalar registers and skips unreachable strips with a
// scalar branch.  A pixel then compares the distance with the record's cutoff
// mcut before evaluating exp().  A pixel has stopped when its transmittance
// is below 1e-4 (pixels outside the image start at 0).
#include "rgp_internal.h"

namespace rgp {

constexpr int kSlots = 4;

// A splat record held in registers.
template <int F4>
struct RecRegs {
  float4 v[F4];
};

// Loads the record of list entry e when begin <= e < end (else zeros).
template <int F4>
__device__ __forceinline__ RecRegs<F4> load_rec(const float4* __restrict__ recs,
                                                const int32_t* __restrict__ ids, int e, int begin,
                                                int end, int* id) {
  RecRegs<F4> r;
  const bool ok = e >= begin && e < end;
  const int g = ok ? ids[e] : 0;
  *id = g;
  const float4* src = recs + (size_t)g * F4;
#pragma unroll
  for (int q = 0; q < F4; ++q) r.v[q] = ok ? src[q] : make_float4(0.f, 0.f, 0.f, 0.f);
  return r;
}

// Bit of the staged mask word set when the record's cutoff mcut is not at
// most 170 (or is NaN).  Below it every distance that passes m <= mcut lies
// in [0, 170], where gauss_of(m) is gauss_half_fast(m), so the range test is
// a scalar branch per entry instead of a per-lane one.  The forward and the
// replay backward use only alpha' = min(0.99, alpha exp(-m/2)), which
// gauss_half gives exactly for every record (see mcut_of), so a wide entry
// takes gauss_half and every other one gauss_half_fast; the colour backward
// also uses exp(-m/2) itself.
constexpr uint32_t kWideExp = 1u << 4;

// Stages one record in LDS, replacing its ry by the mask of strips of tile
// row ty it can reach (bit j: rows ty*16 + 4j .. + 3) and kWideExp, and
// returns that mask word (the compositing loops read it from the staging
// lane's register with v_readlane, not from LDS).  Float4 q of slot s
// sits at s_rec[q * 64 + s]: the wave's 16-byte stores cover the banks once
// per quarter-wave, and a read of one slot is a broadcast.
template <int F4>
__device__ __forceinline__ uint32_t stage_rec(float4* s_rec, int slot, const RecRegs<F4>& r, int ty) {
  uint32_t mask = r.v[1].z <= 170.0f ? 0u : kWideExp;
  const float uy = r.v[0].y, ry = r.v[1].w;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const float yc = (float)(ty * kTile + 4 * j) + 1.5f;
    if (!((uy - yc > ry + 1.5f) || (yc - uy > ry + 1.5f))) mask |= 1u << j;
  }
  s_rec[slot] = r.v[0];
  s_rec[64 + slot] = make_float4(r.v[1].x, r.v[1].y, r.v[1].z, __uint_as_float(mask));
#pragma unroll
  for (int q = 2; q < F4; ++q) s_rec[q * 64 + slot] = r.v[q];
  return mask;
}

// The mask word of staged entry jj (wave-uniform jj), from the register of the
// lane that staged it.
__device__ __forceinline__ uint32_t entry_bits(uint32_t mask, int jj) {
  return __builtin_amdgcn_readlane(mask, jj);
}

// exp(-m/2) for an m that passed m <= mcut; wide (wavefront-uniform) is the
// entry's kWideExp bit, so the branch is a scalar one.
__device__ __forceinline__ float gauss_in(float m, bool wide) {
  if (__builtin_expect(wide, 0)) return gauss_of(m);
  return gauss_half_fast(m);
}


__device__ __forceinline__ uint32_t uniform_bits(float v) {
  return __builtin_amdgcn_readfirstlane(__float_as_uint(v));
}

// Mask of slots in which some lane's pixel is still compositing.
__device__ __forceinline__ uint32_t live_slots(const float* tau) {
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j)
    if (__any(!(tau[j] < 0.0001f))) m |= 1u << j;
  return m;
}

// ---------------------------------------------------------------------------
// forward.  COUNT: also accumulate, over the launch, the list entries each
// pixel evaluates before it stops (counts[0]) and its contributions
// (counts[1]).  STRIPS: also write, for every list entry the tile composites
// (every entry of each batch it stages), the mask of the tile's four row
// strips in which some pixel took a contribution from it (strips[e], one
// byte per list position).  The replay backward, which repeats exactly this
// arithmetic, then evaluates only those strips.  The flags are collected in
// LDS: the contributing lanes of strip j store byte j of the entry's flag
// word (a store in the contribution branch, at the address already formed for
// the entry's weights), and after the batch each lane folds the four bytes of
// its entry into the mask.  CONT = false: the contribution index is neither
// tracked nor written (the training forward, whose replay backward reads the
// strip masks and has no use for contrib).
// ---------------------------------------------------------------------------
template <int DIM, bool COUNT, bool STRIPS = false, bool CONT = true>
__device__ __forceinline__ void draw_fwd_tile(int W, int H, int gx, int tile, const int2* __restrict__ ranges,
                                              const int32_t* __restrict__ ids, const float4* __restrict__ recs,
                                              float empty_tau, float* __restrict__ layer,
                                              int32_t* __restrict__ contrib, float* __restrict__ final_tau,
                                              unsigned long long* __restrict__ counts,
                                              uint8_t* __restrict__ strips, float4* s_rec) {
  constexpr int F4 = RecStride<DIM>::f4;
  const int ty = tile / gx, tx = tile - ty * gx;
  const int lane = threadIdx.x;
  const int px = tx * kTile + (lane & 15);
  const int y0 = ty * kTile + (lane >> 4);
  float fpx = (float)px;
  asm volatile("" : "+v"(fpx));
  const int2 rg = ranges[tile];

  float acc[kSlots][DIM];
  float tau[kSlots], fpy[kSlots];
  int cont[kSlots];
  int n_eval[kSlots], n_con[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) acc[j][k] = 0.f;
    const bool in = px < W && y0 + 4 * j < H;
    tau[j] = in ? 1.f : 0.f;
    fpy[j] = (float)(y0 + 4 * j);
    asm volatile("" : "+v"(fpy[j]));  // kept as a float, not converted again per evaluation
    cont[j] = 0;
    n_eval[j] = 0;
    n_con[j] = 0;
  }
  // STRIPS: byte j of s_flag[jj] is set to 1 when some pixel of strip j took
  // a contribution from entry jj of the staged batch (an LDS byte store from
  // the contributing lanes, at the address of the entry's weights).
  uint32_t* s_flag = reinterpret_cast<uint32_t*>(s_rec + F4 * 64);
  if (STRIPS) s_flag[4 * lane] = 0;
  if (rg.y - rg.x > 0) {
    // Only the next batch's ids are fetched ahead: the records are gathered at
    // the start of their batch (other wavefronts cover that latency), which
    // keeps 16 registers free for the per-pixel state.
    int nid = rg.x + lane < rg.y ? ids[rg.x + lane] : -1;
    for (int b0 = rg.x; b0 < rg.y; b0 += 64) {
      const uint32_t live = live_slots(tau);
      if (live == 0) break;
      RecRegs<F4> cur;
      {
        const float4* src = recs + (size_t)max(nid, 0) * F4;
#pragma unroll
        for (int q = 0; q < F4; ++q) cur.v[q] = nid >= 0 ? src[q] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
      nid = b0 + 64 + lane < rg.y ? ids[b0 + 64 + lane] : -1;
      __syncthreads();  // previous batch fully read
      const uint32_t smask = stage_rec<F4>(s_rec, lane, cur, ty);
      __syncthreads();
      const int cnt = __builtin_amdgcn_readfirstlane(min(64, rg.y - b0));
      for (int jj = 0; jj < cnt; ++jj) {
        const uint32_t ebits = entry_bits(smask, jj);
        const uint32_t slots = ebits & live;
        if (slots == 0) continue;
        // Byte offset of the entry's staged record (and of its flag word), held
        // in one register for the record reads and the strips' flag stores
        // (rematerialised from the scalar jj, it cost one instruction per
        // contributing strip, and a second copy for the reads one per entry).
        int foff = 16 * jj;
        if (STRIPS) asm volatile("" : "+v"(foff));
        const float4* s_ent = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(s_rec) + foff);
        const float4 g0 = s_ent[0];
        const float4 g1 = s_ent[64];
        const float dx = g0.x - fpx;
        const float ta0 = g0.z * dx * dx;  // (cinv.x dx) dx
        const float tc0 = g0.w * dx;       // (2 cinv.y) dx
        const int idx = b0 + jj - rg.x + 1;  // contrib value of a pixel this entry contributes to
        // The entry's channel values, read once for its strips: the contribution
        // step then waits on no LDS read.
        float wv[DIM];
#pragma unroll
        for (int q = 0; q < F4 - 2; ++q) {
          const float4 f = s_ent[(2 + q) * 64];
          if (4 * q + 0 < DIM) wv[4 * q + 0] = f.x;
          if (4 * q + 1 < DIM) wv[4 * q + 1] = f.y;
          if (4 * q + 2 < DIM) wv[4 * q + 2] = f.z;
          if (4 * q + 3 < DIM) wv[4 * q + 3] = f.w;
        }
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
          if (!(slots & (1u << j))) continue;
          const float dy = g0.y - fpy[j];
          const float m = fmaxf(0.0f, fmaf(tc0, dy, fmaf(g1.x * dy, dy, ta0)));
          if (tau[j] < 0.0001f || m > g1.z) continue;  // stopped pixel, or beyond the cutoff
          // exp(-m/2) in the form the entry's cutoff allows (a scalar branch;
          // the loop stays one copy, whose registers the compiler keeps in place)
          float e = gauss_half_fast(m);
          if (__builtin_expect((ebits & kWideExp) != 0u, 0)) e = gauss_half(m);
          const float ap = fminf(0.99f, g1.y * e);
          if (ap < 0.002f) continue;
          const float ta = tau[j] * ap;
#pragma unroll
          for (int k = 0; k < DIM; ++k) acc[j][k] = fmaf(ta, wv[k], acc[j][k]);
          if (CONT || COUNT) cont[j] = idx;
          tau[j] = tau[j] * (1.f - ap);
          if (STRIPS) reinterpret_cast<uint8_t*>(s_flag)[foff + j] = 1;
          if (COUNT) {
            ++n_con[j];
            if (tau[j] < 0.0001f) n_eval[j] = cont[j];
          }
        }
      }
      if (STRIPS) {
        // bytes b0..b3 (each 0 or 1) of entry `lane` -> the 4-bit mask b0 | b1<<1 | b2<<2 | b3<<3:
        // the product with 0x01020408 places byte j's bit at 24 + j and every other partial
        // product below bit 24 or beyond bit 31
        const uint32_t f = s_flag[4 * lane];
        if (lane < cnt) strips[b0 + lane] = (uint8_t)((f * 0x01020408u) >> 24);
        s_flag[4 * lane] = 0;
      }
    }
  }
  const size_t HW = (size_t)W * H;
  unsigned long long ce = 0, cc = 0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int py = y0 + 4 * j;
    if (px < W && py < H) {
      const size_t pix = (size_t)py * W + px;
#pragma unroll
      for (int k = 0; k < DIM; ++k) layer[k * HW + pix] = acc[j][k];
      if (CONT) contrib[pix] = cont[j];
      // the reference kernel returns before writing when the list is empty
      final_tau[pix] = rg.y - rg.x > 0 ? tau[j] : empty_tau;
      if (COUNT) {
        ce += (rg.y - rg.x > 0 && !(tau[j] < 0.0001f)) ? (unsigned long long)(rg.y - rg.x)
                                                       : (unsigned long long)n_eval[j];
        cc += n_con[j];
      }
    }
  }
  if (COUNT) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ce += __shfl_xor(ce, o);
      cc += __shfl_xor(cc, o);
    }
    if (lane == 0) {
      atomicAdd(&counts[0], ce);
      atomicAdd(&counts[1], cc);
    }
  }
}

// Wavefronts per SIMD the compositing kernels are compiled for: the most at
// which the register allocation of the instance still needs no spills (every
// instance compiled at 8, 7, 6 and 5 wavefronts).  These loops are VALU-bound
// with an LDS round trip per staged batch and a record gather per batch, and
// more wavefronts hide more of that latency: the training forward and the
// replay backward at K = 5 went from seven wavefronts (66 and 71 VGPRs) to
// eight (64) without extra instructions.
template <int DIM, bool COUNT, bool CONT>
constexpr int fwd_waves() {
  if (COUNT) return DIM <= 3 ? 8 : DIM == 4 ? 7 : DIM <= 6 ? 6 : 5;
  if (!CONT) return DIM <= 5 ? 8 : DIM <= 7 ? 7 : 6;
  return DIM <= 4 ? 8 : DIM <= 6 ? 7 : 6;
}
template <int DIM, bool STRIPS>
constexpr int bwd_waves() {
  if (STRIPS) return DIM <= 5 ? 8 : DIM <= 7 ? 7 : 6;
  return DIM <= 4 ? 8 : DIM <= 6 ? 7 : DIM == 7 ? 6 : 5;
}

template <int DIM, bool COUNT = false, bool STRIPS = false, bool CONT = true>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(fwd_waves<DIM, COUNT, CONT>())))
void k_draw_fwd(
    int W, int H, int gx, int tiles,
    const int2* __restrict__ ranges, const int32_t* __restrict__ ids, const float4* __restrict__ recs,
    float empty_tau, float* __restrict__ layer, int32_t* __restrict__ contrib, float* __restrict__ final_tau,
    unsigned long long* __restrict__ counts, uint8_t* __restrict__ strips) {
  __shared__ float4 s_rec[64 * (RecStride<DIM>::f4 + (STRIPS ? 1 : 0))];
  draw_fwd_tile<DIM, COUNT, STRIPS, CONT>(W, H, gx, xcd_tile(blockIdx.x, tiles), ranges, ids, recs, empty_tau,
                                          layer, contrib, final_tau, counts, strips, s_rec);
}

// `views` images of one resolution in one launch: block b composites tile
// t % tiles of view t / tiles (t the XCD-banded position of b); view v reads
// ranges + v*tiles, ids + v*id_stride, recs + v*rec_stride (float4 units)
// and writes the v-th (DIM,H,W) layer and (H,W) contrib / final_tau (and,
// STRIPS, its strip masks at strips + v*id_stride).
template <int DIM, bool STRIPS = false, bool CONT = true>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(fwd_waves<DIM, false, CONT>())))
void k_draw_fwd_views(
    int W, int H, int gx, int tiles, int views,
    const int2* __restrict__ ranges, const int32_t* __restrict__ ids, int64_t id_stride,
    const float4* __restrict__ recs, int64_t rec_stride, float empty_tau, float* __restrict__ layer,
    int32_t* __restrict__ contrib, float* __restrict__ final_tau, uint8_t* __restrict__ strips) {
  __shared__ float4 s_rec[64 * (RecStride<DIM>::f4 + (STRIPS ? 1 : 0))];
  const int t = xcd_tile(blockIdx.x, tiles * views);
  const int v = t / tiles, tile = t - v * tiles;
  const size_t HW = (size_t)W * H;
  draw_fwd_tile<DIM, false, STRIPS, CONT>(W, H, gx, tile, ranges + (size_t)v * tiles, ids + v * id_stride,
                                          recs + v * rec_stride, empty_tau, layer + (size_t)v * DIM * HW,
                                          CONT ? contrib + v * HW : nullptr, final_tau + v * HW, nullptr,
                                    STRIPS ? strips + v * id_stride : nullptr, s_rec);
}

// Adds the four-slot lane sums of an entry's gradient terms into dfeat: a
// reduce-scatter over the wavefront leaves channel ch = ws_channel(DIM, lane)
// in the lane, and the owner lanes (one per channel) issue one atomic
// instruction over the entry's DIM floats.  ch and owner are per-lane
// constants set up once per kernel (grad_lanes).
template <int DIM>
__device__ __forceinline__ void flush_grad(const float* part, float* __restrict__ dfeat, int g, int lane,
                                           int ch, bool owner) {
  const float v = wave_reduce_scatter<DIM>(part, lane);
  // The Gaussian's row (g is wavefront-uniform) is formed in scalar registers
  // and pinned there, so the atomic takes it as its scalar base with the
  // lane's 32-bit channel offset; otherwise the compiler folds the lane's
  // channel into a 64-bit vector multiply-add per flush.
  typedef __attribute__((address_space(1))) float gfloat;
  gfloat* row = (gfloat*)(dfeat + (size_t)__builtin_amdgcn_readfirstlane(g) * DIM);
  asm volatile("" : "+s"(row));
  if (owner && v != 0.f)
    __hip_atomic_fetch_add(row + (uint32_t)ch, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sets the per-entry partial sums to zero, two registers per v_mov_b64.  The
// zeros come from asm, so the compiler neither folds them into the first
// slot's multiply-adds nor rematerialises them on every path through the
// slots.
template <int DIM>
__device__ __forceinline__ void zero_part(float* part) {
#pragma unroll
  for (int k = 0; k + 1 < DIM; k += 2) {
    uint64_t z;
    asm volatile("v_mov_b64 %0, 0" : "=v"(z));
    part[k] = __uint_as_float((uint32_t)z);
    part[k + 1] = __uint_as_float((uint32_t)(z >> 32));
  }
  if constexpr ((DIM & 1) != 0) {
    part[DIM - 1] = 0.f;
    asm volatile("" : "+v"(part[DIM - 1]));
  }
}

template <int DIM>
__device__ __forceinline__ void grad_lanes(int lane, int* ch, bool* owner) {
  *ch = (int)((WsLanes<DIM>::chan8 >> (4 * ws_lane_index(lane))) & 15u);
  *owner = (WsLanes<DIM>::owner >> lane) & 1ull;
}

// ---------------------------------------------------------------------------
// backward, reference form (draw_weight_B): entries in reverse list order
// from the tile's largest contrib; a pixel joins at its own contrib and
// recovers the transmittance by division from final_tau.
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(64) void k_draw_bwd(int W, int H, int gx, int tiles,
                                                 const int2* __restrict__ ranges,
                                                 const int32_t* __restrict__ ids,
                                                 const float4* __restrict__ recs,
                                                 const int32_t* __restrict__ contrib,
                                                 const float* __restrict__ final_tau,
                                                 const float* __restrict__ dlayer,
                                                 float* __restrict__ dfeat) {
  constexpr int F4 = RecStride<DIM>::f4;
  __shared__ float4 s_rec[64 * F4];
  __shared__ int32_t s_id[64];
  const int tile = xcd_tile(blockIdx.x, tiles);
  const int ty = tile / gx, tx = tile - ty * gx;
  const int lane = threadIdx.x;
  const int px = tx * kTile + (lane & 15);
  const int y0 = ty * kTile + (lane >> 4);
  const float fpx = (float)px;
  const int2 rg = ranges[tile];
  int ch;
  bool owner;
  grad_lanes<DIM>(lane, &ch, &owner);
  if (rg.y - rg.x <= 0) return;

  const size_t HW = (size_t)W * H;
  float dl[kSlots][DIM];
  float tau[kSlots], fpy[kSlots];
  int cont[kSlots];
  int maxc = 0;
  uint32_t smax_bits[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int py = y0 + 4 * j;
    const bool in = px < W && py < H;
    const size_t pix = (size_t)py * W + px;
#pragma unroll
    for (int k = 0; k < DIM; ++k) dl[j][k] = in ? dlayer[k * HW + pix] : 0.f;
    tau[j] = in ? final_tau[pix] : 0.f;
    cont[j] = in ? contrib[pix] : 0;
    fpy[j] = (float)py;
    int c = cont[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c = max(c, __shfl_xor(c, o));
    smax_bits[j] = (uint32_t)c;
    maxc = max(maxc, c);
  }
  int nid = 0;
  RecRegs<F4> nxt = load_rec<F4>(recs, ids, rg.x + maxc - 1 - lane, rg.x, rg.x + maxc, &nid);
  for (int top = rg.x + maxc; top > rg.x; top -= 64) {
    const int lo = max(rg.x, top - 64);
    const int cnt = top - lo;
    __syncthreads();
    if (lane < cnt) {
      stage_rec<F4>(s_rec, lane, nxt, ty);
      s_id[lane] = nid;
    }
    __syncthreads();
    nxt = load_rec<F4>(recs, ids, top - 65 - lane, rg.x, rg.y, &nid);
    for (int jj = 0; jj < cnt; ++jj) {
      const int fi = top - 1 - jj - rg.x;  // index of the entry in forward order
      const float4 g0 = s_rec[jj];
      const float4 g1 = s_rec[64 + jj];
      uint32_t slots = uniform_bits(g1.w) & 15u;
#pragma unroll
      for (int j = 0; j < kSlots; ++j)
        if ((uint32_t)fi >= smax_bits[j]) slots &= ~(1u << j);
      if (slots == 0) continue;
      const float dx = g0.x - fpx;
      const float ta0 = g0.z * dx * dx;
      const float tc0 = g0.w * dx;
      float part[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) part[k] = 0.f;
      bool on = false;
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        if (!(slots & (1u << j))) continue;
        if (fi >= cont[j]) continue;
        const float dy = g0.y - fpy[j];
        const float m = fmaxf(0.0f, fmaf(tc0, dy, fmaf(g1.x * dy, dy, ta0)));
        if (m > g1.z) continue;
        const float ap = fminf(0.99f, g1.y * gauss_half(m));
        if (ap < 0.002f) continue;
        tau[j] = tau[j] / (1.f - ap);
        const float gv = ap * tau[j];
#pragma unroll
        for (int k = 0; k < DIM; ++k) part[k] = fmaf(dl[j][k], gv, part[k]);
        on = true;
      }
      if (__any(on)) flush_grad<DIM>(part, dfeat, s_id[jj], lane, ch, owner);
    }
  }
}

// ---------------------------------------------------------------------------
// Colour backward with geometry gradients (vanilla drawB,
// vanilla_3dgs/gsplatcu/kernel.cu:811-971): the reference form's reverse
// traversal with tau by division from final_tau, plus the colour behind each
// entry, gamma = a' c + (1 - a') gamma.  Per contributing (pixel, entry):
//   dL/da' = dot(dL/dgamma, tau (c - gamma)) + dL/dtau_f (-tau_f / (1 - a'))
//   du = dL/da' ((-cx dx - cy dy) a', (-cy dx - cz dy) a')
//   dcinv = dL/da' (-0.5 a' dx^2, -a' dx dy, -0.5 a' dy^2)
//   dalpha = dL/da' g,  dcolor = dL/dgamma a' tau
// (the 0.99 clamp and the max(0, .) are not differentiated, as in the
// reference).  The nine per-entry sums go out as one atomic instruction from
// lanes 0..8, each lane to its own array.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void flush_geom(const float* part, float* __restrict__ dus, float* __restrict__ dcinv,
                                           float* __restrict__ dalpha, float* __restrict__ dcolor, int g,
                                           int lane, int ch, bool owner) {
  const float v = across_rows(row_reduce_scatter<9>(part, lane));
  if (owner && v != 0.f) {
    float* dst = ch < 2 ? dus + (size_t)g * 2 + ch
               : ch < 5 ? dcinv + (size_t)g * 3 + (ch - 2)
               : ch == 5 ? dalpha + g
                         : dcolor + (size_t)g * 3 + (ch - 6);
    unsafeAtomicAdd(dst, v);
  }
}

__global__ __launch_bounds__(64) void k_draw_color_bwd(
    int W, int H, int gx, int tiles, const int2* __restrict__ ranges, const int32_t* __restrict__ ids,
    const float4* __restrict__ recs, const int32_t* __restrict__ contrib, const float* __restrict__ final_tau,
    const float* __restrict__ dimg, const float* __restrict__ dtau, float* __restrict__ dus,
    float* __restrict__ dcinv, float* __restrict__ dalpha, float* __restrict__ dcolor) {
  constexpr int F4 = RecStride<3>::f4;
  __shared__ float4 s_rec[64 * F4];
  __shared__ int32_t s_id[64];
  const int tile = xcd_tile(blockIdx.x, tiles);
  const int ty = tile / gx, tx = tile - ty * gx;
  const int lane = threadIdx.x;
  const int px = tx * kTile + (lane & 15);
  const int y0 = ty * kTile + (lane >> 4);
  const float fpx = (float)px;
  const int2 rg = ranges[tile];
  const int ch = (int)((RsLanes<9>::chan >> (4 * (lane & 15))) & 15);
  const bool owner = lane < 16 && ((RsLanes<9>::owner >> lane) & 1);
  if (rg.y - rg.x <= 0) return;

  const size_t HW = (size_t)W * H;
  float dl[kSlots][3], gam[kSlots][3];
  float tau[kSlots], ft[kSlots], dlt[kSlots], fpy[kSlots];
  int cont[kSlots];
  int maxc = 0;
  uint32_t smax_bits[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int py = y0 + 4 * j;
    const bool in = px < W && py < H;
    const size_t pix = (size_t)py * W + px;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dl[j][k] = in ? dimg[k * HW + pix] : 0.f;
      gam[j][k] = 0.f;
    }
    ft[j] = in ? final_tau[pix] : 0.f;
    dlt[j] = in && dtau ? dtau[pix] : 0.f;
    tau[j] = ft[j];
    cont[j] = in ? contrib[pix] : 0;
    fpy[j] = (float)py;
    int c = cont[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c = max(c, __shfl_xor(c, o));
    smax_bits[j] = (uint32_t)c;
    maxc = max(maxc, c);
  }
  int nid = 0;
  RecRegs<F4> nxt = load_rec<F4>(recs, ids, rg.x + maxc - 1 - lane, rg.x, rg.x + maxc, &nid);
  for (int top = rg.x + maxc; top > rg.x; top -= 64) {
    const int lo = max(rg.x, top - 64);
    const int cnt = top - lo;
    __syncthreads();
    if (lane < cnt) {
      stage_rec<F4>(s_rec, lane, nxt, ty);
      s_id[lane] = nid;
    }
    __syncthreads();
    nxt = load_rec<F4>(recs, ids, top - 65 - lane, rg.x, rg.y, &nid);
    for (int jj = 0; jj < cnt; ++jj) {
      const int fi = top - 1 - jj - rg.x;  // index of the entry in forward order
      const float4 g0 = s_rec[jj];
      const float4 g1 = s_rec[64 + jj];
      const float4 g2 = s_rec[128 + jj];
      const uint32_t bits = uniform_bits(g1.w);
      uint32_t slots = bits & 15u;
#pragma unroll
      for (int j = 0; j < kSlots; ++j)
        if ((uint32_t)fi >= smax_bits[j]) slots &= ~(1u << j);
      if (slots == 0) continue;
      const bool wide = bits & kWideExp;
      const float cx = g0.z, cy = 0.5f * g0.w, cz = g1.x;
      const float dx = g0.x - fpx;
      const float ta0 = cx * dx * dx;
      const float tc0 = g0.w * dx;
      const float col[3] = {g2.x, g2.y, g2.z};
      float part[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) part[k] = 0.f;
      bool on = false;
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        if (!(slots & (1u << j))) continue;
        if (fi >= cont[j]) continue;
        const float dy = g0.y - fpy[j];
        const float m = fmaxf(0.0f, fmaf(tc0, dy, fmaf(cz * dy, dy, ta0)));
        if (m > g1.z) continue;
        const float gg = gauss_in(m, wide);
        const float ap = fminf(0.99f, g1.y * gg);
        if (ap < 0.002f) continue;
        tau[j] = tau[j] / (1.f - ap);
        const float t = tau[j];
        const float dg0 = t * (col[0] - gam[j][0]), dg1 = t * (col[1] - gam[j][1]), dg2 = t * (col[2] - gam[j][2]);
        const float dap_c = (dl[j][0] * dg0 + dl[j][1] * dg1) + dl[j][2] * dg2;
        const float dap = dap_c + dlt[j] * (-ft[j] / (1.f - ap));
        const float at = ap * t;
        part[0] += dap * ((-cx * dx - cy * dy) * ap);
        part[1] += dap * ((-cy * dx - cz * dy) * ap);
        part[2] += dap * ((-0.5f * ap) * (dx * dx));
        part[3] += dap * ((-1.0f * ap) * (dx * dy));
        part[4] += dap * ((-0.5f * ap) * (dy * dy));
        part[5] += dap * gg;
#pragma unroll
        for (int k = 0; k < 3; ++k) part[6 + k] += dl[j][k] * at;
#pragma unroll
        for (int k = 0; k < 3; ++k) gam[j][k] = ap * col[k] + (1.f - ap) * gam[j][k];
        on = true;
      }
      if (__any(on)) flush_geom(part, dus, dcinv, dalpha, dcolor, s_id[jj], lane, ch, owner);
    }
  }
}

// ---------------------------------------------------------------------------
// backward, replay form: the forward recurrence is run again front to back;
// the term of a contributing entry is dL/dW_k * tau * alpha' with tau the
// transmittance in front of it, the factor the forward applied to the
// weights, so no per-pixel state from the forward is read.  The set of
// contributing (pixel, entry) pairs is the forward's (same stop rule).
// STRIPS: the forward's strip masks (draw_fwd_tile) replace the reachable
// strips of each entry.  The replay repeats the forward's arithmetic, so a
// strip without a contribution there has none here and is skipped, and an
// entry with a nonzero mask always has a term to flush.
// ---------------------------------------------------------------------------
template <int DIM, bool STRIPS>
__device__ __forceinline__ void draw_bwd_replay_tile(int W, int H, int gx, int tile,
                                                     const int2* __restrict__ ranges,
                                                     const int32_t* __restrict__ ids,
                                                     const float4* __restrict__ recs,
                                                     const uint8_t* __restrict__ strips,
                                                     const float* __restrict__ dlayer,
                                                     float* __restrict__ dfeat, float4* s_rec) {
  constexpr int F4 = RecStride<DIM>::f4;
  const int ty = tile / gx, tx = tile - ty * gx;
  const int lane = threadIdx.x;
  const int px = tx * kTile + (lane & 15);
  const int y0 = ty * kTile + (lane >> 4);
  const float fpx = (float)px;
  const int2 rg = ranges[tile];
  int ch;
  bool owner;
  grad_lanes<DIM>(lane, &ch, &owner);
  if (rg.y - rg.x <= 0) return;

  const size_t HW = (size_t)W * H;
  float dl[kSlots][DIM];
  float tau[kSlots], fpy[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int py = y0 + 4 * j;
    const bool in = px < W && py < H;
    const size_t pix = (size_t)py * W + px;
#pragma unroll
    for (int k = 0; k < DIM; ++k) dl[j][k] = in ? dlayer[k * HW + pix] : 0.f;
    tau[j] = in ? 1.f : 0.f;
    fpy[j] = (float)py;
  }
  // The per-entry partial sums are zero outside a flush: they start at zero
  // and are cleared after each one (zero_part), so an entry without
  // contributions leaves them untouched and the accumulation needs no
  // per-entry re-zeroing.
  float part[DIM];
  zero_part<DIM>(part);
  // Only the next batch's ids (and strip masks) are fetched ahead; the
  // records are gathered at the start of their batch, where the other
  // wavefronts of the SIMD cover the latency (with the 16 registers a
  // prefetched record would hold, the kernel fits eight wavefronts per
  // SIMD).  With strip masks, entries without one are not gathered at all.
  int nid = rg.x + lane < rg.y ? ids[rg.x + lane] : -1;
  uint32_t nstr = 0;
  if (STRIPS && rg.x + lane < rg.y) nstr = strips[rg.x + lane];
  for (int b0 = rg.x; b0 < rg.y; b0 += 64) {
    const uint32_t live = live_slots(tau);
    if (live == 0) break;
    RecRegs<F4> cur;
    {
      const bool ok = nid >= 0 && (!STRIPS || nstr != 0u);
      const float4* src = recs + (size_t)max(nid, 0) * F4;
#pragma unroll
      for (int q = 0; q < F4; ++q) cur.v[q] = ok ? src[q] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const int sid = nid;  // the staged entries' Gaussians, read per entry with v_readlane
    const uint32_t sstr = nstr;
    nid = b0 + 64 + lane < rg.y ? ids[b0 + 64 + lane] : -1;
    if (STRIPS) nstr = b0 + 64 + lane < rg.y ? strips[b0 + 64 + lane] : 0u;
    __syncthreads();  // previous batch fully read
    uint32_t smask;
    if (STRIPS) {
      s_rec[lane] = cur.v[0];
      s_rec[64 + lane] = cur.v[1];
#pragma unroll
      for (int q = 2; q < F4; ++q) s_rec[q * 64 + lane] = cur.v[q];
      smask = sstr | (cur.v[1].z <= 170.0f ? 0u : kWideExp);
    } else {
      smask = stage_rec<F4>(s_rec, lane, cur, ty);
    }
    __syncthreads();
    const int cnt = min(64, rg.y - b0);
    for (int jj = 0; jj < cnt; ++jj) {
      const uint32_t ebits = entry_bits(smask, jj);
      const uint32_t slots = ebits & live;
      if (slots == 0) continue;
      const float4 g0 = s_rec[jj];
      const float4 g1 = s_rec[64 + jj];
      const float dx = g0.x - fpx;
      const float ta0 = g0.z * dx * dx;
      const float tc0 = g0.w * dx;
      bool on = false;
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        if (!(slots & (1u << j))) continue;
        const float dy = g0.y - fpy[j];
        const float m = fmaxf(0.0f, fmaf(tc0, dy, fmaf(g1.x * dy, dy, ta0)));
        // exp(-m/2) in the form the entry's cutoff allows, formed before the
        // lanes' tests so that the scalar branch lies outside their divergent
        // region (a replayed strip has a contributing pixel, so there is no
        // all-lanes-skip to lose)
        float e = gauss_half_fast(m);
        if (__builtin_expect((ebits & kWideExp) != 0u, 0)) e = gauss_half(m);
        if (tau[j] < 0.0001f || m > g1.z) continue;  // stopped pixel, or beyond the cutoff
        const float ap = fminf(0.99f, g1.y * e);
        if (ap < 0.002f) continue;
        const float ta = tau[j] * ap;
#pragma unroll
        for (int k = 0; k < DIM; ++k) part[k] = fmaf(dl[j][k], ta, part[k]);
        tau[j] = tau[j] * (1.f - ap);
        if (!STRIPS) on = true;
      }
      if (STRIPS || __any(on)) {
        flush_grad<DIM>(part, dfeat, __builtin_amdgcn_readlane(sid, jj), lane, ch, owner);
        zero_part<DIM>(part);
      }
    }
  }
}

template <int DIM, bool STRIPS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(bwd_waves<DIM, STRIPS>())))
void k_draw_bwd_replay(
    int W, int H, int gx, int tiles,
    const int2* __restrict__ ranges, const int32_t* __restrict__ ids, const float4* __restrict__ recs,
    const uint8_t* __restrict__ strips, const float* __restrict__ dlayer, float* __restrict__ dfeat) {
  __shared__ float4 s_rec[64 * RecStride<DIM>::f4];
  draw_bwd_replay_tile<DIM, STRIPS>(W, H, gx, xcd_tile(blockIdx.x, tiles), ranges, ids, recs, strips, dlayer, dfeat,
                                    s_rec);
}

// The replay backward of `views` images in one launch, laid out as in
// k_draw_fwd_views; view v reads dlayer + v*DIM*H*W (and its strip masks at
// strips + v*id_stride) and adds into its own gradient rows dfeat + v*n*DIM
// (the SH chain behind them depends on the view).
template <int DIM, bool STRIPS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(bwd_waves<DIM, STRIPS>())))
void k_draw_bwd_replay_views(
    int W, int H, int gx, int tiles, int views,
    const int2* __restrict__ ranges, const int32_t* __restrict__ ids, int64_t id_stride,
    const float4* __restrict__ recs, int64_t rec_stride, const uint8_t* __restrict__ strips,
    const float* __restrict__ dlayer, float* __restrict__ dfeat, int64_t n) {
  __shared__ float4 s_rec[64 * RecStride<DIM>::f4];
  const int t = xcd_tile(blockIdx.x, tiles * views);
  const int v = t / tiles, tile = t - v * tiles;
  const size_t HW = (size_t)W * H;
  draw_bwd_replay_tile<DIM, STRIPS>(W, H, gx, tile, ranges + (size_t)v * tiles, ids + v * id_stride,
                                    recs + v * rec_stride, STRIPS ? strips + v * id_stride : nullptr,
                                    dlayer + (size_t)v * DIM * HW, dfeat + (size_t)v * n * DIM, s_rec);
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

// RGB of the four pixels p4 .. p4+3 (those below hw): rgb_c = sum_k P_kc W_k,
// summed in channel order k = 0 .. K-1 as the oracle does.  VEC: hw is a
// multiple of 4 and both planes are 16-byte aligned, so the group is one
// float4 per plane; else pixel by pixel.
template <int K, bool VEC>
__device__ __forceinline__ void mix_group(int64_t hw, const float* __restrict__ Wp, const float (&P)[K][3],
                                          float* __restrict__ out, int64_t p4) {
  if constexpr (VEC) {
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

__device__ __forceinline__ bool mix_vec_ok(int64_t hw, const void* W, const void* rgb) {
  return (hw & 3) == 0 && (((uintptr_t)W | (uintptr_t)rgb) & 15) == 0;
}

// Batched mix: image blockIdx.y, with its own size, planes and palette; the
// float4 test is per image (uniform over the block).
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
  if (mix_vec_ok(hw, Wp, out))
    mix_group<K, true>(hw, Wp, P, out, p4);
  else
    mix_group<K, false>(hw, Wp, P, out, p4);
}

// Single-image mix with its arguments by value (no device pointer arrays).
template <int K, bool VEC>
__global__ __launch_bounds__(256) void k_mix1(int64_t hw, const float* __restrict__ Wp,
                                              const float* __restrict__ pal, float* __restrict__ out) {
  float P[K][3];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) P[k][c] = pal[k * 3 + c];
  const int64_t p4 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (p4 >= hw) return;
  mix_group<K, VEC>(hw, Wp, P, out, p4);
}

// dL/dW_k = P_k . dL/drgb per pixel, and dL/dP_kc = sum_p W_k dL/drgb_c.
// VEC: hw is a multiple of 4 and the planes are 16-byte aligned; each thread
// then takes two float4 groups of pixels (all 2 (K+3) loads issued before the
// first use).  Else one pixel per thread and step of a grid-stride loop.
template <int K, bool VEC>
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
  if constexpr (VEC) {
    constexpr int G = 2;
    const int64_t n4 = hw >> 2;
    const int64_t q0 = (int64_t)blockIdx.x * (256 * G) + threadIdx.x;
    const float4* d4 = reinterpret_cast<const float4*>(drgb);
    const float4* w4 = reinterpret_cast<const float4*>(Wp);
    float4* o4 = reinterpret_cast<float4*>(dW);
    float4 d[G][3], w[G][K];
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int64_t q = q0 + 256 * i;
      const bool ok = q < n4;
#pragma unroll
      for (int c = 0; c < 3; ++c) d[i][c] = ok ? d4[c * n4 + q] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int k = 0; k < K; ++k) w[i][k] = ok ? w4[k * n4 + q] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int64_t q = q0 + 256 * i;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        float4 r;
        r.x = (P[k][0] * d[i][0].x + P[k][1] * d[i][1].x) + P[k][2] * d[i][2].x;
        r.y = (P[k][0] * d[i][0].y + P[k][1] * d[i][1].y) + P[k][2] * d[i][2].y;
        r.z = (P[k][0] * d[i][0].z + P[k][1] * d[i][1].z) + P[k][2] * d[i][2].z;
        r.w = (P[k][0] * d[i][0].w + P[k][1] * d[i][1].w) + P[k][2] * d[i][2].w;
        if (q < n4) o4[k * n4 + q] = r;
#pragma unroll
        for (int c = 0; c < 3; ++c)
          gp[k][c] = gp[k][c] + ((w[i][k].x * d[i][c].x + w[i][k].y * d[i][c].y) +
                                 (w[i][k].z * d[i][c].z + w[i][k].w * d[i][c].w));
      }
    }
  } else {
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

static int draw_grid(const char* who, int width, int height, int dim, int* gx, int* tiles) {
  if (width <= 0 || height <= 0 || dim < 1 || dim > RGP_MAX_DIM)
    return set_error(RGP_ERR_ARG, "%s: bad size or dim", who);
  *gx = (width + kTile - 1) / kTile;
  *tiles = *gx * ((height + kTile - 1) / kTile);
  return RGP_OK;
}

extern "C" int rgp_draw_forward(int width, int height, int dim, const int32_t* ranges,
                                const int32_t* ids, const float* records, float empty_tau,
                                float* layer, int32_t* contrib, float* final_tau, uint8_t* strips,
                                hipStream_t s) {
  int gx, tiles;
  if (int rc = draw_grid("rgp_draw_forward", width, height, dim, &gx, &tiles)) return rc;
  if (!contrib && !strips) {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd<DD, false, false, false>), dim3(tiles), dim3(64), 0, s,
                                           width, height, gx, tiles, (const int2*)ranges, ids,
                                           (const float4*)records, empty_tau, layer, contrib,
                                           final_tau, (unsigned long long*)nullptr, (uint8_t*)nullptr));
  } else if (!contrib) {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd<DD, false, true, false>), dim3(tiles), dim3(64), 0, s,
                                           width, height, gx, tiles, (const int2*)ranges, ids,
                                           (const float4*)records, empty_tau, layer, contrib,
                                           final_tau, (unsigned long long*)nullptr, strips));
  } else if (strips) {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd<DD, false, true>), dim3(tiles), dim3(64), 0, s,
                                           width, height, gx, tiles, (const int2*)ranges, ids,
                                           (const float4*)records, empty_tau, layer, contrib,
                                           final_tau, (unsigned long long*)nullptr, strips));
  } else {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd<DD, false, false>), dim3(tiles), dim3(64), 0, s,
                                           width, height, gx, tiles, (const int2*)ranges, ids,
                                           (const float4*)records, empty_tau, layer, contrib,
                                           final_tau, (unsigned long long*)nullptr, (uint8_t*)nullptr));
  }
  return check_launch("rgp_draw_forward");
}

static int views_grid(const char* who, int views, int width, int height, int dim, int64_t id_stride,
                      int* gx, int* tiles) {
  if (int rc = draw_grid(who, width, height, dim, gx, tiles)) return rc;
  if (views < 1 || views > RGP_MAX_VIEWS || (int64_t)views * *tiles > INT32_MAX || id_stride < 1)
    return set_error(RGP_ERR_ARG, "%s: bad views (%d) or id stride", who, views);
  return RGP_OK;
}

extern "C" int rgp_draw_forward_views(int views, int width, int height, int dim, const int32_t* ranges,
                                      const int32_t* ids, int64_t id_stride, const float* records,
                                      int64_t n, float empty_tau, float* layer, int32_t* contrib,
                                      float* final_tau, uint8_t* strips, hipStream_t s) {
  int gx, tiles;
  if (int rc = views_grid("rgp_draw_forward_views", views, width, height, dim, id_stride, &gx, &tiles)) return rc;
  if (!contrib && !strips) {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd_views<DD, false, false>), dim3(tiles * views), dim3(64), 0, s,
                                           width, height, gx, tiles, views, (const int2*)ranges, ids, id_stride,
                                           (const float4*)records, n * RecStride<DD>::f4, empty_tau, layer,
                                           contrib, final_tau, (uint8_t*)nullptr));
  } else if (!contrib) {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd_views<DD, true, false>), dim3(tiles * views), dim3(64), 0, s,
                                           width, height, gx, tiles, views, (const int2*)ranges, ids, id_stride,
                                           (const float4*)records, n * RecStride<DD>::f4, empty_tau, layer,
                                           contrib, final_tau, strips));
  } else if (strips) {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd_views<DD, true>), dim3(tiles * views), dim3(64), 0, s,
                                           width, height, gx, tiles, views, (const int2*)ranges, ids, id_stride,
                                           (const float4*)records, n * RecStride<DD>::f4, empty_tau, layer,
                                           contrib, final_tau, strips));
  } else {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd_views<DD, false>), dim3(tiles * views), dim3(64), 0, s,
                                           width, height, gx, tiles, views, (const int2*)ranges, ids, id_stride,
                                           (const float4*)records, n * RecStride<DD>::f4, empty_tau, layer,
                                           contrib, final_tau, (uint8_t*)nullptr));
  }
  return check_launch("rgp_draw_forward_views");
}

extern "C" int rgp_draw_backward_replay_views(int views, int width, int height, int dim, const int32_t* ranges,
                                              const int32_t* ids, int64_t id_stride, const float* records,
                                              int64_t n, const uint8_t* strips, const float* dlayer,
                                              float* dfeat, hipStream_t s) {
  int gx, tiles;
  if (int rc = views_grid("rgp_draw_backward_replay_views", views, width, height, dim, id_stride, &gx, &tiles))
    return rc;
  if (strips) {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_bwd_replay_views<DD, true>), dim3(tiles * views), dim3(64), 0,
                                           s, width, height, gx, tiles, views, (const int2*)ranges, ids, id_stride,
                                           (const float4*)records, n * RecStride<DD>::f4, strips, dlayer, dfeat,
                                           n));
  } else {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_bwd_replay_views<DD, false>), dim3(tiles * views), dim3(64), 0,
                                           s, width, height, gx, tiles, views, (const int2*)ranges, ids, id_stride,
                                           (const float4*)records, n * RecStride<DD>::f4,
                                           (const uint8_t*)nullptr, dlayer, dfeat, n));
  }
  return check_launch("rgp_draw_backward_replay_views");
}

extern "C" int rgp_draw_forward_counts(int width, int height, int dim, const int32_t* ranges,
                                       const int32_t* ids, const float* records, float empty_tau,
                                       float* layer, int32_t* contrib, float* final_tau,
                                       int64_t* counts, hipStream_t s) {
  int gx, tiles;
  if (int rc = draw_grid("rgp_draw_forward_counts", width, height, dim, &gx, &tiles)) return rc;
  RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_fwd<DD, true, false>), dim3(tiles), dim3(64), 0, s,
                                         width, height, gx, tiles, (const int2*)ranges, ids,
                                         (const float4*)records, empty_tau, layer, contrib,
                                         final_tau, (unsigned long long*)counts, (uint8_t*)nullptr));
  return check_launch("rgp_draw_forward_counts");
}

extern "C" int rgp_draw_backward(int width, int height, int dim, const int32_t* ranges,
                                 const int32_t* ids, const float* records, const int32_t* contrib,
                                 const float* final_tau, const float* dlayer, float* dfeat,
                                 hipStream_t s) {
  int gx, tiles;
  if (int rc = draw_grid("rgp_draw_backward", width, height, dim, &gx, &tiles)) return rc;
  RGP_DIM_SWITCH(dim, hipLaunchKernelGGL(k_draw_bwd<DD>, dim3(tiles), dim3(64), 0, s, width,
                                         height, gx, tiles, (const int2*)ranges, ids,
                                         (const float4*)records, contrib, final_tau, dlayer,
                                         dfeat));
  return check_launch("rgp_draw_backward");
}

extern "C" int rgp_draw_color_backward(int width, int height, const int32_t* ranges, const int32_t* ids,
                                       const float* records, const int32_t* contrib, const float* final_tau,
                                       const float* dimage, const float* dfinal_tau, float* dus, float* dcinv,
                                       float* dalphas, float* dcolors, hipStream_t s) {
  int gx, tiles;
  if (int rc = draw_grid("rgp_draw_color_backward", width, height, 3, &gx, &tiles)) return rc;
  if (!dus || !dcinv || !dalphas || !dcolors)
    return set_error(RGP_ERR_ARG, "rgp_draw_color_backward: null gradient buffer");
  hipLaunchKernelGGL(k_draw_color_bwd, dim3(tiles), dim3(64), 0, s, width, height, gx, tiles, (const int2*)ranges,
                     ids, (const float4*)records, contrib, final_tau, dimage, dfinal_tau, dus, dcinv, dalphas,
                     dcolors);
  return check_launch("rgp_draw_color_backward");
}

extern "C" int rgp_draw_backward_replay(int width, int height, int dim, const int32_t* ranges,
                                        const int32_t* ids, const float* records, const uint8_t* strips,
                                        const float* dlayer, float* dfeat, hipStream_t s) {
  int gx, tiles;
  if (int rc = draw_grid("rgp_draw_backward_replay", width, height, dim, &gx, &tiles)) return rc;
  if (strips) {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_bwd_replay<DD, true>), dim3(tiles), dim3(64), 0, s,
                                           width, height, gx, tiles, (const int2*)ranges, ids,
                                           (const float4*)records, strips, dlayer, dfeat));
  } else {
    RGP_DIM_SWITCH(dim, hipLaunchKernelGGL((k_draw_bwd_replay<DD, false>), dim3(tiles), dim3(64), 0, s,
                                           width, height, gx, tiles, (const int2*)ranges, ids,
                                           (const float4*)records, (const uint8_t*)nullptr, dlayer, dfeat));
  }
  return check_launch("rgp_draw_backward_replay");
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

extern "C" int rgp_mix(int K, int64_t hw, const float* W, const float* palette, float* rgb,
                       hipStream_t s) {
  if (K < 1 || K > RGP_MAX_DIM || hw < 0) return set_error(RGP_ERR_ARG, "rgp_mix: bad K or size");
  if (hw == 0) return RGP_OK;
  const int64_t blocks = (hw + 1023) / 1024;
  if (blocks > 0x7fffffff) return set_error(RGP_ERR_ARG, "rgp_mix: image too large");
  if (((hw & 3) == 0) && ((((uintptr_t)W) | ((uintptr_t)rgb)) & 15) == 0) {
    RGP_DIM_SWITCH(K, hipLaunchKernelGGL((k_mix1<DD, true>), dim3((unsigned)blocks), dim3(256), 0, s, hw, W,
                                         palette, rgb));
  } else {
    RGP_DIM_SWITCH(K, hipLaunchKernelGGL((k_mix1<DD, false>), dim3((unsigned)blocks), dim3(256), 0, s, hw, W,
                                         palette, rgb));
  }
  return check_launch("rgp_mix");
}

extern "C" int rgp_mix_backward(int K, int64_t hw, const float* W, const float* palette,
                                const float* drgb, float* dW, float* dpalette, hipStream_t s) {
  if (K < 1 || K > RGP_MAX_DIM || hw < 0)
    return set_error(RGP_ERR_ARG, "rgp_mix_backward: bad K or size");
  if (hw == 0) return RGP_OK;
  const bool vec = hw % 4 == 0 && (((uintptr_t)W | (uintptr_t)drgb | (uintptr_t)dW) & 15) == 0;
  if (vec) {
    const int64_t blocks = (hw / 4 + 511) / 512;
    if (blocks > 0x7fffffff) return set_error(RGP_ERR_ARG, "rgp_mix_backward: image too large");
    RGP_DIM_SWITCH(K, hipLaunchKernelGGL((k_mix_bwd<DD, true>), dim3((unsigned)blocks), dim3(256), 0, s, hw,
                                         W, palette, drgb, dW, dpalette));
  } else {
    int64_t blocks = (hw + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    RGP_DIM_SWITCH(K, hipLaunchKernelGGL((k_mix_bwd<DD, false>), dim3((unsigned)blocks), dim3(256), 0, s, hw,
                                         W, palette, drgb, dW, dpalette));
  }
  return check_launch("rgp_mix_backward");
}
