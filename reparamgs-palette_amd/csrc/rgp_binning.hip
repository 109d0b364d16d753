// Tile binning for gfx950: the pair count M, the tile lists and per-tile
// ranges.  Replaces thrust::inclusive_scan, createKeys, thrust::sort_by_key
// and getRanges of gsplatcu/gausplat.cu:199-226 / kernel.cu:96-187,
// producing the same order: tile ascending, quantised depth ascending,
// Gaussian id ascending (the input order a stable sort keeps).
//
// Nothing here synchronises the host.  Sizes that the reference reads back
// (the pair count M, gausplat.cu:201) stay on the device: every kernel is
// launched for a caller-given pair capacity and reads M from the workspace
// header; if M exceeds the capacity the overflow flag is raised and all
// downstream kernels see zero pairs.
//
// Two paths.  Grids up to 3840x2160 take the tile-bucket path (two-level
// bucketing into the tiles' lists, then a sort inside each tile; below).
// Larger grids take an LSD radix sort of the (tile, depth) keys with 8-bit
// digits, one pass per key byte; passes whose digit is the same for every key
// (depth bytes above the largest quantised depth, tile bytes above the tile
// count) are skipped, decided on the device.  Each pass is histogram ->
// per-digit scan over blocks -> scatter, where the scatter ranks keys stably
// inside the block with wavefront ballots, reorders them in LDS and writes
// coalesced runs.
#include "rgp_internal.h"

namespace rgp {

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// Exclusive scan over the block (blockDim.x <= 1024, multiple of 64).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total,
                                                    uint32_t* lds /*16 words*/) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (int j = 0; j < nw; ++j) {
    const uint32_t x = lds[j];
    if (j < w) before += x;
    all += x;
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* dst, uint32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

// ---------------------------------------------------------------------------
// scan
// ---------------------------------------------------------------------------
// Sums the counts of one chunk and, when dqpart is given, takes the maximum
// of its per-64-Gaussian quantised-depth maxima (written by the rect kernels
// without atomics; ndq of them).
__global__ __launch_bounds__(256) void k_scan_reduce(int n, const uint32_t* __restrict__ counts,
                                                     int ndq, const uint32_t* __restrict__ dqw,
                                                     uint32_t* __restrict__ partials,
                                                     uint32_t* __restrict__ dqpart) {
  static_assert(kScanChunk == 64 * 64, "one wavefront reads the chunk's 64 group maxima");
  __shared__ uint32_t lds[16];
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  uint32_t s = 0;
  if (base + kScanChunk <= n) {  // a full chunk: coalesced 16-byte loads (integer sum, any order)
    const uint4* c4 = reinterpret_cast<const uint4*>(counts + base);
#pragma unroll
    for (int j = 0; j < kScanChunk / 1024; ++j) {
      const uint4 x = c4[j * 256 + threadIdx.x];
      s += (x.x + x.y) + (x.z + x.w);
    }
  } else {
#pragma unroll
    for (int j = 0; j < kScanChunk / 256; ++j) {
      const int64_t i = base + j * 256 + threadIdx.x;
      if (i < n) s += counts[i];
    }
  }
  if (dqpart && threadIdx.x < 64) {
    const int64_t g = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const uint32_t dq = wave_max_u32(g < ndq ? dqw[g] : 0u);
    if (threadIdx.x == 0) dqpart[blockIdx.x] = dq;
  }
  uint32_t tot;
  block_excl_scan(s, &tot, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Scans the block sums, publishes M, the overflow flag and the sort plan.
__global__ __launch_bounds__(1024) void k_scan_top(int nscan, uint32_t* __restrict__ partials,
                                                   const uint32_t* __restrict__ dqpart,
                                                   uint32_t* __restrict__ hdr, int64_t cap,
                                                   int tile_bytes) {
  __shared__ uint32_t lds[16];
  __shared__ uint32_t s_dq[16];
  uint32_t carry = 0, dqm = 0;
  for (int base = 0; base < nscan; base += blockDim.x) {
    const int i = base + threadIdx.x;
    const uint32_t v = i < nscan ? partials[i] : 0u;
    if (dqpart && i < nscan) dqm = max(dqm, dqpart[i]);
    uint32_t tot;
    const uint32_t ex = block_excl_scan(v, &tot, lds);
    if (i < nscan) partials[i] = carry + ex;
    carry += tot;
  }
  dqm = wave_max_u32(dqm);
  if ((threadIdx.x & 63) == 0) s_dq[threadIdx.x >> 6] = dqm;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t dq = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) dq = max(dq, s_dq[w]);
    hdr[kHdrDqMax] = dq;
    const uint32_t M = carry;
    const bool ovf = (int64_t)M > cap;
    const uint32_t meff = ovf ? 0u : M;
    hdr[kHdrM] = M;
    hdr[kHdrOverflow] = ovf ? 1u : 0u;
    hdr[kHdrMEff] = meff;
    int depth_bytes = 0;
    while (depth_bytes < 4 && (dq >> (8 * depth_bytes)) != 0) ++depth_bytes;
    uint32_t cur = 0;
    int last = -1;
    for (int p = 0; p < 8; ++p) {
      bool skip = p < 4 ? (p >= depth_bytes) : (p - 4 >= tile_bytes);
      if (meff <= 1) skip = true;
      hdr[kHdrCur + p] = cur;
      hdr[kHdrSkip + p] = skip ? 1u : 0u;
      if (!skip) { cur ^= 1u; last = p; }
    }
    hdr[kHdrCur + 8] = cur;
    hdr[kHdrLast] = (uint32_t)last;
    hdr[kHdrNblk] = (meff + kSortTile - 1) / kSortTile;
  }
}

// counts -> inclusive offsets, in place.
__global__ __launch_bounds__(256) void k_scan_down(int n, uint32_t* __restrict__ counts,
                                                   const uint32_t* __restrict__ partials) {
  __shared__ uint32_t lds[16];
  constexpr int PER = kScanChunk / 256;
  static_assert(PER % 4 == 0, "a thread's counts are whole 16-byte groups");
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * PER;
  // a full chunk (the workspace arrays are 256-byte aligned): 16-byte loads
  // and stores, without a bounds test per element
  const bool full = (int64_t)(blockIdx.x + 1) * kScanChunk <= n;
  uint4* c4 = reinterpret_cast<uint4*>(counts + base);
  uint32_t v[PER];
  uint32_t s = 0;
  if (full) {
#pragma unroll
    for (int q = 0; q < PER / 4; ++q) {
      const uint4 x = c4[q];
      v[4 * q] = x.x;
      v[4 * q + 1] = x.y;
      v[4 * q + 2] = x.z;
      v[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < PER; ++j) v[j] = (base + j < n) ? counts[base + j] : 0u;
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) s += v[j];
  uint32_t tot;
  uint32_t run = block_excl_scan(s, &tot, lds) + partials[blockIdx.x];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    run += v[j];
    v[j] = run;
  }
  if (full) {
#pragma unroll
    for (int q = 0; q < PER / 4; ++q) c4[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  } else {
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (base + j < n) counts[base + j] = v[j];
  }
}

// ---------------------------------------------------------------------------
// createKeys (kernel.cu:96-124), load balanced: a block expands the pairs of
// its 256 Gaussians cooperatively, so one large splat does not serialise a
// thread.  Pair p of Gaussian i (row-major over its rect) goes to
// offset[i-1] + p, as in the reference.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dup(int n, const uint32_t* __restrict__ offs,
                                             const uint2* __restrict__ rects,
                                             const float* __restrict__ depths, int gx,
                                             const uint32_t* __restrict__ hdr,
                                             uint64_t* __restrict__ keys,
                                             int32_t* __restrict__ vals) {
  if (hdr[kHdrMEff] == 0) return;
  __shared__ uint32_t s_incl[256];
  __shared__ uint2 s_rect[256];
  __shared__ uint32_t s_dq[256];
  const int t = threadIdx.x;
  const int i0 = blockIdx.x * 256;
  const uint32_t base = i0 == 0 ? 0u : offs[i0 - 1];
  const int i = i0 + t;
  const int cnt_blk = min(256, n - i0);
  uint32_t incl = 0;
  if (i < n) {
    incl = offs[i] - base;
    const uint32_t excl = (i == 0 ? 0u : offs[i - 1]) - base;
    if (incl > excl) {
      s_rect[t] = rects[i];
      s_dq[t] = sat_u32(depths[i] * 1000.f);
    }
  }
  s_incl[t] = incl;
  __syncthreads();
  const uint32_t total = s_incl[cnt_blk - 1];
  for (uint32_t p = t; p < total; p += 256) {
    // smallest g with s_incl[g] > p
    int lo = 0, hi = cnt_blk - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_incl[mid] > p) hi = mid; else lo = mid + 1;
    }
    const uint32_t li = p - (lo == 0 ? 0u : s_incl[lo - 1]);
    const uint2 r = s_rect[lo];
    const uint32_t x0 = r.x & 0xffffu, y0 = r.x >> 16, x1 = r.y & 0xffffu;
    const uint32_t w = x1 - x0;
    const uint32_t yy = li / w, xx = li - yy * w;
    const uint32_t tile = (y0 + yy) * (uint32_t)gx + (x0 + xx);
    const uint32_t gp = base + p;
    keys[gp] = ((uint64_t)tile << 32) | (uint64_t)s_dq[lo];
    vals[gp] = i0 + lo;
  }
}

// ---------------------------------------------------------------------------
// radix sort passes
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t digit_of(uint64_t k, int p) {
  return (uint32_t)(k >> (8 * p)) & 0xffu;
}

__global__ __launch_bounds__(256) void k_sort_hist(int p, const uint32_t* __restrict__ hdr,
                                                   const uint64_t* __restrict__ k0,
                                                   const uint64_t* __restrict__ k1,
                                                   uint32_t* __restrict__ hist, int nblk_cap) {
  if (hdr[kHdrSkip + p]) return;
  const uint32_t M = hdr[kHdrMEff];
  const int b = blockIdx.x;
  if ((uint32_t)b >= hdr[kHdrNblk]) return;
  const uint64_t* src = hdr[kHdrCur + p] ? k1 : k0;
  __shared__ uint32_t h[4][256];
  const int t = threadIdx.x, w = t >> 6;
#pragma unroll
  for (int q = 0; q < 4; ++q) h[q][t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)b * kSortTile;
#pragma unroll 4
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = base + j * 256 + t;
    if (e < M) atomicAdd(&h[w][digit_of(src[e], p)], 1u);
  }
  __syncthreads();
  hist[(size_t)t * nblk_cap + b] = h[0][t] + h[1][t] + h[2][t] + h[3][t];
}

// One block per digit: exclusive scan of its counts over the sort blocks.
__global__ __launch_bounds__(256) void k_sort_scan(int p, const uint32_t* __restrict__ hdr,
                                                   uint32_t* __restrict__ hist,
                                                   uint32_t* __restrict__ totals, int nblk_cap) {
  if (hdr[kHdrSkip + p]) return;
  __shared__ uint32_t lds[16];
  const int nblk = (int)hdr[kHdrNblk];
  uint32_t* row = hist + (size_t)blockIdx.x * nblk_cap;
  uint32_t carry = 0;
  for (int base = 0; base < nblk; base += 256) {
    const int i = base + threadIdx.x;
    const uint32_t v = i < nblk ? row[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan(v, &tot, lds);
    if (i < nblk) row[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ __launch_bounds__(256) void k_sort_scatter(int p, const uint32_t* __restrict__ hdr,
                                                      uint64_t* __restrict__ k0,
                                                      uint64_t* __restrict__ k1,
                                                      int32_t* __restrict__ v0,
                                                      int32_t* __restrict__ v1,
                                                      int32_t* __restrict__ ids_out,
                                                      const uint32_t* __restrict__ hist,
                                                      const uint32_t* __restrict__ totals,
                                                      int nblk_cap) {
  if (hdr[kHdrSkip + p]) return;
  const uint32_t M = hdr[kHdrMEff];
  const int b = blockIdx.x;
  if ((uint32_t)b >= hdr[kHdrNblk]) return;
  const bool from1 = hdr[kHdrCur + p] != 0;
  const uint64_t* sk = from1 ? k1 : k0;
  const int32_t* sv = from1 ? v1 : v0;
  uint64_t* dk = from1 ? k0 : k1;
  int32_t* dv = ((int)hdr[kHdrLast] == p) ? ids_out : (from1 ? v0 : v1);

  __shared__ uint64_t s_key[kSortTile];
  __shared__ int32_t s_val[kSortTile];
  __shared__ uint32_t s_wcnt[4][256];
  __shared__ uint32_t s_run[256];
  __shared__ uint32_t s_start[256];
  __shared__ uint32_t s_gbase[256];
  __shared__ uint32_t lds[16];

  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // global start of every digit bucket for this block
  {
    uint32_t tot;
    const uint32_t ex = block_excl_scan(totals[t], &tot, lds);
    s_gbase[t] = ex + hist[(size_t)t * nblk_cap + b];
  }
  s_run[t] = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) s_wcnt[q][t] = 0;
  __syncthreads();

  const int64_t base = (int64_t)b * kSortTile;
  uint64_t key[kSortItems];
  int32_t val[kSortItems];
  uint32_t off[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = base + j * 256 + t;
    const bool valid = e < M;
    key[j] = valid ? sk[e] : 0ull;
    val[j] = valid ? sv[e] : 0;
    const uint32_t d = digit_of(key[j], p);
    // peers: lanes of this wavefront holding the same digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool on = (d >> bit) & 1u;
      const uint64_t bb = __ballot(on);
      peers &= on ? bb : ~bb;
    }
    const uint32_t rank = __popcll(peers & lt_mask);
    if (valid && rank == 0) s_wcnt[w][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t o = s_run[d] + rank;
      for (int q = 0; q < w; ++q) o += s_wcnt[q][d];
      off[j] = o;
    }
    __syncthreads();
    s_run[t] += s_wcnt[0][t] + s_wcnt[1][t] + s_wcnt[2][t] + s_wcnt[3][t];
#pragma unroll
    for (int q = 0; q < 4; ++q) s_wcnt[q][t] = 0;
    __syncthreads();
  }
  {
    uint32_t tot;
    s_start[t] = block_excl_scan(s_run[t], &tot, lds);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = base + j * 256 + t;
    if (e < M) {
      const uint32_t pos = s_start[digit_of(key[j], p)] + off[j];
      s_key[pos] = key[j];
      s_val[pos] = val[j];
    }
  }
  __syncthreads();
  const int cnt = (int)min((int64_t)kSortTile, (int64_t)M - base);
  for (int e = t; e < cnt; e += 256) {
    const uint64_t k = s_key[e];
    const uint32_t d = digit_of(k, p);
    const uint32_t g = s_gbase[d] + (uint32_t)e - s_start[d];
    dk[g] = k;
    dv[g] = s_val[e];
  }
}

// Final buffers: ids when no pass ran, keys on request.
__global__ __launch_bounds__(256) void k_sort_finish(const uint32_t* __restrict__ hdr,
                                                     const uint64_t* __restrict__ k0,
                                                     const uint64_t* __restrict__ k1,
                                                     const int32_t* __restrict__ v0,
                                                     int32_t* __restrict__ ids_out,
                                                     uint64_t* __restrict__ keys_out) {
  const uint32_t M = hdr[kHdrMEff];
  const bool copy_ids = (int)hdr[kHdrLast] < 0;
  if (!copy_ids && !keys_out) return;
  const bool from1 = hdr[kHdrCur + 8] != 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < M;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (copy_ids) ids_out[e] = v0[e];
    if (keys_out) keys_out[e] = from1 ? k1[e] : k0[e];
  }
}

// getRanges (kernel.cu:165-187) by binary search per tile; (0,0) for empty
// tiles and, as in the reference, for every tile when M == 1.
__global__ __launch_bounds__(256) void k_ranges(int tiles, const uint32_t* __restrict__ hdr,
                                                const uint64_t* __restrict__ k0,
                                                const uint64_t* __restrict__ k1,
                                                int2* __restrict__ ranges) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tiles) return;
  const uint32_t M = hdr[kHdrMEff];
  if (M <= 1) { ranges[t] = make_int2(0, 0); return; }
  const uint64_t* keys = hdr[kHdrCur + 8] ? k1 : k0;
  auto lower = [&](uint32_t tile) {
    uint32_t lo = 0, hi = M;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((uint32_t)(keys[mid] >> 32) < tile) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  const uint32_t a = lower((uint32_t)t), e = lower((uint32_t)t + 1);
  ranges[t] = e > a ? make_int2((int)a, (int)e) : make_int2(0, 0);
}

// ---------------------------------------------------------------------------
// Tile-bucket path.  The order the reference's stable (tile, depth) sort
// produces inside one tile is ascending (uint32(depth*1000), Gaussian id):
// createKeys writes the pairs of Gaussian i before those of i + 1
// (kernel.cu:96-124) and the radix sort keeps that order among equal keys.
// So the tile lists may be filled in any order and each sorted by the key
// (depth_q << 32 | id), whose values are all distinct:
//   1. k_st_count: blocks of Gaussians count their (super-tile, Gaussian)
//      pairs in LDS and reserve slots of each super-tile's coarse list;
//   2. k_st_scatter: every block forms M (from the tile counts) and the
//      coarse lists' offsets, block 0 publishes them with the overflow flag,
//      and each coarse entry carries the Gaussian's id, tile rect and
//      quantised depth (written by the preprocess in Gaussian order);
//   3. k_st_fine_count: one block per super-tile counts its tiles' pairs and
//      adds them to the per-tile-row sums;
//   4. k_st_fine_scatter: one block per super-tile forms its tiles' list
//      offsets from the row sums and the counts, and writes the complete
//      lists of sort keys (every line of a list written by one block);
//   5. k_tb_sort (lists <= 1024, registers), k_tb_sort_long (longer ones).
// Super-tiles are 4 x 4 tiles up to 2560x1440 and 8 x 8 above, a few
// hundred of them either way.
// ---------------------------------------------------------------------------

// Super-tile rect [sx0, sx1) x [sy0, sy1) of a packed tile rect (x1, y1
// exclusive, non-empty), super-tiles of 2^sts x 2^sts tiles.
struct StRect {
  uint32_t x0, y0, x1, y1;
};
__device__ __forceinline__ StRect st_rect(uint2 r, uint32_t sts) {
  const uint32_t x0 = r.x & 0xffffu, y0 = r.x >> 16, x1 = r.y & 0xffffu, y1 = r.y >> 16;
  return StRect{x0 >> sts, y0 >> sts, ((x1 - 1) >> sts) + 1, ((y1 - 1) >> sts) + 1};
}

// Calls fn(s, i, rect, w) for every super-tile s that the rect of Gaussian i
// in [g0, g1) reaches (w = extra[i], or 0 without extra), and returns the
// thread's sum of the tile counts of its Gaussians: a thread walks its own
// Gaussian's super-tiles when there are at most kTbBig of them; more are
// walked by the whole wavefront, 64 a step.  The count, rect and extra word
// of a thread's next Gaussian are loaded before its current one is walked
// (the rect and extra unconditionally: words of culled Gaussians are read
// and ignored), so each step waits on loads issued a step earlier.
template <class F>
__device__ __forceinline__ uint32_t st_for_pairs(int g0, int g1, const uint32_t* __restrict__ counts,
                                                 const uint2* __restrict__ rects,
                                                 const uint32_t* __restrict__ extra, uint32_t sts, uint32_t sgx,
                                                 F fn) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t fine = 0;
  uint32_t nc = 0, nw = 0;
  uint2 nr = make_uint2(0u, 0u);
  auto fetch = [&](int i) {
    nc = 0;
    if (i < g1) {
      nc = counts[i];
      nr = rects[i];
      if (extra) nw = extra[i];
    }
  };
  fetch(g0 + (int)threadIdx.x);
  for (int b0 = g0; b0 < g1; b0 += (int)blockDim.x) {
    const int i = b0 + (int)threadIdx.x;
    const uint32_t c = nc, w = nw;
    const uint2 r = nr;
    fetch(i + (int)blockDim.x);
    fine += c;
    StRect q = c ? st_rect(r, sts) : StRect{0, 0, 0, 0};
    const uint32_t sw = q.x1 - q.x0, sc = sw * (q.y1 - q.y0);
    if (sc != 0 && sc <= (uint32_t)kTbBig) {
      uint32_t sx = q.x0, row = q.y0 * sgx;
      for (uint32_t k = 0; k < sc; ++k) {
        fn(row + sx, i, r, w);
        if (++sx == q.x1) {
          sx = q.x0;
          row += sgx;
        }
      }
    }
    uint64_t big = __ballot(sc > (uint32_t)kTbBig);
    while (big) {
      const int src = __ffsll((unsigned long long)big) - 1;
      big &= big - 1;
      const uint32_t bc = __shfl(sc, src), bw = __shfl(sw, src);
      const uint32_t bx0 = __shfl(q.x0, src), by0 = __shfl(q.y0, src);
      const int bi = __shfl(i, src);
      const uint32_t bwd = (uint32_t)__shfl((int)w, src);
      const uint2 br = make_uint2((uint32_t)__shfl((int)r.x, src), (uint32_t)__shfl((int)r.y, src));
      for (uint32_t k = lane; k < bc; k += 64) {
        const uint32_t yy = k / bw, xx = k - yy * bw;
        fn((by0 + yy) * sgx + bx0 + xx, bi, br, bwd);
      }
    }
  }
  return fine;
}

// In-tile sort key of Gaussian g: (uint32(depth * 1000) << 32) | g.
__device__ __forceinline__ uint64_t tb_key(const float* __restrict__ depths, int32_t g) {
  return ((uint64_t)sat_u32(depths[g] * 1000.f) << 32) | (uint32_t)g;
}

// Coarse count: block b counts its Gaussians' (super-tile, Gaussian) pairs per
// super-tile in LDS, reserves that many slots of each super-tile's coarse list
// with one atomic on sttot[s] (cleared with the header), and records the end
// of its slots, relative to the list's start, in boff[b * st + s]; bpart[b] =
// the tile pairs of its Gaussians.
__global__ __launch_bounds__(kTbThreads) void k_st_count(int n, int chunk, int st, int sts, int sgx,
                                                         const uint32_t* __restrict__ counts,
                                                         const uint2* __restrict__ rects,
                                                         uint32_t* __restrict__ sttot,
                                                         uint32_t* __restrict__ boff,
                                                         uint32_t* __restrict__ bpart) {
  extern __shared__ uint32_t s_hist[];  // st
  __shared__ uint32_t lds[16];
  const int b = blockIdx.x;
  for (int q = threadIdx.x; q < st; q += kTbThreads) s_hist[q] = 0;
  __syncthreads();
  const int g0 = b * chunk, g1 = min(n, g0 + chunk);
  const uint32_t fine = st_for_pairs(g0, g1, counts, rects, nullptr, (uint32_t)sts, (uint32_t)sgx,
                                     [&](uint32_t s, int, uint2, uint32_t) { atomicAdd(&s_hist[s], 1u); });
  uint32_t tot;
  block_excl_scan(fine, &tot, lds);  // its barriers also order the histogram's atomics before the reads
  for (int q = threadIdx.x; q < st; q += kTbThreads) {
    const uint32_t h = s_hist[q];
    if (h) boff[(size_t)b * st + q] = atomicAdd(&sttot[q], h) + h;
  }
  if (threadIdx.x == 0) bpart[b] = tot;
}

// Coarse scatter.  Every block first forms M = the sum of bpart (the pair
// count of the reference's scan, gausplat.cu:199-201) and the exclusive scan
// of the super-tile totals in LDS (the same values in every block; block 0
// publishes M, the overflow flag and the offsets, stoff[st] = the coarse
// pairs, so super-tile s's coarse list is [stoff[s], stoff[s + 1])).  Then
// block b writes, for each (super-tile, Gaussian) pair of its Gaussians, the
// Gaussian's id, tile rect and quantised depth into its slots of the
// super-tile's coarse list.  Coarse pairs never outnumber tile pairs, so the
// lists fit in the pair capacity whenever M does.
__global__ __launch_bounds__(kTbThreads) void k_st_scatter(int n, int chunk, int st, int sts, int sgx, int nb,
                                                           const uint32_t* __restrict__ counts,
                                                           const uint2* __restrict__ rects,
                                                           const uint32_t* __restrict__ dqs,
                                                           const uint32_t* __restrict__ sttot,
                                                           const uint32_t* __restrict__ boff,
                                                           const uint32_t* __restrict__ bpart,
                                                           uint32_t* __restrict__ hdr, uint32_t* __restrict__ stoff,
                                                           int64_t cap, uint4* __restrict__ clist) {
  extern __shared__ uint32_t s_cur[];  // st; entries of super-tiles this block has no pair in are not used
  __shared__ uint32_t lds[16];
  const int b = blockIdx.x;
  uint32_t m = 0;
  for (int q = threadIdx.x; q < nb; q += kTbThreads) m += bpart[q];
  for (int q = threadIdx.x; q < st; q += kTbThreads) s_cur[q] = sttot[q];
  uint32_t M;
  block_excl_scan(m, &M, lds);  // its barriers also order the s_cur stores before the reads below
  {
    const int per = (st + kTbThreads - 1) / kTbThreads;
    const int lo = min(st, per * (int)threadIdx.x), hi = min(st, lo + per);
    uint32_t sum = 0;
    for (int q = lo; q < hi; ++q) sum += s_cur[q];
    uint32_t ctot;
    uint32_t run = block_excl_scan(sum, &ctot, lds);
    for (int q = lo; q < hi; ++q) {
      const uint32_t v = s_cur[q];
      s_cur[q] = run;
      run += v;
    }
    __syncthreads();
    const bool ovf = (int64_t)M > cap;
    if (b == 0) {
      for (int q = threadIdx.x; q < st; q += kTbThreads) stoff[q] = s_cur[q];
      if (threadIdx.x == 0) {
        stoff[st] = ctot;
        hdr[kHdrM] = M;
        hdr[kHdrOverflow] = ovf ? 1u : 0u;
        hdr[kHdrMEff] = ovf ? 0u : M;
      }
    }
    if (ovf || M == 0) return;
  }
  for (int q = threadIdx.x; q < st; q += kTbThreads) s_cur[q] += boff[(size_t)b * st + q];
  __syncthreads();
  const int g0 = b * chunk, g1 = min(n, g0 + chunk);
  st_for_pairs(g0, g1, counts, rects, dqs, (uint32_t)sts, (uint32_t)sgx, [&](uint32_t s, int g, uint2 r, uint32_t dq) {
    clist[atomicSub(&s_cur[s], 1u) - 1u] = make_uint4((uint32_t)g, r.x, r.y, dq);
  });
}

// Calls fn(t) for every tile t (0..kst*kst-1, row-major inside the super-tile
// of edge kst = 2^sts whose first tile is (tx0, ty0)) that the packed tile
// rect (rx, ry) covers.
template <class F>
__device__ __forceinline__ void st_tiles(uint32_t rx, uint32_t ry, uint32_t tx0, uint32_t ty0, uint32_t sts,
                                         F fn) {
  const uint32_t kst = 1u << sts;
  const uint32_t x0 = rx & 0xffffu, y0 = rx >> 16, x1 = ry & 0xffffu, y1 = ry >> 16;
  const uint32_t ix0 = max(x0, tx0) - tx0, ix1 = min(x1, tx0 + kst) - tx0;
  const uint32_t iy0 = max(y0, ty0) - ty0, iy1 = min(y1, ty0 + kst) - ty0;
  for (uint32_t y = iy0; y < iy1; ++y)
    for (uint32_t x = ix0; x < ix1; ++x) fn((y << sts) + x);
}

// Coarse entries a thread of the fine count / scatter loads at once (all its
// loads in flight before the first entry is walked).
constexpr int kFineBatch = 4;

// Fine count: one block per super-tile counts the pairs of each of its tiles
// from its coarse list, toff[tile] = the count, and adds each of its tile
// rows' count to rowsum[row] (cleared with the header).
__global__ __launch_bounds__(1024) void k_st_fine_count(int sts, int sgx, int gx, int gy,
                                                        const uint32_t* __restrict__ stoff,
                                                        const uint32_t* __restrict__ hdr,
                                                        const uint4* __restrict__ clist,
                                                        uint32_t* __restrict__ toff,
                                                        uint32_t* __restrict__ rowsum) {
  __shared__ uint32_t cnt[kStMax * kStMax];
  if (hdr[kHdrMEff] == 0) return;
  const int s = blockIdx.x;
  const uint32_t kst = 1u << sts, nt = kst * kst;
  const uint32_t tx0 = (uint32_t)(s % sgx) << sts, ty0 = (uint32_t)(s / sgx) << sts;
  if (threadIdx.x < nt) cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t start = stoff[s], end = stoff[s + 1];
  for (uint32_t e0 = start + threadIdx.x; e0 < end; e0 += kFineBatch * blockDim.x) {
    uint4 c[kFineBatch];
#pragma unroll
    for (int u = 0; u < kFineBatch; ++u) {
      const uint32_t e = e0 + u * blockDim.x;
      c[u] = e < end ? clist[e] : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < kFineBatch; ++u)
      if (e0 + u * blockDim.x < end)
        st_tiles(c[u].y, c[u].z, tx0, ty0, sts, [&](uint32_t t) { atomicAdd(&cnt[t], 1u); });
  }
  __syncthreads();
  if (threadIdx.x < nt) {
    const uint32_t tx = tx0 + (threadIdx.x & (kst - 1)), ty = ty0 + (threadIdx.x >> sts);
    if (tx < (uint32_t)gx && ty < (uint32_t)gy) toff[ty * gx + tx] = cnt[threadIdx.x];
    // tiles past the grid's edge hold no pairs (tile rects are clipped to the grid)
    if ((threadIdx.x & (kst - 1)) == 0 && ty < (uint32_t)gy) {
      uint32_t r = 0;
      for (uint32_t x = 0; x < kst; ++x) r += cnt[threadIdx.x + x];
      if (r) atomicAdd(&rowsum[ty], r);
    }
  }
}

// Fine scatter: one block per super-tile writes the sort keys of its tiles'
// pairs, each tile's list filled from its start by this block alone (LDS
// cursors), so the lines of the key list are written whole.  The keys come
// with the coarse entries, so the in-tile sorts read them in list order
// instead of gathering depths per pair.
__global__ __launch_bounds__(1024) void k_st_fine_scatter(int sts, int sgx, int gx, int gy,
                                                          const uint32_t* __restrict__ stoff,
                                                          const uint32_t* __restrict__ hdr,
                                                          const uint4* __restrict__ clist,
                                                          const uint32_t* __restrict__ toff,
                                                          const uint32_t* __restrict__ rowsum,
                                                          uint32_t* __restrict__ toffx,
                                                          uint64_t* __restrict__ tkeys) {
  __shared__ uint32_t cur[kStMax * kStMax];
  const uint32_t M = hdr[kHdrMEff];
  if (M == 0) return;
  const int s = blockIdx.x;
  const uint32_t kst = 1u << sts, nt = kst * kst;
  const uint32_t tx0 = (uint32_t)(s % sgx) << sts, ty0 = (uint32_t)(s / sgx) << sts;
  // List offsets of the block's tiles, in row-major tile order: wavefront w
  // takes tile row ty0 + w and adds the row sums of the rows above it, the
  // counts of the row's tiles left of the super-tile, and the exclusive scan
  // of the counts inside it.  They go to cur (the fill cursors) and to toffx
  // for the in-tile sorts; toffx[tiles] = M.
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (w < kst) {
    const uint32_t ty = ty0 + w;
    uint32_t c = 0;
    if (ty < (uint32_t)gy) {
      uint32_t part = 0;
      for (uint32_t r0 = 0; r0 < ty; r0 += 64) part += r0 + lane < ty ? rowsum[r0 + lane] : 0u;
      for (uint32_t x0 = 0; x0 < tx0; x0 += 64) part += x0 + lane < tx0 ? toff[ty * gx + x0 + lane] : 0u;
      const uint32_t before = __builtin_amdgcn_readlane(wave_incl_scan(part), 63);
      c = lane < kst && tx0 + lane < (uint32_t)gx ? toff[ty * gx + tx0 + lane] : 0u;
      const uint32_t off = before + wave_incl_scan(c) - c;
      if (lane < kst) {
        cur[w * kst + lane] = off;
        if (tx0 + lane < (uint32_t)gx) toffx[ty * gx + tx0 + lane] = off;
      }
    } else if (lane < kst) {
      cur[w * kst + lane] = 0u;  // rows past the grid hold no pairs
    }
  }
  if (s == 0 && threadIdx.x == 0) toffx[gx * gy] = M;
  __syncthreads();
  const uint32_t start = stoff[s], end = stoff[s + 1];
  for (uint32_t e0 = start + threadIdx.x; e0 < end; e0 += kFineBatch * blockDim.x) {
    uint4 c[kFineBatch];
#pragma unroll
    for (int u = 0; u < kFineBatch; ++u) {
      const uint32_t e = e0 + u * blockDim.x;
      c[u] = e < end ? clist[e] : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < kFineBatch; ++u) {
      if (e0 + u * blockDim.x >= end) break;
      const uint64_t key = ((uint64_t)c[u].w << 32) | c[u].x;
      st_tiles(c[u].y, c[u].z, tx0, ty0, sts, [&](uint32_t t) { tkeys[atomicAdd(&cur[t], 1u)] = key; });
    }
  }
}

// Bitonic network whose comparators all put the smaller key at the lower
// index, over P = 2^p slots of which the first n hold keys: the slots >= n
// act as +infinity, so comparators reaching them are no-ops and no padding
// is stored.  Comparator q of a stage has its lower index
// i = ((q & ~(h - 1)) << 1) | (q & (h - 1)), h = k/2 for the "flip" of block
// size k (partner i ^ (k - 1)) and h = j for a half-cleaner (partner i + j).
__device__ __forceinline__ void tb_cx(uint64_t* a, uint32_t i, uint32_t l, uint32_t n) {
  if (l < n) {
    const uint64_t x = a[i], y = a[l];
    if (y < x) {
      a[i] = y;
      a[l] = x;
    }
  }
}

__device__ __forceinline__ void tb_flip(uint64_t* a, uint32_t n, uint32_t P, uint32_t k) {
  const uint32_t h = k >> 1;
  for (uint32_t q = threadIdx.x; q < (P >> 1); q += blockDim.x) {
    const uint32_t i = ((q & ~(h - 1)) << 1) | (q & (h - 1));
    tb_cx(a, i, i ^ (k - 1), n);
  }
}

__device__ __forceinline__ void tb_half(uint64_t* a, uint32_t n, uint32_t P, uint32_t j) {
  for (uint32_t q = threadIdx.x; q < (P >> 1); q += blockDim.x) {
    const uint32_t i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
    tb_cx(a, i, i + j, n);
  }
}

// Complete sort of a[0, n), P = 2^p >= n, by all threads of the block.
__device__ __forceinline__ void tb_bitonic(uint64_t* a, uint32_t n, uint32_t P) {
  for (uint32_t k = 2; k <= P; k <<= 1) {
    tb_flip(a, n, P, k);
    __syncthreads();
    for (uint32_t j = k >> 2; j > 0; j >>= 1) {
      tb_half(a, n, P, j);
      __syncthreads();
    }
  }
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t n) {
  uint32_t P = 1;
  while (P < n) P <<= 1;
  return P;
}

__device__ __forceinline__ void tb_range(int tile, const uint32_t* __restrict__ toff, uint32_t* start,
                                         uint32_t* end) {
  *start = toff[tile];
  *end = toff[tile + 1];
}

// ---------------------------------------------------------------------------
// In-register bitonic sort of one tile list by one wavefront.  Each lane holds
// KPL keys; key index i = lane * KPL + r (r the register), so compare-exchange
// stages with index stride below KPL run inside a lane's registers and the
// others exchange registers between lanes (DPP or swizzle within a row of 16,
// ds_bpermute across rows).  The network is tb_bitonic's (every comparator
// puts the smaller key at the lower index: a "flip" with partner i ^ (k - 1)
// per block size k, then half-cleaners with partner i ^ j), over P = 64 KPL
// slots; slots past the list hold ~0 and stay last.  No LDS and no barrier.
// ---------------------------------------------------------------------------
template <int M>
__device__ __forceinline__ uint32_t lane_xor_u32(uint32_t v) {
  if constexpr (M == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
  else if constexpr (M == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
  else if constexpr (M == 3) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x1B, 0xF, 0xF, false);
  else if constexpr (M == 7) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
  else if constexpr (M == 8) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
  else if constexpr (M == 15) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);
  else if constexpr (M < 32) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (M << 10) | 0x1F);
  else return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((threadIdx.x ^ M) << 2), (int)v);
}

template <int M>
__device__ __forceinline__ uint64_t lane_xor_u64(uint64_t v) {
  return ((uint64_t)lane_xor_u32<M>((uint32_t)(v >> 32)) << 32) | lane_xor_u32<M>((uint32_t)v);
}

// Keys below 0x7FEFFFFF << 32 (every key whose quantised depth is below
// 0x7FEFFFFF) are the bit patterns of finite non-negative doubles, mostly
// subnormal, and as doubles they order exactly as they do as integers (the
// kernels run with FP64 subnormals kept, IEEE mode).  Then one v_min_f64 and
// one v_max_f64 do a register compare-exchange, where the integer form takes
// a compare and four selects; the padding is the largest finite double,
// above every such key.  FD selects that form.
constexpr uint64_t kPadFd = 0x7FEFFFFFFFFFFFFFull;
constexpr uint32_t kFdMaxHi = 0x7FEFFFFFu;  // keys with a smaller high word qualify

template <bool FD>
__device__ __forceinline__ uint64_t sort_pad() { return FD ? kPadFd : ~0ull; }

template <bool FD>
__device__ __forceinline__ void reg_cx(uint64_t& a, uint64_t& b) {
  if constexpr (FD) {
    const double x = __longlong_as_double((long long)a), y = __longlong_as_double((long long)b);
    double lo, hi;
    asm("v_min_f64 %0, %2, %3\n\tv_max_f64 %1, %2, %3" : "=&v"(lo), "=&v"(hi) : "v"(x), "v"(y));
    a = (uint64_t)__double_as_longlong(lo);
    b = (uint64_t)__double_as_longlong(hi);
  } else {
    const bool sw = b < a;
    const uint64_t lo = sw ? b : a, hi = sw ? a : b;
    a = lo;
    b = hi;
  }
}

// Flip stage of block size K (K > KPL): register r pairs with register
// KPL-1-r of lane ^ (K/KPL - 1); the lane whose index bit K/2 is set keeps
// the larger key.
template <int KPL, int K>
__device__ __forceinline__ void wave_flip(uint64_t (&k)[KPL], int lane) {
  constexpr int LM = K / KPL - 1;
  const bool upper = (lane & ((K / KPL) >> 1)) != 0;
#pragma unroll
  for (int r = 0; r < (KPL + 1) / 2; ++r) {
    const int q = KPL - 1 - r;
    const uint64_t tr = lane_xor_u64<LM>(k[q]);  // partner's register r's pair
    const uint64_t tq = r == q ? tr : lane_xor_u64<LM>(k[r]);
    // branch-free: the lower lane keeps the smaller key, the upper one the larger (keys are distinct
    // except the ~0 padding, for which either choice is the same value)
    const bool take_r = (tr < k[r]) != upper;
    const bool take_q = (tq < k[q]) != upper;
    k[r] = take_r ? tr : k[r];
    if (r != q) k[q] = take_q ? tq : k[q];
  }
}

// Half-cleaner of index stride J (J >= KPL): register r pairs with register r
// of lane ^ (J/KPL); the lane with that bit set keeps the larger key.
template <int KPL, int J>
__device__ __forceinline__ void wave_half(uint64_t (&k)[KPL], int lane) {
  constexpr int LM = J / KPL;
  const bool upper = (lane & LM) != 0;
#pragma unroll
  for (int r = 0; r < KPL; ++r) {
    const uint64_t t = lane_xor_u64<LM>(k[r]);
    const bool take = (t < k[r]) != upper;
    k[r] = take ? t : k[r];
  }
}

template <bool FD, int KPL, int K, int J>
__device__ __forceinline__ void wave_halves(uint64_t (&k)[KPL], int lane) {
  if constexpr (J >= 1) {
    if constexpr (J >= KPL) {
      wave_half<KPL, J>(k, lane);
    } else {
#pragma unroll
      for (int r = 0; r < KPL; ++r)
        if ((r & J) == 0) reg_cx<FD>(k[r], k[r | J]);
    }
    wave_halves<FD, KPL, K, J / 2>(k, lane);
  }
}

template <bool FD, int KPL, int K>
__device__ __forceinline__ void wave_bitonic_from(uint64_t (&k)[KPL], int lane) {
  if constexpr (K <= 64 * KPL) {
    if constexpr (K <= KPL) {
#pragma unroll
      for (int r = 0; r < KPL; ++r) {
        const int q = r ^ (K - 1);
        if (r < q) reg_cx<FD>(k[r], k[q]);
      }
    } else {
      wave_flip<KPL, K>(k, lane);
    }
    wave_halves<FD, KPL, K, K / 4>(k, lane);
    wave_bitonic_from<FD, KPL, K * 2>(k, lane);
  }
}

// Loads keys[e] (e = r * 64 + lane below n, coalesced; any placement is
// sorted alike) into k and sorts them over the wavefront, with the double
// compare-exchange when every key qualifies (decided for the wavefront) and
// the integer one otherwise.  Slots past n hold the form's padding.
template <int KPL>
__device__ __forceinline__ void wave_sort_keys(const uint64_t* __restrict__ keys, uint32_t n, int lane,
                                               uint64_t (&k)[KPL]) {
  uint32_t hmax = 0;
#pragma unroll
  for (int r = 0; r < KPL; ++r) {
    const uint32_t e = (uint32_t)(r * 64 + lane);
    k[r] = e < n ? keys[e] : 0ull;
    hmax = max(hmax, (uint32_t)(k[r] >> 32));
  }
  hmax = wave_max_u32(hmax);
  if (hmax < kFdMaxHi) {
#pragma unroll
    for (int r = 0; r < KPL; ++r)
      if ((uint32_t)(r * 64 + lane) >= n) k[r] = kPadFd;
    wave_bitonic_from<true, KPL, 2>(k, lane);
  } else {
#pragma unroll
    for (int r = 0; r < KPL; ++r)
      if ((uint32_t)(r * 64 + lane) >= n) k[r] = ~0ull;
    wave_bitonic_from<false, KPL, 2>(k, lane);
  }
}

// Index of sorted position x in the LDS buffers of the in-tile sorts: two
// spare slots after every 16 keys, so a lane's 16 consecutive keys start 144
// bytes after its neighbour's and the 16-byte accesses of a wavefront spread
// over the banks.
__device__ __forceinline__ uint32_t tb_pad(uint32_t x) { return x + 2 * (x >> 4); }
constexpr int kTbPadded = kTbSortSmall + 2 * (kTbSortSmall / 16);
constexpr int kTbPaddedReg = kTbSortReg + 2 * (kTbSortReg / 16);

// Writes sorted keys s[tb_pad(e)], e < n, as ids (and tile-tagged keys)
// with consecutive threads on consecutive positions.
__device__ __forceinline__ void tb_emit(const uint64_t* __restrict__ s, uint32_t n, int32_t* __restrict__ ids,
                                        uint64_t* __restrict__ keys_out, uint64_t thi, uint32_t start) {
  for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
    const uint64_t key = s[tb_pad(e)];
    ids[start + e] = (int32_t)(uint32_t)key;
    if (keys_out) keys_out[start + e] = thi | (key >> 32);
  }
}

// Sort the tile list tkeys[start, start + n) (n <= 64 KPL) by (depth_q, id)
// into ids[start, start + n).  Lane l ends with the keys of positions
// l*KPL .. l*KPL+KPL-1; they pass through LDS (s) so that the id stores are
// coalesced.
template <int KPL>
__device__ __forceinline__ void tb_wave_sort(const uint64_t* __restrict__ tkeys, int32_t* __restrict__ ids,
                                             uint64_t* __restrict__ keys_out, uint64_t thi, uint32_t start,
                                             uint32_t n, int lane, uint64_t* __restrict__ s) {
  uint64_t k[KPL];
  wave_sort_keys<KPL>(tkeys + start, n, lane, k);
#pragma unroll
  for (int r = 0; r < KPL; ++r) s[tb_pad((uint32_t)(lane * KPL + r))] = k[r];
  __syncthreads();
  tb_emit(s, n, ids, keys_out, thi, start);
}

// One wavefront per tile: its range, and the sort of its ids when the list
// has at most kTbSortReg entries (in registers, KPL the smallest power of two
// with 64 KPL >= n).  Longer lists are left to k_tb_sort_long.
__global__ __launch_bounds__(64) void k_tb_sort(int tiles, const uint32_t* __restrict__ hdr,
                                                const uint32_t* __restrict__ toff,
                                                const uint64_t* __restrict__ tkeys,
                                                int32_t* __restrict__ ids,
                                                uint64_t* __restrict__ keys_out,
                                                int2* __restrict__ ranges) {
  static_assert(kTbSortReg == 64 * 16, "at most 16 keys per lane");
  const int tile = blockIdx.x;
  const int lane = threadIdx.x;
  const uint32_t M = hdr[kHdrMEff];
  uint32_t start, end;
  tb_range(tile, toff, &start, &end);
  // empty tiles keep (0,0), and so does every tile when M == 1 (kernel.cu:165-187)
  if (lane == 0) ranges[tile] = (M <= 1 || end == start) ? make_int2(0, 0) : make_int2((int)start, (int)end);
  if (M == 0 || end == start) return;
  const uint32_t n = end - start;
  if (n > (uint32_t)kTbSortReg) return;  // left to k_tb_sort_long
  const uint64_t thi = (uint64_t)tile << 32;
  __shared__ uint64_t s[kTbPaddedReg];
  if (n <= 64) tb_wave_sort<1>(tkeys, ids, keys_out, thi, start, n, lane, s);
  else if (n <= 128) tb_wave_sort<2>(tkeys, ids, keys_out, thi, start, n, lane, s);
  else if (n <= 256) tb_wave_sort<4>(tkeys, ids, keys_out, thi, start, n, lane, s);
  else if (n <= 512) tb_wave_sort<8>(tkeys, ids, keys_out, thi, start, n, lane, s);
  else tb_wave_sort<16>(tkeys, ids, keys_out, thi, start, n, lane, s);
}


// Merge path over two sorted runs of L keys each, s[a0, a0 + L) and
// s[b0, b0 + L) (unpadded positions; keys distinct except the ~0 padding):
// the cnt outputs from diagonal d, handed to out(q, key) in order.  The split
// point is the number of keys the first d outputs take from the first run,
// found by binary search; the first run wins ties, so the merge is stable.
template <class F>
__device__ __forceinline__ void tb_merge_path(const uint64_t* __restrict__ s, uint32_t a0, uint32_t b0, uint32_t L,
                                              uint32_t d, int cnt, F out) {
  uint32_t lo = d > L ? d - L : 0u, hi = d < L ? d : L;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s[tb_pad(a0 + mid)] <= s[tb_pad(b0 + d - 1 - mid)]) lo = mid + 1;
    else hi = mid;
  }
  uint32_t i = lo, j = d - lo;
  uint64_t x = i < L ? s[tb_pad(a0 + i)] : ~0ull;
  uint64_t y = j < L ? s[tb_pad(b0 + j)] : ~0ull;
  for (int q = 0; q < cnt; ++q) {
    const bool ta = j >= L || (i < L && x <= y);
    out(q, ta ? x : y);
    if (ta) {
      ++i;
      x = i < L ? s[tb_pad(a0 + i)] : ~0ull;
    } else {
      ++j;
      y = j < L ? s[tb_pad(b0 + j)] : ~0ull;
    }
  }
}

// Lists of kTbSortReg + 1 .. kTbSortSmall entries, by the 256 threads of a
// block.  Wavefront w sorts keys [1024 w, 1024 w + 1024) in registers (the
// runs past the list are ~0), the runs go to LDS, and merge path joins them:
// two runs in one pass, four in two; each pass holds its outputs in registers
// until every thread has read its inputs, then writes them back to the same
// buffer, and the result leaves through tb_emit (coalesced id stores).
__device__ __forceinline__ void tb_sort_runs(uint64_t* __restrict__ s, const uint64_t* __restrict__ tkeys,
                                             int32_t* __restrict__ ids, uint64_t* __restrict__ keys_out,
                                             int tile, uint32_t start, uint32_t n) {
  static_assert(kTbSortReg == 1024 && kTbSortSmall == 4 * kTbSortReg, "up to four register runs of 1024 keys");
  const uint32_t w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const uint32_t runs = n <= 2u * kTbSortReg ? 2u : 4u;
  if (w < runs) {
    // each run is sorted with its own form; its padding, sorted last, leaves as ~0 so that under the merge's
    // integer compares it follows every key of every run (another run's integer form may hold keys above the
    // double form's padding)
    const uint32_t base = w * kTbSortReg;
    uint64_t k[16];
    if (base < n) {
      wave_sort_keys<16>(tkeys + start + base, n - base, lane, k);
    } else {
#pragma unroll
      for (int r = 0; r < 16; ++r) k[r] = ~0ull;
    }
#pragma unroll
    for (int r = 0; r < 16; r += 2) {
      const uint32_t i = base + (uint32_t)lane * 16 + r;
      *reinterpret_cast<ulonglong2*>(&s[tb_pad(i)]) =
          make_ulonglong2(i < n ? k[r] : ~0ull, i + 1 < n ? k[r + 1] : ~0ull);
    }
  }
  __syncthreads();
  const uint64_t thi = (uint64_t)tile << 32;
  uint64_t o[16];
  if (runs == 2) {
    const uint32_t d = threadIdx.x * 8;
    tb_merge_path(s, 0, kTbSortReg, kTbSortReg, d, 8, [&](int q, uint64_t key) { o[q] = key; });
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; q += 2)
      *reinterpret_cast<ulonglong2*>(&s[tb_pad(d + q)]) = make_ulonglong2(o[q], o[q + 1]);
  } else {
    {
      // runs (0,1) and (2,3), 128 threads each, 16 outputs per thread
      const uint32_t a0 = (threadIdx.x >> 7) * 2 * kTbSortReg, d = (threadIdx.x & 127) * 16;
      tb_merge_path(s, a0, a0 + kTbSortReg, kTbSortReg, d, 16, [&](int q, uint64_t key) { o[q] = key; });
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 16; q += 2)
        *reinterpret_cast<ulonglong2*>(&s[tb_pad(a0 + d + q)]) = make_ulonglong2(o[q], o[q + 1]);
    }
    __syncthreads();
    const uint32_t d = threadIdx.x * 16;
    tb_merge_path(s, 0, 2 * kTbSortReg, 2 * kTbSortReg, d, 16, [&](int q, uint64_t key) { o[q] = key; });
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q += 2)
      *reinterpret_cast<ulonglong2*>(&s[tb_pad(d + q)]) = make_ulonglong2(o[q], o[q + 1]);
  }
  __syncthreads();
  tb_emit(s, n, ids, keys_out, thi, start);
}

// Lists longer than kTbSortSmall: chunks of kTbSortLds keys sorted in LDS,
// then merged by the network's stages that cross chunks in global memory
// (tkeys) and the rest of each merge in LDS.
__device__ __forceinline__ void tb_sort_huge(uint64_t* __restrict__ s, uint64_t* __restrict__ tkeys,
                                             int32_t* __restrict__ ids, uint64_t* __restrict__ keys_out,
                                             int tile, uint32_t start, uint32_t n) {
  constexpr uint32_t C = kTbSortLds;
  const uint64_t thi = (uint64_t)tile << 32;
  uint64_t* seg = tkeys + start;
  const uint32_t P = pow2_at_least(n);
  // chunks sorted in LDS (the network's blocks of size <= C)
  for (uint32_t c0 = 0; c0 < n; c0 += C) {
    const uint32_t nc = min(C, n - c0);
    for (uint32_t e = threadIdx.x; e < nc; e += blockDim.x) s[e] = seg[c0 + e];
    __syncthreads();
    tb_bitonic(s, nc, C);
    for (uint32_t e = threadIdx.x; e < nc; e += blockDim.x) seg[c0 + e] = s[e];
    __threadfence_block();
    __syncthreads();
  }
  for (uint32_t k = 2 * C; k <= P; k <<= 1) {
    tb_flip(seg, n, P, k);
    __threadfence_block();
    __syncthreads();
    for (uint32_t j = k >> 2; j >= C; j >>= 1) {
      tb_half(seg, n, P, j);
      __threadfence_block();
      __syncthreads();
    }
    for (uint32_t c0 = 0; c0 < n; c0 += C) {
      const uint32_t nc = min(C, n - c0);
      for (uint32_t e = threadIdx.x; e < nc; e += blockDim.x) s[e] = seg[c0 + e];
      __syncthreads();
      for (uint32_t j = C >> 1; j > 0; j >>= 1) {
        tb_half(s, nc, C, j);
        __syncthreads();
      }
      for (uint32_t e = threadIdx.x; e < nc; e += blockDim.x) seg[c0 + e] = s[e];
      __threadfence_block();
      __syncthreads();
    }
  }
  for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
    const uint64_t k = seg[e];
    ids[start + e] = (int32_t)(uint32_t)k;
    if (keys_out) keys_out[start + e] = thi | (k >> 32);
  }
}

// Every list longer than kTbSortReg entries, one 256-thread block per tile
// (the blocks of tiles with shorter lists, which k_tb_sort sorted, exit at
// once).
__global__ __launch_bounds__(256) void k_tb_sort_long(const uint32_t* __restrict__ hdr,
                                                      const uint32_t* __restrict__ toff,
                                                      uint64_t* __restrict__ tkeys, int32_t* __restrict__ ids,
                                                      uint64_t* __restrict__ keys_out) {
  static_assert(kTbSortLds <= kTbPadded, "the chunk buffer fits in the run buffer");
  __shared__ __align__(16) uint64_t s[kTbPadded];
  if (hdr[kHdrMEff] == 0) return;
  const int tile = blockIdx.x;
  uint32_t start, end;
  tb_range(tile, toff, &start, &end);
  const uint32_t n = end - start;
  if (n <= (uint32_t)kTbSortReg) return;
  if (n <= (uint32_t)kTbSortSmall) tb_sort_runs(s, tkeys, ids, keys_out, tile, start, n);
  else tb_sort_huge(s, tkeys, ids, keys_out, tile, start, n);
}

int bin_reset_header(void* ws, int words, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3((words + 255) / 256), dim3(256), 0, s, (uint32_t*)ws, 0u, (int64_t)words);
  return check_launch("bin_reset_header");
}

}  // namespace rgp

using namespace rgp;

extern "C" size_t rgp_bin_workspace_bytes(int n, int64_t capacity, int width, int height) {
  if (n < 0 || width <= 0 || height <= 0) return 0;
  return bin_layout(n, capacity, width, height).total;
}

extern "C" const uint32_t* rgp_bin_header(const void* ws) { return (const uint32_t*)ws; }

extern "C" int rgp_bin_scan(int n, int width, int height, void* ws, int64_t capacity,
                            hipStream_t s) {
  if (n < 0 || width <= 0 || height <= 0) return set_error(RGP_ERR_ARG, "rgp_bin_scan: bad size");
  BinLayout L = bin_layout(n, capacity, width, height);
  char* base = (char*)ws;
  uint32_t* hdr = (uint32_t*)(base + L.header);
  int tile_bytes = 0;
  while (tile_bytes < 4 && ((uint32_t)(L.tiles - 1) >> (8 * tile_bytes)) != 0) ++tile_bytes;
  if (L.tb) {
    // coarse count -> M, the coarse offsets and the coarse lists -> tile counts and tile-row sums
    // (the tile offsets are formed by the fine scatter of rgp_bin_sort); with no Gaussians the
    // cleared header already says M = 0
    uint32_t* sttot = (uint32_t*)(base + L.sttot);
    uint32_t* stoff = (uint32_t*)(base + L.stoff);
    uint32_t* boff = (uint32_t*)(base + L.boff);
    uint32_t* bpart = (uint32_t*)(base + L.bpart);
    uint4* clist = (uint4*)(base + L.clist);
    uint32_t* toff = (uint32_t*)(base + L.toff);
    const size_t st_lds = (size_t)L.st * 4;
    if (n > 0) {
      hipLaunchKernelGGL(k_st_count, dim3(L.nb), dim3(kTbThreads), st_lds, s, n, L.chunk, L.st, L.sts, L.sgx,
                         (const uint32_t*)(base + L.counts), (const uint2*)(base + L.rects), sttot, boff, bpart);
      hipLaunchKernelGGL(k_st_scatter, dim3(L.nb), dim3(kTbThreads), st_lds, s, n, L.chunk, L.st, L.sts, L.sgx,
                         L.nb, (const uint32_t*)(base + L.counts), (const uint2*)(base + L.rects),
                         (const uint32_t*)(base + L.dqs), (const uint32_t*)sttot, (const uint32_t*)boff,
                         (const uint32_t*)bpart, hdr, stoff, L.cap, clist);
      hipLaunchKernelGGL(k_st_fine_count, dim3(L.st), dim3(1024), 0, s, L.sts, L.sgx, L.gx, L.gy,
                         (const uint32_t*)stoff, (const uint32_t*)hdr, (const uint4*)clist, toff,
                         (uint32_t*)(base + L.rowsum));
    }
    return check_launch("rgp_bin_scan");
  }
  if (L.nscan > 0)
    hipLaunchKernelGGL(k_scan_reduce, dim3(L.nscan), dim3(256), 0, s, n,
                       (const uint32_t*)(base + L.counts), (n + 63) / 64,
                       (const uint32_t*)(base + L.dqw), (uint32_t*)(base + L.partials),
                       (uint32_t*)(base + L.dqpart));
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, L.nscan,
                     (uint32_t*)(base + L.partials), (const uint32_t*)(base + L.dqpart), hdr, L.cap,
                     tile_bytes);
  if (L.nscan > 0)
    hipLaunchKernelGGL(k_scan_down, dim3(L.nscan), dim3(256), 0, s, n,
                       (uint32_t*)(base + L.counts), (const uint32_t*)(base + L.partials));
  return check_launch("rgp_bin_scan");
}

extern "C" int rgp_bin_sort(int n, const float* depths, int width, int height, void* ws,
                            int64_t capacity, int32_t* ids_out, uint64_t* keys_out,
                            int32_t* ranges_out, hipStream_t s) {
  if (n < 0 || width <= 0 || height <= 0) return set_error(RGP_ERR_ARG, "rgp_bin_sort: bad size");
  if (capacity > (int64_t)INT32_MAX)
    return set_error(RGP_ERR_ARG, "rgp_bin_sort: capacity above 2^31 - 1");
  BinLayout L = bin_layout(n, capacity, width, height);
  char* base = (char*)ws;
  const uint32_t* hdr = (const uint32_t*)(base + L.header);
  if (L.tb) {
    const uint32_t* toffx = (const uint32_t*)(base + L.toffx);
    uint64_t* tkeys = (uint64_t*)(base + L.tkeys);
    (void)depths;  // the coarse lists carry the quantised depths written by the preprocess
    hipLaunchKernelGGL(k_st_fine_scatter, dim3(L.st), dim3(1024), 0, s, L.sts, L.sgx, L.gx, L.gy,
                       (const uint32_t*)(base + L.stoff), hdr, (const uint4*)(base + L.clist),
                       (const uint32_t*)(base + L.toff), (const uint32_t*)(base + L.rowsum),
                       (uint32_t*)toffx, tkeys);
    hipLaunchKernelGGL(k_tb_sort, dim3(L.tiles), dim3(64), 0, s, L.tiles, hdr, toffx, (const uint64_t*)tkeys,
                       ids_out, keys_out, (int2*)ranges_out);
    hipLaunchKernelGGL(k_tb_sort_long, dim3(L.tiles), dim3(256), 0, s, hdr, toffx, tkeys, ids_out, keys_out);
    return check_launch("rgp_bin_sort");
  }
  uint64_t* k0 = (uint64_t*)(base + L.keys0);
  uint64_t* k1 = (uint64_t*)(base + L.keys1);
  int32_t* v0 = (int32_t*)(base + L.vals0);
  int32_t* v1 = (int32_t*)(base + L.vals1);
  uint32_t* hist = (uint32_t*)(base + L.hist);
  uint32_t* totals = (uint32_t*)(base + L.totals);
  if (n > 0 && L.cap > 0)
    hipLaunchKernelGGL(k_dup, dim3((n + 255) / 256), dim3(256), 0, s, n,
                       (const uint32_t*)(base + L.counts), (const uint2*)(base + L.rects), depths,
                       L.gx, hdr, k0, v0);
  if (L.nblk_cap > 0) {
    for (int p = 0; p < 8; ++p) {
      hipLaunchKernelGGL(k_sort_hist, dim3(L.nblk_cap), dim3(256), 0, s, p, hdr, k0, k1, hist,
                         L.nblk_cap);
      hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(256), 0, s, p, hdr, hist, totals,
                         L.nblk_cap);
      hipLaunchKernelGGL(k_sort_scatter, dim3(L.nblk_cap), dim3(256), 0, s, p, hdr, k0, k1, v0,
                         v1, ids_out, hist, totals, L.nblk_cap);
    }
    const int64_t fin = (L.cap + 255) / 256;
    hipLaunchKernelGGL(k_sort_finish, dim3((unsigned)(fin < 1024 ? fin : 1024)), dim3(256), 0, s, hdr,
                       k0, k1, v0, ids_out, keys_out);
  }
  hipLaunchKernelGGL(k_ranges, dim3((L.tiles + 255) / 256), dim3(256), 0, s, L.tiles, hdr, k0, k1,
                     (int2*)ranges_out);
  return check_launch("rgp_bin_sort");
}

extern "C" int rgp_fill_u32(void* dst, uint32_t value, int64_t n, hipStream_t s) {
  if (n < 0) return set_error(RGP_ERR_ARG, "rgp_fill_u32: n < 0");
  if (n == 0) return RGP_OK;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     (uint32_t*)dst, value, n);
  return check_launch("rgp_fill_u32");
}

extern "C" int rgp_bin_scan_views(int views, int n, int width, int height, void* ws, size_t ws_stride,
                                  int64_t capacity, hipStream_t s) {
  if (views < 1 || views > RGP_MAX_VIEWS) return set_error(RGP_ERR_ARG, "rgp_bin_scan_views: bad views");
  for (int v = 0; v < views; ++v)
    if (int rc = rgp_bin_scan(n, width, height, (char*)ws + (size_t)v * ws_stride, capacity, s)) return rc;
  return RGP_OK;
}

extern "C" int rgp_bin_sort_views(int views, int n, const float* depths, int width, int height, void* ws,
                                  size_t ws_stride, int64_t capacity, int32_t* ids_out, int32_t* ranges_out,
                                  hipStream_t s) {
  if (views < 1 || views > RGP_MAX_VIEWS) return set_error(RGP_ERR_ARG, "rgp_bin_sort_views: bad views");
  const int64_t tiles = (int64_t)((width + kTile - 1) / kTile) * ((height + kTile - 1) / kTile);
  for (int v = 0; v < views; ++v)
    if (int rc = rgp_bin_sort(n, depths + (size_t)v * n, width, height, (char*)ws + (size_t)v * ws_stride,
                              capacity, ids_out + v * capacity, nullptr, ranges_out + v * tiles * 2, s))
      return rc;
  return RGP_OK;
}
