// Internal declarations shared by the translation units of librgp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rgp_device.h"

namespace rgp {

// Error reporting: thread-local message, negative return codes.
int set_error(int code, const char* fmt, ...);
int check_launch(const char* who);

// ---------------------------------------------------------------------------
// Binning workspace layout.  All regions 256-byte aligned.
//   header   : 64 u32 control words (below)
//   counts   : N u32, tile counts, turned into inclusive offsets by the scan
//   rects    : N uint2, packed (x0 | y0 << 16, x1 | y1 << 16)
//   partials : scan block sums
//   keys0/1  : capacity u64 (tile << 32 | uint32(depth * 1000))
//   vals0/1  : capacity i32 Gaussian ids
//   hist     : 256 digits x sort blocks, per-pass digit counts
//   totals   : 256 digit totals
// ---------------------------------------------------------------------------
constexpr int kHdrM = 0;         // pairs produced by the scan (u32)
constexpr int kHdrOverflow = 1;  // 1 when M > capacity
constexpr int kHdrDqMax = 2;     // max uint32(depth*1000) over binned Gaussians
constexpr int kHdrMEff = 3;      // M when it fits, else 0
constexpr int kHdrCur = 4;       // 9 words: source buffer of sort pass p (p = 0..8)
constexpr int kHdrSkip = 13;     // 8 words: pass p is skipped (all digits equal)
constexpr int kHdrLast = 21;     // index of the last executed pass, or -1
constexpr int kHdrNblk = 22;     // active sort blocks
constexpr int kHdrWords = 64;

constexpr int kScanChunk = 4096;  // counts per scan block (256 threads x 16)
constexpr int kSortItems = 16;    // keys per thread per sort block
constexpr int kSortTile = 256 * kSortItems;

struct BinLayout {
  size_t header, counts, rects, dqw, partials, dqpart, keys0, keys1, vals0, vals1, hist, totals, total;
  int gx, gy, tiles, nscan, nblk_cap;
  int64_t cap;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

inline BinLayout bin_layout(int n, int64_t cap, int width, int height) {
  BinLayout L;
  L.gx = (width + kTile - 1) / kTile;
  L.gy = (height + kTile - 1) / kTile;
  L.tiles = L.gx * L.gy;
  L.cap = cap < 0 ? 0 : cap;
  L.nscan = (n + kScanChunk - 1) / kScanChunk;
  L.nblk_cap = (int)((L.cap + kSortTile - 1) / kSortTile);
  size_t o = 0;
  L.header = o; o = align256(o + kHdrWords * 4);
  L.counts = o; o = align256(o + (size_t)n * 4);
  L.rects = o; o = align256(o + (size_t)n * 8);
  L.dqw = o; o = align256(o + (size_t)((n + 63) / 64) * 4);
  L.partials = o; o = align256(o + (size_t)(L.nscan + 1) * 4);
  L.dqpart = o; o = align256(o + (size_t)(L.nscan + 1) * 4);
  L.keys0 = o; o = align256(o + (size_t)L.cap * 8);
  L.keys1 = o; o = align256(o + (size_t)L.cap * 8);
  L.vals0 = o; o = align256(o + (size_t)L.cap * 4);
  L.vals1 = o; o = align256(o + (size_t)L.cap * 4);
  L.hist = o; o = align256(o + (size_t)256 * (L.nblk_cap > 0 ? L.nblk_cap : 1) * 4);
  L.totals = o; o = align256(o + 256 * 4);
  L.total = o;
  return L;
}

int bin_reset_header(void* ws, hipStream_t s);

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_xor(v, o);
    v = t > v ? t : v;
  }
  return v;
}

}  // namespace rgp
