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
// Binning workspace layout.  All regions 256-byte aligned.  Two sort paths:
//
// Tile-bucket path (tiles <= kTbMaxTiles, every resolution up to 4K):
//   header   : 64 u32 control words (below)
//   counts   : N u32 per-Gaussian tile counts
//   rects    : N uint2, packed (x0 | y0 << 16, x1 | y1 << 16)
//   dqw      : ceil(N / 64) u32, max uint32(depth * 1000) per 64 Gaussians
//   tbpart   : tiles x nb u32, pairs of tile t from Gaussian block b at
//              [t * nb + b]; the scan turns it into inclusive offsets
//   partials : scan block sums of tbpart
//   tkeys    : capacity u64 (uint32(depth * 1000) << 32 | Gaussian id), the
//              keys of tile lists too long to sort in LDS
//   bigq     : tiles u32, the queue of those tiles
//   (the ids are scattered into the caller's id buffer and sorted there)
//
// Radix path (larger grids):
//   counts are turned into inclusive offsets by the scan; keys0/1 hold
//   (tile << 32 | uint32(depth * 1000)), vals0/1 the ids; hist/totals are
//   the per-pass digit counts of an LSD radix sort with 8-bit digits.
// ---------------------------------------------------------------------------
constexpr int kHdrM = 0;         // pairs produced by the scan (u32)
constexpr int kHdrOverflow = 1;  // 1 when M > capacity
constexpr int kHdrDqMax = 2;     // max uint32(depth*1000) over binned Gaussians (radix path)
constexpr int kHdrMEff = 3;      // M when it fits, else 0
constexpr int kHdrCur = 4;       // 9 words: source buffer of sort pass p (p = 0..8)
constexpr int kHdrSkip = 13;     // 8 words: pass p is skipped (all digits equal)
constexpr int kHdrLast = 21;     // index of the last executed pass, or -1
constexpr int kHdrNblk = 22;     // active sort blocks
constexpr int kHdrBig = 23;      // tiles queued for the long-list sort (tile-bucket path)
constexpr int kHdrWords = 64;

constexpr int kScanChunk = 4096;  // counts per scan block (256 threads x 16)
constexpr int kSortItems = 16;    // keys per thread per sort block
constexpr int kSortTile = 256 * kSortItems;

constexpr int kTbMaxTiles = 32768;   // per-tile LDS counters of one block: 128 KiB
constexpr int kTbThreads = 1024;     // threads of the count / scatter blocks
constexpr int kTbMinChunk = 4096;    // Gaussians per count / scatter block, at least
constexpr int kTbMaxBlocks = 256;
constexpr int kTbBig = 32;           // rects above this many tiles are walked per wavefront
constexpr int kTbSortReg = 1024;     // lists sorted in registers by the one-wavefront-per-tile kernel
constexpr int kTbSortSmall = 2048;   // lists sorted in LDS by the one-block-per-tile kernel (16 KiB)
constexpr int kTbSortLds = 4096;     // keys of one long list sorted in LDS; longer lists
                                     // are merged through global memory
constexpr int kTbBigBlocks = 1024;   // blocks of the long-list sort (grid-stride over its queue)

struct BinLayout {
  size_t header, counts, rects, dqw, partials, dqpart, keys0, keys1, vals0, vals1, hist, totals;
  size_t tbpart, tkeys, bigq, total;
  int gx, gy, tiles, nscan, nblk_cap;
  bool tb;      // tile-bucket path
  int nb;       // count / scatter blocks (tile-bucket path)
  int chunk;    // Gaussians per count / scatter block
  int64_t npart;  // tiles * nb
  int64_t cap;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

inline BinLayout bin_layout(int n, int64_t cap, int width, int height) {
  BinLayout L;
  L.gx = (width + kTile - 1) / kTile;
  L.gy = (height + kTile - 1) / kTile;
  L.tiles = L.gx * L.gy;
  L.cap = cap < 0 ? 0 : cap;
  L.tb = L.tiles <= kTbMaxTiles;
  size_t o = 0;
  L.header = o; o = align256(o + kHdrWords * 4);
  L.counts = o; o = align256(o + (size_t)n * 4);
  L.rects = o; o = align256(o + (size_t)n * 8);
  L.dqw = o; o = align256(o + (size_t)((n + 63) / 64) * 4);
  L.keys0 = L.keys1 = L.vals0 = L.vals1 = L.hist = L.totals = L.dqpart = L.tbpart = L.tkeys = L.bigq = 0;
  L.nblk_cap = 0;
  if (L.tb) {
    int nb = (n + kTbMinChunk - 1) / kTbMinChunk;
    nb = nb < 1 ? 1 : (nb > kTbMaxBlocks ? kTbMaxBlocks : nb);
    L.nb = nb;
    L.chunk = (n + nb - 1) / nb;
    L.npart = (int64_t)L.tiles * nb;
    L.nscan = (int)((L.npart + kScanChunk - 1) / kScanChunk);
    L.tbpart = o; o = align256(o + (size_t)(L.npart + 1) * 4);
    L.partials = o; o = align256(o + (size_t)(L.nscan + 1) * 4);
    L.tkeys = o; o = align256(o + (size_t)L.cap * 8);
    L.bigq = o; o = align256(o + (size_t)L.tiles * 4);
  } else {
    L.nb = 0;
    L.chunk = 0;
    L.npart = 0;
    L.nscan = (n + kScanChunk - 1) / kScanChunk;
    L.nblk_cap = (int)((L.cap + kSortTile - 1) / kSortTile);
    L.partials = o; o = align256(o + (size_t)(L.nscan + 1) * 4);
    L.dqpart = o; o = align256(o + (size_t)(L.nscan + 1) * 4);
    L.keys0 = o; o = align256(o + (size_t)L.cap * 8);
    L.keys1 = o; o = align256(o + (size_t)L.cap * 8);
    L.vals0 = o; o = align256(o + (size_t)L.cap * 4);
    L.vals1 = o; o = align256(o + (size_t)L.cap * 4);
    L.hist = o; o = align256(o + (size_t)256 * (L.nblk_cap > 0 ? L.nblk_cap : 1) * 4);
    L.totals = o; o = align256(o + 256 * 4);
  }
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
