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
// Binning workspace layout.  All regions 256-byte aligned.  Two sort paths.
//
// Common regions:
//   header   : 64 u32 control words (below); on the tile-bucket path followed
//              by sttot, S + 1 u32 (S super-tiles), cleared with the header
//   counts   : N u32 per-Gaussian tile counts
//   rects    : N uint2, packed (x0 | y0 << 16, x1 | y1 << 16)
//   dqw      : ceil(N / 64) u32, max uint32(depth * 1000) per 64 Gaussians
//   dqs      : N u32, uint32(depth * 1000) of each binned Gaussian
//
// Tile-bucket path (tiles <= kTbMaxTiles, every resolution up to 4K), two
// levels: Gaussians are first bucketed into super-tiles of kst x kst tiles
// (kst = 4 up to kStSmallGrid tiles, 8 above, so that either way there are a
// few hundred super-tiles),
// then each super-tile's block writes the complete lists of its tiles.
//   sttot    : coarse pairs per super-tile (blocks reserve their slots with
//              atomics), then scanned into exclusive offsets, sttot[S] = total
//   boff     : nb x S u32, end of Gaussian block b's slots in super-tile s's
//              coarse list, relative to the list's start, at [b * S + s]
//   bpart    : nb u32, the tile pairs of Gaussian block b (sum: M)
//   clist    : capacity uint4 (Gaussian id, rect.x, rect.y, its dqs), the
//              coarse lists, super-tile major (no more entries than M)
//   toff     : tiles + 1 u32, exclusive offsets of the tile lists
//   tkeys    : capacity u64 (uint32(depth * 1000) << 32 | Gaussian id), every
//              tile list's sort keys, written by the fine scatter (and the
//              merge space of lists too long to sort in LDS)
//   (the in-tile sorts read tkeys and write the caller's id buffer)
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
constexpr int kHdrWords = 64;

constexpr int kScanChunk = 4096;  // counts per scan block (256 threads x 16)
constexpr int kSortItems = 16;    // keys per thread per sort block
constexpr int kSortTile = 256 * kSortItems;

constexpr int kTbMaxTiles = 32768;   // largest grid of the tile-bucket path (3840x2160 is 32400)
constexpr int kStMax = 8;            // largest super-tile edge in tiles (128 x 128 pixels)
constexpr int kStSmallGrid = 16384;  // grids up to this many tiles (2560x1440 has 14400) use edge 4
constexpr int kTbThreads = 1024;     // threads of the coarse count / scatter blocks
constexpr int kTbMinChunk = 4096;    // Gaussians per coarse count / scatter block, at least
constexpr int kTbMaxBlocks = 256;
constexpr int kTbBig = 32;           // rects above this many super-tiles are walked per wavefront
constexpr int kTbSortReg = 1024;     // lists sorted in registers by the one-wavefront-per-tile kernel
constexpr int kTbSortSmall = 4096;   // lists sorted as register runs of kTbSortReg keys merged through LDS
constexpr int kTbSortLds = 4096;     // keys of one longer list sorted in LDS at a time; the rest of
                                     // its merges run in global memory

struct BinLayout {
  size_t header, counts, rects, dqw, dqs, partials, dqpart, keys0, keys1, vals0, vals1, hist, totals;
  size_t sttot, rowsum, stoff, boff, bpart, clist, toff, toffx, tkeys, total;
  int gx, gy, tiles, nscan, nblk_cap;
  int reset_words;  // words cleared before binning: the header (and sttot, rowsum)
  bool tb;      // tile-bucket path
  int sts, sgx, sgy, st;  // super-tile edge log2 and grid (tile-bucket path)
  int nb;       // coarse count / scatter blocks (tile-bucket path)
  int chunk;    // Gaussians per coarse count / scatter block
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
  L.sts = L.sgx = L.sgy = L.st = 0;
  if (L.tb) {
    L.sts = L.tiles > kStSmallGrid ? 3 : 2;
    const int kst = 1 << L.sts;
    L.sgx = (L.gx + kst - 1) / kst;
    L.sgy = (L.gy + kst - 1) / kst;
    L.st = L.sgx * L.sgy;
  }
  size_t o = 0;
  L.header = o;
  L.sttot = o + kHdrWords * 4;
  L.rowsum = L.sttot + (size_t)(L.tb ? L.st + 1 : 0) * 4;
  L.reset_words = kHdrWords + (L.tb ? L.st + 1 + L.gy : 0);
  o = align256(o + (size_t)L.reset_words * 4);
  L.counts = o; o = align256(o + (size_t)n * 4);
  L.rects = o; o = align256(o + (size_t)n * 8);
  L.dqw = o; o = align256(o + (size_t)((n + 63) / 64) * 4);
  L.dqs = o; o = align256(o + (size_t)n * 4);
  L.keys0 = L.keys1 = L.vals0 = L.vals1 = L.hist = L.totals = L.dqpart = L.partials = 0;
  L.stoff = L.boff = L.bpart = L.clist = L.toff = L.toffx = L.tkeys = 0;
  L.nblk_cap = 0;
  if (L.tb) {
    int nb = (n + kTbMinChunk - 1) / kTbMinChunk;
    nb = nb < 1 ? 1 : (nb > kTbMaxBlocks ? kTbMaxBlocks : nb);
    L.nb = nb;
    L.chunk = (n + nb - 1) / nb;
    L.nscan = 0;
    L.stoff = o; o = align256(o + (size_t)(L.st + 1) * 4);
    L.boff = o; o = align256(o + (size_t)nb * L.st * 4);
    L.bpart = o; o = align256(o + (size_t)nb * 4);
    L.clist = o; o = align256(o + (size_t)L.cap * 16);
    L.toff = o; o = align256(o + (size_t)(L.tiles + 1) * 4);
    L.toffx = o; o = align256(o + (size_t)(L.tiles + 1) * 4);
    L.tkeys = o; o = align256(o + (size_t)L.cap * 8);
  } else {
    L.nb = 0;
    L.chunk = 0;
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

// Clears the first `words` words of a binning workspace (L.reset_words).
int bin_reset_header(void* ws, int words, hipStream_t s);


__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_xor(v, o);
    v = t > v ? t : v;
  }
  return v;
}

}  // namespace rgp
