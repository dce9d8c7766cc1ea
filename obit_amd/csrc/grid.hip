// grid.hip -- convolutional uv gridding for gfx950.
//
// Replaces Obit's CUDA gridder (reference src/ObitCUDAGrid.cu:90-920) behind
// the same C entry points (ObitCUDAGrid.h) plus resident extensions
// (ObitHIPVis.h).  Per-sample arithmetic restates Obit's CPU gridder
// (src/ObitThreadGrid.c:1263-1387 fast_prep_grid, :2045-2177 fast_rot,
// :2002-2027 fast_grid): cell indices and GCF taps are bit-identical to the
// CPU's; the facet phase shift is formed in FP64 turns with an FP32
// quadrant-reduced series, so weighted sample values agree with the CPU's
// within a few FP32 ulps (prep_vals).
//
// Accumulation: every facet-plane grid is cut into TU x TV-cell tiles of
// base cells (the first column and row of a sample's 7x7 footprint).  A
// binning pass lists, for each tile, the (row, plane) tracks whose base
// cells can fall in it (a track's base cells are bounded by the samples at
// the plane's lowest and highest frequency, since cells are monotonic in
// frequency).  Lists are cut into work items of at most kChunk tracks; a
// wave takes an item, accumulates the samples whose base cell is in the
// tile into a private LDS tile that carries an 8-cell halo to the right and
// a 6-cell halo below, so every tap of every kept sample lands in it
// (lanes = the cells of a 16x8 window round the footprint, see
// k_grid_taps), and adds the tile into the
// grid in HBM: with plain loads/stores for the cells no other tile reaches
// when the item is the tile's whole list, with device atomics otherwise.
#include "common.h"
#include "ObitCUDAGrid.h"
#include "host.h"

#include <hipfft/hipfft.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>
#include "ObitCUDAUtil.h"
#include "ObitHIPVis.h"
#include "phasor.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace obit {
namespace grid {

constexpr int TU = 64;        // tile width in grid columns (complex cells)
// tile height in grid rows: the LDS of 3 four-wave blocks per CU (12 waves;
// measured on workload 4: 50.0 ms a pass, against 58.3 at 20 rows and 58.9
// at 12 with 8 waves, 59.9 at 6 rows)
constexpr int TV = 8;
constexpr int kGcf = 708;     // tabulated GCF, CPU gridding order
constexpr int kGcf8 = 101 * 8; // the same, rows padded to 8 on the device


struct FacetDev {
  float *grid;
  int sizeGrid;  // floats per plane
  int nx, ny, lrow, hw;
  float uscale, vscale;
  float shift[3];
  double sturn[3];       // shift / 2 pi: the phase shift in turns per wavelength
  float rot[9];
  float guardu, guardv;  // wavelengths, CPU arithmetic
  float maxBL2, minBL2, bmTaper;
  int doBeam, doTape;
  int ntu, ntv;
  long long listBase;  // first list of this facet (plane 0, tile 0)
};

struct Common {
  int nrparm, lenvis, nchan, nplane, convWidth, convNperCell;
  const float *freq;    // FP32 fscale per channel
  const float *sig1, *sig2, *sig3;
  int hasSigma;
  const int *pchan;     // channels of each plane [nplane][maxpc]
  const int *pcount;    // [nplane]
  const float *pfmin, *pfmax;  // plane frequency extremes
  int maxpc;
  const float *pfreq;   // FP32 frequency factor of each plane's channels [nplane][maxpc]
  const int *pmono;     // [nplane]: 1 if pfreq is monotone along the plane's channel list
  const float *gcf;     // [101][8]: row it = taps 0..6 of fractional offset it, then 0
};

// ---------------- device: CPU-exact sample arithmetic ----------------

__device__ __forceinline__ int lround_obit(float in) {
#pragma clang fp contract(off)
  if (in >= 0.0f) return (int)((double)in + 0.5);
  return (int)((double)in - 0.5);
}

struct Sample {
  int iu, iv, itu, itv;  // first column, first row, GCF blocks
  float vr, vi;          // weighted visibility; (0,0) means "not gridded"
};

// Rotated uvw of one row for a facet (float, left-to-right, no FMA).
__device__ __forceinline__ void rotate3(const FacetDev &f, float u, float v, float w, float &uu,
                                        float &vv, float &ww) {
#pragma clang fp contract(off)
  uu = u * f.rot[0] + v * f.rot[1] + w * f.rot[2];
  vv = u * f.rot[3] + v * f.rot[4] + w * f.rot[5];
  ww = u * f.rot[6] + v * f.rot[7] + w * f.rot[8];
}

__device__ __forceinline__ void rotate_uvw(const FacetDev &f, const float *row, float &uu,
                                           float &vv, float &ww) {
  rotate3(f, row[0], row[1], row[2], uu, vv, ww);
}

// Cells of a channel at FP32 frequency factor freq (no validity test):
// used by the binning pass to bound a track's footprint.
__device__ __forceinline__ void cells_at(const FacetDev &f, float uu, float vv, float freq,
                                         int &iu, int &ivrow) {
#pragma clang fp contract(off)
  const float phaseSign = (uu <= 0.0f) ? -1.0f : 1.0f;
  const float ff = phaseSign * freq;
  float u = uu * ff, v = vv * ff;
  u = u * f.uscale;
  v = v * f.vscale;
  iu = (int)((double)u + 0.5);
  ivrow = lround_obit(v) - f.hw + f.ny / 2;
}

// fast_prep_grid of a rotated sample: validity, taper, weighting, cells
// and GCF taps, in the CPU's FP32 operations (no contraction).
__device__ __forceinline__ Sample prep_cells(const FacetDev &f, const Common &cm, float u, float v,
                                             int ich, float vr, float vi, float vwt) {
#pragma clang fp contract(off)
  Sample s;
  float vw = vwt;
  bool want = (vw > 0.0f) && (u < f.guardu) && (fabsf(v) < f.guardv);
  const float BL2 = u * u + v * v;
  if ((f.maxBL2 > 0.0f) && want) want = f.maxBL2 > BL2;
  if ((f.minBL2 > 0.0f) && want) want = f.minBL2 < BL2;
  if (!want) {
    s.iu = f.hw; s.iv = f.ny / 2; s.itu = 0; s.itv = 0; s.vr = 0.0f; s.vi = 0.0f;
    return s;
  }
  if (f.doTape) {
    if (cm.hasSigma && cm.sig1[ich] != 0.0f) {
      const float ftemp = u * u * (cm.sig2[ich] + f.bmTaper) + v * v * (cm.sig1[ich] + f.bmTaper) +
                          u * v * (cm.sig3[ich]);
      vw *= expf(ftemp);
    } else if (fabsf(f.bmTaper) > 0.0f) {
      vw *= expf(f.bmTaper * BL2);
    }
  }
  s.vr = vr * vw;
  s.vi = vi * vw;
  u = u * f.uscale;
  v = v * f.vscale;
  const int iu = (int)((double)u + 0.5);
  const int iv = lround_obit(v);
  const float N = (float)cm.convNperCell, hN = 0.5f * N;
  const float du = (float)iu - u;
  const float dv = (float)iv - v;
  s.itu = cm.convNperCell + lround_obit(fmaf(N, du, -hN));
  s.itv = cm.convNperCell + lround_obit(fmaf(N, dv, -hN));
  if (s.itu < 0 || s.itu > cm.convNperCell || s.itv < 0 || s.itv > cm.convNperCell) {
    // only for non-positive frequencies
    s.iu = f.hw; s.iv = f.ny / 2; s.itu = 0; s.itv = 0; s.vr = 0.0f; s.vi = 0.0f;
    return s;
  }
  s.iu = iu;
  s.iv = iv - f.hw + f.ny / 2;
  return s;
}

// fast_rot + fast_prep_grid for one (row, channel) of one facet, given
// the channel's frequency factor and (re, im, weight).
//
// Cell indices and GCF taps are bit-identical to the CPU's: its FP32
// products are repeated unfused, its double-promoted roundings kept where
// they matter (iu, iv and the taps' lround of a float), and its tap
// fraction (float)(N * ((double)du - 0.5)) formed as fmaf(N, du, -N/2),
// which rounds the same exact value once (checked for every float du in
// (-0.5, 0.5]).  The facet phase shift is formed in turns in FP64 from the
// same FP32 u, v, w (the CPU forms it in radians in FP64 and rounds the
// reduced angle to FP32 for a Cephes sincos); here the reduced turn is
// rounded to FP32 and goes through a quadrant reduction and Taylor
// series, so the weighted value differs from the CPU's by a few FP32
// ulps of the phase, within both evaluations' own error.
__device__ __forceinline__ Sample prep_vals(const FacetDev &f, const Common &cm, float uu, float vv,
                                            float ww, int ich, float fsc, float vre, float vim,
                                            float vwt) {
  float u, v, w;
  {
#pragma clang fp contract(off)
    const float phaseSign = (uu <= 0.0f) ? -1.0f : 1.0f;
    const float freqFact = phaseSign * fsc;
    u = uu * freqFact; v = vv * freqFact; w = ww * freqFact;
  }
  // fast_rot: phase shift of the facet
  float vr, vi;
  if (f.doBeam) {
    vr = 1.0f; vi = 0.0f;
  } else {
    const double t = (double)u * f.sturn[0] + (double)v * f.sturn[1] + (double)w * f.sturn[2];
    float sn, cs;
    sincos_turns_poly((float)(t - rint(t)), &sn, &cs);
    const float vvi = (uu <= 0.0f) ? -vim : vim;
    vr = vre * cs - vvi * sn;
    vi = vre * sn + vvi * cs;
  }
  return prep_cells(f, cm, u, v, ich, vr, vi, vwt);
}

__device__ Sample prep_sample(const FacetDev &f, const Common &cm, const float *row, float uu,
                              float vv, float ww, int ich) {
  const float *vis = row + cm.nrparm + ich * 3;
  return prep_vals(f, cm, uu, vv, ww, ich, cm.freq[ich], vis[0], vis[1], vis[2]);
}

// ---------------- binning ----------------

// Clear two int arrays with a kernel (kept out of memset nodes so the
// sequence stays capturable in a hipGraph).
__global__ void k_zero2(int *a, int *b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) { a[i] = 0; b[i] = 0; }
}

// Order-independent checksum of the rows' u, v, w bit patterns: the track
// lists of a pass may be reused only for rows with the same uvw.
__global__ void k_uvw_hash(const float *vis, long long nrow, int lenvis, unsigned long long *out) {
  unsigned long long h = 0;
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < nrow;
       r += (long long)gridDim.x * blockDim.x) {
    const float *row = vis + r * (long long)lenvis;
    const unsigned long long a = __float_as_uint(row[0]), b = __float_as_uint(row[1]),
                             c = __float_as_uint(row[2]);
    unsigned long long x = (a * 0x9E3779B97F4A7C15ULL) ^ (b * 0xC2B2AE3D27D4EB4FULL) ^
                           (c * 0x165667B19E3779F9ULL) ^ ((unsigned long long)r * 0xD6E8FEB86659FD93ULL);
    x ^= x >> 29;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 32;
    h += x;
  }
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
  if ((threadIdx.x & 63) == 0 && h) atomicAdd(out, h);
}

// Every (plane, tile of base cells) a track's samples have base cells in,
// with the channels that do and the window block where they enter:
// fn(L, wkey, lo, hi) for channels lo <= i < hi of the plane's channel list.
// A track's base cells are bounded by the cells at the plane's lowest and
// highest frequency (cells are monotonic in frequency), which gives the
// candidate tiles.  Where the plane's frequencies are monotonic along its
// channel list, so are the cells (the same FP32 arithmetic as prep_cells,
// each operation monotonic), and a binary search over the list finds
// exactly the channels whose base cell lies in the tile; a tile no channel
// lands in is left out.  Other planes give every channel to every
// candidate tile.  wkey is the 8x2 block of the tile holding the cell of
// the first channel given (clamped into the tile; block row * 8 + block
// column, < 80).
template <typename Fn>
__device__ __forceinline__ void track_tiles(const FacetDev &f, const Common &cm, float uu, float vv,
                                            Fn fn) {
  const int ncol = f.lrow / 2;
  for (int p = 0; p < cm.nplane; ++p) {
    const int npc = cm.pcount[p];
    if (npc <= 0) continue;
    int iu0, iv0, iu1, iv1;
    cells_at(f, uu, vv, cm.pfmin[p], iu0, iv0);
    cells_at(f, uu, vv, cm.pfmax[p], iu1, iv1);
    int cA = min(iu0, iu1), cB = max(iu0, iu1);
    int rA = min(iv0, iv1), rB = max(iv0, iv1);
    cA = max(cA, 0); rA = max(rA, 0);
    cB = min(cB, ncol - 1); rB = min(rB, f.ny - 1);
    if (cA > cB || rA > rB) continue;
    const float *pf = cm.pfreq + (size_t)p * cm.maxpc;
    const bool mono = cm.pmono[p] != 0;
    // cells of the first and last channel of the list: the directions in
    // which the cells move along it
    int iuF, ivF, iuL, ivL;
    cells_at(f, uu, vv, pf[0], iuF, ivF);
    cells_at(f, uu, vv, pf[npc - 1], iuL, ivL);
    const int tu0 = cA / TU, tu1 = cB / TU, tv0 = rA / TV, tv1 = rB / TV;
    const long long base = f.listBase + (long long)p * f.ntu * f.ntv;
    for (int tv = tv0; tv <= tv1; ++tv)
      for (int tu = tu0; tu <= tu1; ++tu) {
        const int C0 = tu * TU, C1 = C0 + TU, R0 = tv * TV, R1 = R0 + TV;
        int lo = 0, hi = npc;
        if (mono) {
          // first index in [a, b) where pred holds (pred false, then true)
          auto first = [&](auto pred) {
            int a = 0, b = npc;
            while (a < b) {
              const int m = (a + b) >> 1;
              int iu, iv;
              cells_at(f, uu, vv, pf[m], iu, iv);
              if (pred(iu, iv)) b = m; else a = m + 1;
            }
            return a;
          };
          if (min(iuF, iuL) < C0 || max(iuF, iuL) >= C1) {
            if (iuF <= iuL) {
              lo = max(lo, first([&](int iu, int) { return iu >= C0; }));
              hi = min(hi, first([&](int iu, int) { return iu >= C1; }));
            } else {
              lo = max(lo, first([&](int iu, int) { return iu < C1; }));
              hi = min(hi, first([&](int iu, int) { return iu < C0; }));
            }
          }
          if (min(ivF, ivL) < R0 || max(ivF, ivL) >= R1) {
            if (ivF <= ivL) {
              lo = max(lo, first([&](int, int iv) { return iv >= R0; }));
              hi = min(hi, first([&](int, int iv) { return iv >= R1; }));
            } else {
              lo = max(lo, first([&](int, int iv) { return iv < R1; }));
              hi = min(hi, first([&](int, int iv) { return iv < R0; }));
            }
          }
          if (lo >= hi) continue;
        }
        int iue, ive;
        cells_at(f, uu, vv, pf[lo], iue, ive);
        const int cu = min(max(iue - C0, 0), TU - 1);
        const int cv = min(max(ive - R0, 0), TV - 1);
        fn(base + (long long)tv * f.ntu + tu, (unsigned)((cv >> 1) * (TU / 8) + (cu >> 3)), lo, hi);
      }
  }
}

// Binning, pass 1: the number of (plane, tile) entries of every track of
// every facet, one thread per (row, facet), facet fastest.
__global__ void k_bin_count(const float *vis, long long nrow, const FacetDev *facets, int nf, Common cm,
                            unsigned *cnt) {
  const long long total = nrow * (long long)nf;
  for (long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x; it < total;
       it += (long long)gridDim.x * blockDim.x) {
    const int fi = (int)(it % nf);
    const long long r = it / nf;
    const FacetDev &f = facets[fi];
    float uu, vv, ww;
    rotate_uvw(f, vis + r * (long long)cm.lenvis, uu, vv, ww);
    unsigned n = 0;
    track_tiles(f, cm, uu, vv, [&](long long, unsigned, int, int) { ++n; });
    cnt[it] = n;
  }
}

// Binning, pass 2: every entry as a sort key (list << 7 | window block)
// and its (row, lo | hi << 16) channel range, written at the thread's
// scanned offset.  Sorting the pairs by key gives each tile's list ordered
// by the window its tracks enter first, then by row (the sort is stable
// and the emission ordered).
// The entries of a block's 256 (row, facet) tracks are one contiguous
// range of the output ([off[it0], off[it1])): where it fits, the block
// stages it in LDS and writes it out whole (the tracks' own runs are short
// and scattered); a larger range is written run by run.
constexpr int kEmitCap = 4096;  // entries a block stages (48 KB)

__global__ __launch_bounds__(256) void k_bin_emit(const float *vis, long long nrow, const FacetDev *facets,
                                                  int nf, Common cm, const unsigned long long *off,
                                                  unsigned *keys, int2 *rows) {
  __shared__ unsigned s_key[kEmitCap];
  __shared__ int2 s_row[kEmitCap];
  const long long total = nrow * (long long)nf;
  const long long nchunk = (total + 255) / 256;
  for (long long ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
    const long long it0 = ch * 256, it1 = min(total, it0 + 256);
    const unsigned long long base = off[it0], span = off[it1] - base;
    const bool stage = span <= (unsigned long long)kEmitCap;  // the same for the whole block
    const long long it = it0 + threadIdx.x;
    if (it < total) {
      const int fi = (int)(it % nf);
      const long long r = it / nf;
      const FacetDev &f = facets[fi];
      float uu, vv, ww;
      rotate_uvw(f, vis + r * (long long)cm.lenvis, uu, vv, ww);
      unsigned long long o = off[it];
      track_tiles(f, cm, uu, vv, [&](long long L, unsigned wkey, int lo, int hi) {
        const unsigned k = ((unsigned)L << 7) | wkey;
        const int2 v = make_int2((int)r, lo | (hi << 16));
        if (stage) {
          s_key[o - base] = k;
          s_row[o - base] = v;
        } else {
          keys[o] = k;
          rows[o] = v;
        }
        ++o;
      });
    }
    if (stage) {
      __syncthreads();
      for (int j = threadIdx.x; j < (int)span; j += 256) {
        keys[base + j] = s_key[j];
        rows[base + j] = s_row[j];
      }
      __syncthreads();
    }
  }
}

// List extents from the sorted keys: first[L] at a list's first entry and
// last[L] one past its last (both zero for empty lists).
__global__ void k_list_bounds(const unsigned *keys, long long n, long long *first, long long *last) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const unsigned L = keys[i] >> 7;
    if (i == 0 || (keys[i - 1] >> 7) != L) first[L] = i;
    if (i == n - 1 || (keys[i + 1] >> 7) != L) last[L] = i + 1;
  }
}

__global__ void k_list_counts(const long long *first, const long long *last, long long nl, int *counts) {
  for (long long L = (long long)blockIdx.x * blockDim.x + threadIdx.x; L < nl;
       L += (long long)gridDim.x * blockDim.x)
    counts[L] = (int)(last[L] - first[L]);
}

__global__ void k_zero64(long long *a, long long *b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) { a[i] = 0; b[i] = 0; }
}

// ---------------- tile gridding ----------------

constexpr int kW = 7;          // GCF support in cells (checked on the host)
constexpr int kChunk = 128;    // tracks per work item (swept 64..2048 on workload 4: 128 best)
constexpr int kWaves = 4;      // waves per workgroup, each with a private tile
// A sample's footprint is accumulated through a 16 (u) x 8 (v) window of
// cells whose origin is its base cell rounded down to a multiple of 8 in u
// and of 2 in v, so a window serves the 8x2 base cells of an aligned
// block.  Tile row stride in floats 12 mod 16 and the im plane 32 mod 64
// floats after the re plane: the 64 (row, column, plane) cells one
// hand-over instruction touches fall on distinct banks.
constexpr int kWinU = 16, kWinV = 8;              // window columns and rows
constexpr int kBlkU = 8, kBlkV = 2;               // base cells a window serves
constexpr int kS = TU + 12;
constexpr int kTR = TV + kWinV - kBlkV;            // tile rows with the halo below
constexpr int kTC = TU + kWinU - kBlkU;            // tile columns with the halo to the right
constexpr int kPlane = ((kTR * kS + 31) / 64) * 64 + 32;  // floats per re / im plane
constexpr int kHalf = 64;      // kept samples staged at a time (one pass)
constexpr int kGrp = kHalf / 4;  // 4-sample groups of a pass (K of one matrix step)
constexpr int kGcfRow = 24;    // padded GCF row: 7 zeros, 7 taps, 10 zeros
static_assert(kBlkU + kW - 1 <= kWinU && kBlkV + kW - 1 <= kWinV, "the window holds every footprint");
static_assert(TU % kBlkU == 0 && TV % kBlkV == 0, "windows are aligned within the tile");
static_assert(kS >= kTC && (4 * kS) % 64 == 48, "tile stride: room for the halo, 4 rows = 48 banks");
static_assert(kPlane % 64 == 32 && kPlane >= kTR * kS, "im plane offset must be 32 mod 64");
static_assert(kPlane % 4 == 0, "planes are cleared as float4");

// a in the lanes whose bit is set in the wave-uniform mask m, else 0: one
// select on a scalar lane mask.  The operand column of lane l is l / 16,
// so samples k0 <= k < k1 of a group are lanes 16 k0 .. 16 k1 - 1.  The
// result feeds a matrix step, which may read a VGPR only 2 wait states
// after a VALU instruction wrote it; the compiler does not see into the
// asm, so the asm waits them itself.
__device__ __forceinline__ float keep_lanes(float a, unsigned long long m) {
  float r;
  asm("v_cndmask_b32_e64 %0, 0, %1, %2\n\ts_nop 1" : "=v"(r) : "v"(a), "s"(m));
  return r;
}

// Work items are (tile list, chunk of <= kChunk tracks).  Each wave takes
// its own item and owns a private (TV + 6) x (TU + 8) tile (separate re /
// im planes) in LDS.
//
// The item's samples, each entry's channels whose base cells the binning
// placed in this tile, are visited 64 at a time, lanes = samples in entry
// order, channels in list order; the rows for the next 64 are fetched
// while the current 64 are gridded.  Each sample goes through
// the CPU gridder's arithmetic (prep_vals); those whose base cell lies in
// the tile are kept, in order, a 64-bit mask marks the kept
// samples whose window differs from the previous one's, and they pass
// through a per-wave stage as four scalars: where the sample's u and v
// taps start in the padded GCF table (its GCF rows less its base cell's
// offsets in the window) and the weighted visibility.  While the window
// stays put its 16x8 complex sums are a rank-k update, run on the matrix
// cores 4 samples at a time: D[16][16] += A[16][4] B[4][16] with A's rows
// the re and im parts of the v-weight x visibility of the window's 8 rows,
// B's columns the u-weights of its 16 columns and K the samples.  Each
// lane gathers its own A and B element from the GCF table, whose rows are
// padded with zeros so that any window column or row reads its tap or a
// zero with no test; a group of 4 that straddles a move is split by
// zeroing A outside each segment.  Four waves share a block and its table.  The accumulator holds the window's running sums
// and is handed to the tile when the mask says the window moves.  A
// hand-over is a read-modify-write whose write is deferred to the next
// hand-over, so no step waits on a fresh LDS read (LDS float atomics
// measured 3x slower); all of one hand-over's writes are issued before any
// of the next one's reads, which keeps successive overlapping windows
// exact.
// Cells that only this tile reaches are added into the
// grid with plain stores when the item is the tile's whole list; the
// halo, and every cell of an item of a longer list, add with device
// atomics.
__global__ __launch_bounds__(64 * kWaves) void k_grid_taps(const float *vis, const FacetDev *facets,
                                                           const int *listFacet, const int2 *items,
                                                           long long nitems, Common cm,
                                                           const int *counts, const long long *offsets,
                                                           const int2 *lists,
                                                           unsigned long long *stats) {
  __shared__ float4 s_tile4[kWaves][2 * kPlane / 4];  // re plane, then im plane
  __shared__ int2 s_si[kWaves][kHalf];      // staged sample: u and v tap-row offsets in s_gcfp
  __shared__ float2 s_sv[kWaves][kHalf];    // staged sample: weighted visibility
  __shared__ int s_cell[kWaves][64];
  __shared__ int s_flag[kWaves][64];        // entry starts within a window of slots
  // padded GCF rows: 7 zeros, the 7 taps of fractional offset it, 10 zeros,
  // so every window column or row reads its tap (or a zero) with no test
  __shared__ float s_gcfp[101 * kGcfRow];
  for (int i = threadIdx.x; i < 101 * kGcfRow; i += blockDim.x) {
    const int it = i / kGcfRow, j = i - it * kGcfRow - 7;
    s_gcfp[i] = (j >= 0 && j < kW) ? cm.gcf[8 * it + j] : 0.0f;
  }
  __syncthreads();
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // matrix-step roles of this lane: operand column k (a sample of the
  // group) and row / column m; accumulator rows 4k..4k+3 of column m
  typedef float f4 __attribute__((ext_vector_type(4)));
  const int mk = lane >> 4, mm = lane & 15;
  const int laneOff = (mk >> 1) * kPlane + 4 * (mk & 1) * kS + mm;
  const float *gcf = s_gcfp;
  const int mm7 = mm & 7;
  float *re = reinterpret_cast<float *>(s_tile4[wv]);
  float *im = re + kPlane;
  int2 *si = s_si[wv];
  float2 *sv = s_sv[wv];
  // this lane's part of a staged visibility: sample mk of a group, real
  // part for window rows (mm < 8), imaginary for the rest
  const float *svl = reinterpret_cast<const float *>(sv) + 2 * mk + (mm >= 8 ? 1 : 0);
  int *scell = s_cell[wv];
  int *flag = s_flag[wv];
  for (long long w = (long long)blockIdx.x * kWaves + wv; w < nitems;
       w += (long long)gridDim.x * kWaves) {
    const int2 item = items[w];
    const long long L = __builtin_amdgcn_readfirstlane(item.x);
    const int n = counts[L];
    const int e0 = __builtin_amdgcn_readfirstlane(item.y) * kChunk;
    const int e1 = min(n, e0 + kChunk);
    const bool multi = n > kChunk;
    const int fi = listFacet[L];
    // a register copy: the wave fences below would otherwise make every
    // pass reload the descriptor from memory and wait on the row prefetch
    const FacetDev f = facets[fi];
    const long long rel = L - f.listBase;
    const int per = f.ntu * f.ntv;
    const int p = (int)(rel / per);
    const int t = (int)(rel - (long long)p * per);
    const int tv = t / f.ntu, tu = t - tv * f.ntu;
    const int col0 = tu * TU, row0 = tv * TV;
    for (int i = lane; i < 2 * kPlane / 4; i += 64) s_tile4[wv][i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int2 *lst = lists + offsets[L];
    const int npc = cm.pcount[p];
    if (npc <= 0) continue;  // planes without channels have no tracks
    const int *pch = cm.pchan + (size_t)p * cm.maxpc;
    unsigned nStaged = 0, nMoves = 0;  // walk statistics of this item
    // the sums start at zero for window 0, and the first deferred write
    // stores zero into cells that are still zero
    int cur = 0;                     // tile index of the window origin of the running sums
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};  // running sums: 4 window rows of column mm
    int pend = laneOff;              // tile index of the deferred write (row 0 of 4)
    f4 pv = {0.0f, 0.0f, 0.0f, 0.0f};   // tile values read for it
    f4 pa = {0.0f, 0.0f, 0.0f, 0.0f};   // sums to add to them
    auto handover = [&]() {
#pragma unroll
      for (int r = 0; r < 4; ++r) re[pend + r * kS] = pv[r] + pa[r];
      pend = cur + laneOff;
#pragma unroll
      for (int r = 0; r < 4; ++r) pv[r] = re[pend + r * kS];
      pa = acc;
    };
    // The item's entries (at most kChunk = 128: lane j holds entries j and
    // 64 + j): row, first channel lo (in the plane's channel list) and
    // channel count, and where its slots start, P (an exclusive scan of the
    // counts).  Slot s of entry k is channel lo_k + s - P_k of the list;
    // pass q takes slots 64 q .. 64 q + 63, lanes in order.
    const int ne = e1 - e0;
    int rowA = 0, rowB = 0, cntA = 0, cntB = 0, loA = 0, loB = 0;
    if (lane < ne) {
      const int2 x = lst[e0 + lane];
      rowA = x.x; loA = x.y & 0xFFFF; cntA = (x.y >> 16) - loA;
    }
    if (64 + lane < ne) {
      const int2 x = lst[e0 + 64 + lane];
      rowB = x.x; loB = x.y & 0xFFFF; cntB = (x.y >> 16) - loB;
    }
    int incA = cntA, incB = cntB;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int ya = __shfl_up(incA, o), yb = __shfl_up(incB, o);
      if (lane >= o) { incA += ya; incB += yb; }
    }
    incB += __builtin_amdgcn_readlane(incA, 63);
    const int nslot = __builtin_amdgcn_readlane(incB, 63);
    const int PA = incA - cntA, PB = incB - cntB;
    const int baseA = loA - PA, baseB = loB - PB;
    // Each lane's row, list index of its channel and liveness for the
    // window of slots S .. S + 63.  jA is the entry holding slot S; lane j
    // looks at entry jA + 1 + j, whose start (if in the window) it marks in
    // flag, and lane l's entry is jA plus the starts at or before S + l.
    // Slots past the item's end take its last entry (always valid) so
    // every pass issues the same loads.
    int jA = 0;
    auto map = [&](int S, int &row, int &ci, bool &live) {
      const int idx = jA + 1 + lane;
      const int pa = __shfl(PA, idx & 63), pb = __shfl(PB, idx & 63);
      const int pos = idx < ne ? (idx < 64 ? pa : pb) - S : 65;  // 1 .. ; 65: no entry
      flag[lane] = 0;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (pos < 64) flag[pos] = 1;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const unsigned long long starts = __ballot(flag[lane] != 0);
      const int k = min(jA + (int)__popcll(starts & (~0ull >> (63 - lane))), ne - 1);
      const int ra = __shfl(rowA, k & 63), rb2 = __shfl(rowB, k & 63);
      const int ba = __shfl(baseA, k & 63), bb = __shfl(baseB, k & 63);
      row = k < 64 ? ra : rb2;
      ci = min((k < 64 ? ba : bb) + S + lane, npc - 1);
      live = S + lane < nslot;
      // the entry holding slot S + 64: one more if an entry starts there
      const int c = (int)__popcll(starts);
      jA += c + (__builtin_amdgcn_readlane(pos, c) == 64 ? 1 : 0);
    };
    float xu, xv, xw, xre, xim, xwt, xfs;
    int xch;
    auto fetch = [&](int r, int ch) {
      const float *row = vis + (long long)r * cm.lenvis;
      xch = ch;
      xfs = cm.freq[ch];
      xu = row[0]; xv = row[1]; xw = row[2];
      const float *x = row + cm.nrparm + 3 * ch;
      xre = x[0]; xim = x[1]; xwt = x[2];
    };
    // the row and channel of a slot are read two passes ahead and its data
    // one pass ahead
    bool liveCur, liveNext;
    int rb, chb;
    {
      int r0, c0;
      map(0, r0, c0, liveCur);
      fetch(r0, pch[c0]);
      map(64, rb, c0, liveNext);
      chb = pch[c0];
    }
    const int npass = (nslot + 63) / 64;
    for (int pass = 0; pass < npass; ++pass) {
      const bool live = liveCur;
      const float u0 = xu, v0 = xv, w0 = xw, re0 = xre, im0 = xim, wt0 = xwt, fs0 = xfs;
      const int ch0 = xch;
      // the slot after next: its row and channel; then the next pass's data
      int rc, cic;
      bool liveC;
      map(64 * (pass + 2), rc, cic, liveC);
      const int chc = pch[cic];
      fetch(rb, chb);
      bool keep = false;
      Sample sm;
      if (live) {
        float uu, vv, ww;
        rotate3(f, u0, v0, w0, uu, vv, ww);
        sm = prep_vals(f, cm, uu, vv, ww, ch0, fs0, re0, im0, wt0);
        keep = (sm.vr != 0.0f || sm.vi != 0.0f) && sm.iu >= col0 && sm.iu < col0 + TU &&
               sm.iv >= row0 && sm.iv < row0 + TV;
      }
      liveCur = liveNext; liveNext = liveC; rb = rc; chb = chc;
      const unsigned long long mask = __ballot(keep);
      if (mask == 0) continue;
      const int nv = __popcll(mask);
      int pos = 0;  // place of this lane's sample among the kept ones
      if (keep) {
        pos = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
        scell[pos] = ((sm.iv - row0) & -kBlkV) * kS + ((sm.iu - col0) & -kBlkU);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int cellq = scell[lane];  // lane q: tile index of kept sample q's window
      const int before = scell[lane > 0 ? lane - 1 : 0];
      const int cprev = (lane == 0) ? cur : before;
      const unsigned long long moves = __ballot(lane < nv && cellq != cprev);
      nStaged += nv;
      nMoves += __popcll(moves);
      // Stage the kept samples' scalars (zeros past the last one), read the
      // pass's 16 groups back in matrix layout, gather the operands from
      // the GCF table and walk the groups.
      if (keep) {
        const int cu = sm.iu - col0, cv = sm.iv - row0;
        si[pos] = make_int2(kGcfRow * sm.itu + 7 - (cu & (kBlkU - 1)), kGcfRow * sm.itv + 7 - (cv & (kBlkV - 1)));
        sv[pos] = make_float2(sm.vr, sm.vi);
      }
      if (lane >= nv) {
        si[lane] = make_int2(0, 0);  // any tap row: the visibility is zero
        sv[lane] = make_float2(0.0f, 0.0f);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      float A[kGrp], B[kGrp];
#pragma unroll
      for (int g = 0; g < kGrp; ++g) {
        const int2 q = si[4 * g + mk];
        B[g] = gcf[q.x + mm];
        A[g] = gcf[q.y + mm7] * svl[8 * g];
      }
#pragma unroll
      for (int g = 0; g < kGrp; ++g) {
        if (4 * g >= nv) break;
        unsigned mg = (unsigned)(moves >> (4 * g)) & 15u;  // moves at the group's samples
        int s0 = 0;
        while (mg) {
          const int p1 = __builtin_ctz(mg);
          if (p1 > s0) {
            // samples s0 .. p1 - 1 of the group (p1 <= 3)
            const unsigned long long m = (~0ull << (16 * s0)) ^ (~0ull << (16 * p1));
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(keep_lanes(A[g], m), B[g], acc, 0, 0, 0);
          }
          handover();
          cur = __builtin_amdgcn_readlane(cellq, 4 * g + p1);
          acc = f4{0.0f, 0.0f, 0.0f, 0.0f};
          s0 = p1;
          mg &= mg - 1;
        }
        // samples s0 .. 3 of the group
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(keep_lanes(A[g], ~0ull << (16 * s0)), B[g], acc, 0, 0, 0);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    handover();
#pragma unroll
    for (int r = 0; r < 4; ++r) re[pend + r * kS] = pv[r] + pa[r];
    if (stats && lane == 0) {
      atomicAdd(&stats[0], (unsigned long long)nStaged);
      atomicAdd(&stats[1], (unsigned long long)nMoves);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float *g = f.grid + (size_t)p * f.sizeGrid;
    const int ncol = f.lrow / 2;
    for (int i = lane; i < kTR * kTC; i += 64) {
      const int rr = i / kTC, cc = i - rr * kTC;
      const int gr = row0 + rr, gc = col0 + cc;
      if (gr >= f.ny || gc >= ncol) continue;
      const float vr = re[rr * kS + cc], vi = im[rr * kS + cc];
      if (vr == 0.0f && vi == 0.0f) continue;
      float *q = g + (size_t)gr * f.lrow + 2 * gc;
      // the windows of the tiles to the left and above reach kWin - kBlk
      // cells into this one
      const bool own = !multi && rr >= kWinV - kBlkV && rr < TV && cc >= kWinU - kBlkU && cc < TU;
      if (own) {
        q[0] += vr;
        q[1] += vi;
      } else {
        unsafeAtomicAdd(q, vr);
        unsafeAtomicAdd(q + 1, vi);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// ThreadFlip (src/ObitThreadGrid.c:998-1045) as a parallel update: thread
// (facet plane, vrow in 1..ny/2, iu in 0..hw).  Each thread reads the
// cells at hw-iu (never written) and hw+iu (written only by itself).
__global__ void k_flip(const FacetDev *facets, int nf, int nplane, int hwmax) {
  const int fi = blockIdx.z / nplane, p = blockIdx.z - (blockIdx.z / nplane) * nplane;
  if (fi >= nf) return;
  const FacetDev &f = facets[fi];
  const int ny2 = f.ny / 2;
  const int hw = f.hw;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int vrow = 1 + idx / (hwmax + 1);
  const int iu = idx - (idx / (hwmax + 1)) * (hwmax + 1);
  if (vrow > ny2 || iu > hw) return;
  float *g = f.grid + (size_t)p * f.sizeGrid;
  const int vc = f.ny - vrow;
  float *gxo = g + (size_t)vrow * f.lrow + 2 * (hw + iu);
  float *gxi = g + (size_t)vrow * f.lrow + 2 * (hw - iu);
  float *gco = g + (size_t)vc * f.lrow + 2 * (hw + iu);
  float *gci = g + (size_t)vc * f.lrow + 2 * (hw - iu);
  const float cji0 = gci[0], cji1 = gci[1], cjo0 = gco[0], cjo1 = gco[1];
  const float xxi0 = gxi[0], xxi1 = gxi[1], xxo0 = gxo[0], xxo1 = gxo[1];
  gxo[0] = xxo0 + cji0;
  gxo[1] = xxo1 - cji1;
  gco[0] = cjo0 + xxi0;
  gco[1] = cjo1 - xxi1;
}

// Trim hw columns and swap v halves: out[p][y][x] = grid[p][(y+ny/2)%ny][hw+x].
__global__ void k_trim_centre(const FacetDev *facets, int fi, int nplane, float *out) {
  const FacetDev &f = facets[fi];
  const int nxo = f.nx / 2 + 1;
  const long long total = (long long)nplane * f.ny * nxo;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(i % nxo);
    const long long yp = i / nxo;
    const int y = (int)(yp % f.ny);
    const int p = (int)(yp / f.ny);
    const int ys = (y + f.ny / 2) % f.ny;
    const float *src = f.grid + (size_t)p * f.sizeGrid + (size_t)ys * f.lrow + 2 * (f.hw + x);
    out[2 * i] = src[0];
    out[2 * i + 1] = src[1];
  }
}

// ObitFArray2DCenter's quadrant swap and ObitFArrayMulColRow's correction
// (src/ObitFArray.c:3116-3123) of a C2R of the planes as they lie in the
// gridder (v halves not swapped): a cyclic shift of the v axis by ny/2 before the transform
// is the factor (-1)^y after it, so pixel (x, y) of the centred image is
// (-1)^ys out[ys][xs] with xs = (x + nx/2) % nx, ys = (y + ny/2) % ny,
// times xCorr[y]/norm and yCorr[x].  One thread per 4 pixels of a row
// (nx a multiple of 8, so the 4 source pixels are contiguous).
__global__ void k_center_correct4(const float *fft, int nx, int ny, const float *xc, const float *yc,
                                  const float *invNorm, float *img) {
  const int x0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  const int y = blockIdx.y, p = blockIdx.z;
  if (x0 >= nx) return;
  int ys = y + ny / 2;
  if (ys >= ny) ys -= ny;
  int xs = x0 + nx / 2;
  if (xs >= nx) xs -= nx;
  const float4 v = *reinterpret_cast<const float4 *>(fft + ((size_t)p * ny + ys) * nx + xs);
  const float4 c = *reinterpret_cast<const float4 *>(yc + x0);
  const float r = ((ys & 1) ? -xc[y] : xc[y]) * invNorm[p];
  *reinterpret_cast<float4 *>(img + ((size_t)p * ny + y) * nx + x0) =
      make_float4(v.x * r * c.x, v.y * r * c.y, v.z * r * c.z, v.w * r * c.w);
}

// ---- the v transform of FFT2Image: columns of the padded half planes ----

// V[k] = sum_m v[m] exp(+2 pi i m k / R) for R = 4, 8, 16, in natural
// order: radix-2 stages on a bit-reversed copy, the stage twiddles exact
// FP32 constants (W16^k).
template <int R>
__device__ __forceinline__ void dft_small(float2 (&v)[R]) {
  constexpr float kc[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508979f,
                            0.0f, -0.38268343236508979f, -0.70710678118654752f, -0.92387953251128674f,
                            -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508979f,
                            0.0f, 0.38268343236508979f, 0.70710678118654752f, 0.92387953251128674f};
  constexpr int kBits = R == 4 ? 2 : R == 8 ? 3 : 4;
  float2 a[R];
#pragma unroll
  for (int m = 0; m < R; ++m) {
    int r = 0;
#pragma unroll
    for (int b = 0; b < kBits; ++b) r |= ((m >> b) & 1) << (kBits - 1 - b);
    a[r] = v[m];
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
    const int half = len / 2;
#pragma unroll
    for (int i = 0; i < R; i += len) {
#pragma unroll
      for (int j = 0; j < half; ++j) {
        const int e = j * (16 / len);          // W_len^j = W16^e
        const float2 x = a[i + j], y = a[i + j + half];
        float2 t;
        if (e == 0) {
          t = y;
        } else if (e == 4) {                   // +i
          t = make_float2(-y.y, y.x);
        } else {
          const float wc = kc[e], ws = kc[(e + 12) & 15];   // cos, sin(2 pi e / 16)
          t = make_float2(y.x * wc - y.y * ws, y.x * ws + y.y * wc);
        }
        a[i + j] = make_float2(x.x + t.x, x.y + t.y);
        a[i + j + half] = make_float2(x.x - t.x, x.y - t.y);
      }
    }
  }
#pragma unroll
  for (int m = 0; m < R; ++m) v[m] = a[m];
}

// LDS layout of the transforms: point n of a column at lidx(n) = n + n / 8
// (a free slot after every 8 points, so the stride-8 writes of a Stockham
// pass fall on distinct banks), columns kColStride<N> float2 apart (one
// more, so a row's CB columns fall on distinct banks too).
__device__ __forceinline__ int lidx(int n) { return n + (n >> 3); }
template <int N>
constexpr int kColStride = N + N / 8 + 1;

// One Stockham pass over CB columns of N points held in LDS (column c at
// lds[c * kColStride<N>], point n at lidx(n)): for every j < N / R, with Ns
// the length of the transforms already formed and k = j mod Ns, the R
// points j + r N / R times W_{Ns R}^{r k}, a radix-R DFT, written to
// (j / Ns) Ns R + k + r Ns.
// With HALF, tw holds W_N^m for m < N / 2 only (W_N^{m + N/2} = -W_N^m).
// tid is the thread's index in the block.
template <int N, int R, int CB, int NT = 256, bool HALF = false>
__device__ __forceinline__ void stockham_pass(float2 *lds, int Ns, const float2 *__restrict__ tw,
                                              int tid = (int)threadIdx.x) {
  constexpr int kItems = CB * N / R;
  constexpr int kPer = (kItems + NT - 1) / NT;
  float2 v[kPer][R];
  int dcol[kPer], dn[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int it = q * NT + tid;
    dcol[q] = -1;
    if (it < kItems) {
      const int c = it / (N / R), j = it - c * (N / R);
      const int k = j & (Ns - 1);
      const int step = N / (Ns * R);            // W_{Ns R}^{r k} = W_N^{r k step}
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = lds[c * kColStride<N> + lidx(j + r * (N / R))];
        if (r) {
          const int m = (r * k * step) & (N - 1);
          float2 w;
          if (HALF) {
            w = tw[m & (N / 2 - 1)];
            if (m & (N / 2)) w = make_float2(-w.x, -w.y);
          } else {
            w = tw[m];
          }
          x = make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
        }
        v[q][r] = x;
      }
      dft_small<R>(v[q]);
      dcol[q] = c;
      dn[q] = (j - k) * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPer; ++q)
    if (dcol[q] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) lds[dcol[q] * kColStride<N> + lidx(dn[q] + r * Ns)] = v[q][r];
    }
  __syncthreads();
}

// In-place backward (exp +i) complex transform of length N along v of
// columns of a facet's padded half planes: thread t of a block reads rows
// t + (N / R1) r, r < R1, of its columns and runs the first radix-R1 pass
// on them in registers; the two other Stockham passes (N = R1 R2 R3) go
// through LDS; the columns are written back in place.  tw[m] = exp(+2 pi
// i m / N), formed in FP64 on the host.
//
// The same transform, 8 columns at a time by persistent blocks of 512
// threads, so each row's 64-byte segment is read and written whole by one
// wave: wave w, lane L takes rows q + 256 r (q = 32 w + L mod 32) and,
// with h = L / 32, columns 2h, 2h+1 (bytes 16h..16h+15 of the segment) and
// 4+2h, 5+2h, so each 16-byte access instruction covers 32 contiguous
// bytes of 32 rows.  The work is the (plane, 8-column block) list of
// columns c0 .. c0 + 8 nfull - 1 (c0 even) of nplane planes: the blocks of
// XCD x (blockIdx mod 8; gridDim.x a multiple of 8) take its x-th eighth,
// so an XCD works through neighbouring column blocks.  A block loads its
// next column block into registers as soon as the current one's first
// pass is in LDS, so the loads run under the LDS passes and the write-back.  The half twiddle table is copied to LDS, so the passes make
// no global loads that would wait behind those.  Then the blocks take the
// column side of every plane, one plane at a time.
template <int N, int R1, int R2, int R3>
__global__ __launch_bounds__(512) void k_fft_cols8(float2 *grid, int pitchC, long long planeC, int c0, int nplane,
                                                   int nfull, int side, const float2 *__restrict__ tw) {
  constexpr int CB = 8, NT = 512;
  static_assert(R1 * R2 * R3 == N && N / R1 == 256, "radices: the first pass has one item a thread");
  __shared__ float2 lds[CB * kColStride<N>];
  __shared__ float2 twl[N / 2];
  const int t = threadIdx.x;
  for (int i = t; i < N / 2; i += NT) twl[i] = tw[i];
  const int total = nplane * nfull, span = (total + 7) / 8, nper = (int)(gridDim.x >> 3);
  const int x = (int)(blockIdx.x & 7);
  const int eEnd = min(total, (x + 1) * span);
  auto base = [&](int e) {
    const int p = e / nfull, blk = e - p * nfull;
    return grid + (long long)p * planeC + c0 + (long long)blk * CB;
  };
  float2 v[4][R1];
  // rows r0 <= r < r1 of column block e into v, for thread tl
  auto load = [&](int e, int r0, int r1, int tl) {
    const float2 *g = base(e);
    const int h = (tl >> 5) & 1, q = (tl >> 6) * 32 + (tl & 31);
#pragma unroll
    for (int r = r0; r < r1; ++r) {
      const float2 *row = g + (long long)(q + 256 * r) * pitchC;
      const float4 a = *reinterpret_cast<const float4 *>(row + 2 * h);
      const float4 b = *reinterpret_cast<const float4 *>(row + 4 + 2 * h);
      v[0][r] = make_float2(a.x, a.y);
      v[1][r] = make_float2(a.z, a.w);
      v[2][r] = make_float2(b.x, b.y);
      v[3][r] = make_float2(b.z, b.w);
    }
  };
  int e = x * span + (int)(blockIdx.x >> 3);
  if (e < eEnd) load(e, 0, R1, t);
  for (; e < eEnd; e += nper) {
    // the thread index through an empty asm, so the compiler forms the
    // addresses and twiddle indices in each iteration instead of keeping
    // them all in registers across the loop
    int tl = t;
    asm volatile("" : "+v"(tl));
    const int h = (tl >> 5) & 1, q = (tl >> 6) * 32 + (tl & 31);
    const int cid[4] = {2 * h, 2 * h + 1, 4 + 2 * h, 5 + 2 * h};
    // the first pass (transforms of length 1 so far: no twiddles)
#pragma unroll
    for (int c = 0; c < 4; ++c) dft_small<R1>(v[c]);
    __syncthreads();  // the last column block's write-back has read the LDS
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int r = 0; r < R1; ++r) lds[cid[c] * kColStride<N> + lidx(q * R1 + r)] = v[c][r];
    // the next column block's rows, in flight under the passes
    if (e + nper < eEnd) load(e + nper, 0, R1, tl);
    __syncthreads();
    stockham_pass<N, R2, CB, NT, true>(lds, R1, twl, tl);
    stockham_pass<N, R3, CB, NT, true>(lds, R1 * R2, twl, tl);
    float2 *g = base(e);
#pragma unroll
    for (int i = 0; i < N / 256; ++i) {
      const int row = q + 256 * i;
      float2 *dst = g + (long long)row * pitchC;
      const int l = lidx(row);
      const float2 x0 = lds[cid[0] * kColStride<N> + l], x1 = lds[cid[1] * kColStride<N> + l];
      const float2 x2 = lds[cid[2] * kColStride<N> + l], x3 = lds[cid[3] * kColStride<N> + l];
      *reinterpret_cast<float4 *>(dst + 2 * h) = make_float4(x0.x, x0.y, x1.x, x1.y);
      *reinterpret_cast<float4 *>(dst + 4 + 2 * h) = make_float4(x2.x, x2.y, x3.x, x3.y);
    }
  }
  // column side of plane sp: threads t < 256 hold rows t + 256 r
  for (int sp = (int)blockIdx.x; sp < nplane; sp += (int)gridDim.x) {
    float2 *g = grid + (long long)sp * planeC + side;
    __syncthreads();
    if (t < 256) {
      float2 w[R1];
#pragma unroll
      for (int r = 0; r < R1; ++r) w[r] = g[(long long)(t + 256 * r) * pitchC];
      dft_small<R1>(w);
#pragma unroll
      for (int r = 0; r < R1; ++r) lds[lidx(t * R1 + r)] = w[r];
    }
    __syncthreads();
    stockham_pass<N, R2, 1, NT, true>(lds, R1, twl);
    stockham_pass<N, R3, 1, NT, true>(lds, R1 * R2, twl);
    for (int i = t; i < N; i += NT) g[(long long)i * pitchC] = lds[lidx(i)];
  }
}

// The u transform of FFT2Image fused with its centring and correction:
// a C2R of length N (= nx) of every row of the column-transformed half
// planes, as a complex transform of length M = N / 2 (the even and odd
// output samples as its real and imaginary parts):
//   Z[k] = Fe[k] + i Fo[k],  Fe = X[k] + conj X[M-k],
//   Fo = (X[k] - conj X[M-k]) exp(+2 pi i k / N),  k < M,
//   z = backward transform of Z,  x[2m] = Re z[m],  x[2m+1] = Im z[m],
// with the imaginary parts of X[0] and X[M] taken as zero (FFTW's C2R).
// Output row y of plane p is input row r = (y + ny/2) mod ny times
// (-1)^r xc[y] invNorm[p], and column x is sample (x + N/2) mod N times
// yc[x] (ObitFArray2DCenter and ObitFArrayMulColRow, as k_center_correct4
// applies them).  Block (RW rows, plane); NT threads.  twN[k] = exp(+2 pi
// i k / N) and twM = the M-point table (every other entry of twN).
template <int N, int R1, int R2, int R3, int RW = 4, int NT = 256>
__global__ __launch_bounds__(NT) void k_fft_rows_c2r(const float2 *grid, int pitchC, long long planeC, int hw,
                                                      int ny, const float2 *__restrict__ twN,
                                                      const float2 *__restrict__ twM, const float *xc,
                                                      const float *yc, const float *invNorm, float *img) {
  constexpr int M = N / 2;
  static_assert(R1 * R2 * R3 == M, "radices");
  __shared__ float2 lds[RW * kColStride<M>];
  const int t = threadIdx.x, p = blockIdx.y;
  const int r0 = blockIdx.x * RW;
  const float2 *g = grid + (long long)p * planeC + hw;
  // rows r0 .. r0+3, samples 0 .. M (row c at lds[c * kColStride<M>]):
  // every load issued before the first is waited on
  static_assert(M % NT == 0 && NT >= RW, "whole sample blocks per thread");
  constexpr int kPerRow = M / NT;
  float2 xin[RW][kPerRow];
  float2 xlast = make_float2(0.0f, 0.0f);
#pragma unroll
  for (int c = 0; c < RW; ++c)
#pragma unroll
    for (int i = 0; i < kPerRow; ++i) xin[c][i] = g[(long long)(r0 + c) * pitchC + t + NT * i];
  if (t < RW) xlast = g[(long long)(r0 + t) * pitchC + M];
#pragma unroll
  for (int c = 0; c < RW; ++c)
#pragma unroll
    for (int i = 0; i < kPerRow; ++i) lds[c * kColStride<M> + lidx(t + NT * i)] = xin[c][i];
  if (t < RW) lds[t * kColStride<M> + lidx(M)] = xlast;
  __syncthreads();
  // Z in place: the pair (k, M - k) by one thread; all of a thread's
  // loads first
  constexpr int kPairs = RW * (M / 2 + 1), kPP = (kPairs + NT - 1) / NT;
  float2 za[kPP], zb[kPP], wa[kPP], wb[kPP];
#pragma unroll
  for (int i = 0; i < kPP; ++i) {
    const int q = t + NT * i;
    const int c = q / (M / 2 + 1), k = q - c * (M / 2 + 1);
    if (q < kPairs) {
      const float2 *row = lds + c * kColStride<M>;
      za[i] = row[lidx(k)];
      zb[i] = row[lidx(M - k)];
      wa[i] = twN[k];
      wb[i] = twN[(M - k) & (N - 1)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kPP; ++i) {
    const int q = t + NT * i;
    const int c = q / (M / 2 + 1), k = q - c * (M / 2 + 1);
    if (q < kPairs) {
      float2 *row = lds + c * kColStride<M>;
      float2 a = za[i], b = zb[i];
      if (k == 0) { a.y = 0.0f; b.y = 0.0f; }
      // Z[k] from X[k], X[M-k] and exp(+2 pi i k / N)
      auto zk = [](float2 xk, float2 xm, float2 w) {
        const float2 fe = make_float2(xk.x + xm.x, xk.y - xm.y);
        const float2 d = make_float2(xk.x - xm.x, xk.y + xm.y);
        const float2 fo = make_float2(d.x * w.x - d.y * w.y, d.x * w.y + d.y * w.x);
        return make_float2(fe.x - fo.y, fe.y + fo.x);
      };
      row[lidx(k)] = zk(a, b, wa[i]);
      if (k != 0) row[lidx(M - k)] = zk(b, a, wb[i]);
    }
  }
  __syncthreads();
  stockham_pass<M, R1, RW, NT>(lds, 1, twM);
  stockham_pass<M, R2, RW, NT>(lds, R1, twM);
  stockham_pass<M, R3, RW, NT>(lds, R1 * R2, twM);
  // centre, correct, store: thread takes output column pairs
  const float inv = invNorm[p];
#pragma unroll
  for (int i = 0; i < RW * M / NT; ++i) {
    const int q = t + NT * i;
    const int c = q / M, m = q - c * M;
    const int r = r0 + c;
    int y = r + ny / 2;
    if (y >= ny) y -= ny;
    const float rs = ((r & 1) ? -xc[y] : xc[y]) * inv;
    const float2 z = lds[c * kColStride<M> + lidx(m)];
    int x = 2 * m + N / 2;
    if (x >= N) x -= N;
    const float2 yy = *reinterpret_cast<const float2 *>(yc + x);
    *reinterpret_cast<float2 *>(img + ((long long)p * ny + y) * N + x) = make_float2(z.x * rs * yy.x, z.y * rs * yy.y);
  }
}

// Beam normalisation from the column-transformed planes: the uncorrected
// image's (0, 0) sample is the sum of row 0's Hermitian extension,
// X[0] + X[M] + 2 sum_{0<k<M} Re X[k] (imaginary parts of X[0], X[M]
// dropped as the C2R drops them), times xc[ny/2] yc[nx/2]; one block a plane.
// The norm also goes straight to hostNorm (coherent pinned memory).
__global__ __launch_bounds__(256) void k_beam_norm_rows(const float2 *grid, long long planeC, int hw, int nx,
                                                        const float *xc, const float *yc, int ny, float *norm,
                                                        float *invNorm, float *hostNorm) {
  __shared__ double s[256];
  const int p = blockIdx.x, M = nx / 2;
  const float2 *row = grid + (long long)p * planeC + hw;
  double acc = 0.0;
  for (int k = threadIdx.x; k <= M; k += 256) acc += (k == 0 || k == M) ? (double)row[k].x : 2.0 * (double)row[k].x;
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float v = (float)s[0] * xc[ny / 2] * yc[nx / 2];
    norm[p] = v;
    invNorm[p] = (float)(1.0 / fmax(1.0e-20, (double)v));
    hostNorm[p] = v;
  }
}

// Beam normalisation of every plane on the device: the centre pixel of the
// centred image, (-1)^0 out[p][0][0] x xCorr[ny/2] x yCorr[nx/2]; norm and
// its reciprocal (as the reference forms it, 1 / max(1e-20, norm)).
__global__ void k_beam_norm(const float *fft, int nx, int ny, int nplane, const float *xc, const float *yc,
                            float *norm, float *invNorm, float *hostNorm) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nplane) return;
  const float v = fft[(size_t)p * ny * nx] * xc[ny / 2] * yc[nx / 2];
  norm[p] = v;
  invNorm[p] = (float)(1.0 / fmax(1.0e-20, (double)v));
  hostNorm[p] = v;
}

// Per-sample outputs of prep_sample for the parity tests.
__global__ void k_prep_debug(const float *vis, long long nrow, const FacetDev *facets, int fi,
                             Common cm, int *iu, int *iv, int *itu, int *itv, float *vout) {
  const long long total = nrow * cm.nchan;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long r = i / cm.nchan;
    const int c = (int)(i - r * cm.nchan);
    const FacetDev &f = facets[fi];
    const float *row = vis + r * (long long)cm.lenvis;
    float uu, vv, ww;
    rotate_uvw(f, row, uu, vv, ww);
    const Sample s = prep_sample(f, cm, row, uu, vv, ww, c);
    iu[i] = s.iu; iv[i] = s.iv; itu[i] = s.itu; itv[i] = s.itv;
    vout[2 * i] = s.vr; vout[2 * i + 1] = s.vi;
  }
}

// ---------------- host state ----------------

struct GpuState {
  int device = 0;
  hipStream_t stream = nullptr;
  int nCU = 256;
  int tapsPerCU = 0;        // resident k_grid_taps blocks per CU (0: not queried yet)
  std::vector<int> facets;  // facet indices on this GPU
  FacetDev *d_facets = nullptr;
  std::vector<FacetDev> h_facets;
  int *d_listFacet = nullptr;
  long long nlist = 0;
  int *d_counts = nullptr;
  int2 *d_lists = nullptr;          // per list entry: row, lo | hi << 16 (channels of the plane list)
  long long *d_offsets = nullptr;  // first entry of every list (64-bit: passes may exceed 2^31 entries)
  long long *d_lend = nullptr;     // one past the last entry of every list
  long long listCap = 0;
  unsigned *d_tcnt = nullptr;      // entries per (row, facet) track
  unsigned long long *d_toff = nullptr;  // their exclusive scan
  long long tcap = 0;
  unsigned *d_keys = nullptr, *d_keys2 = nullptr;  // sort keys, in and out
  int2 *d_rows = nullptr;          // (row, channel range) before the sort (d_lists after it)
  int2 *d_items = nullptr;
  long long itemCap = 0;
  long long *d_itemOff = nullptr;  // first work item of every list (exclusive scan of chunk counts)
  void *d_scanTmp = nullptr;       // scan scratch
  size_t scanTmpBytes = 0;
  long long *h_tot = nullptr;      // pinned: list entries and work items of the last pass
  long long nitems = 0;
  bool listsValid = false;         // the lists of the last pass, for ...
  long long listsRows = 0;         // ... a pass over as many rows ...
  unsigned long long listsHash = 0;  // ... with the same uvw (k_uvw_hash)
  unsigned long long *d_hash = nullptr;
  unsigned long long *d_walk = nullptr;  // staged samples, base-cell moves of the last pass
  // FFT to images: C2R plan for (nx, ny, nplane), complex scratch, corrections
  hipfftHandle fftPlan = 0;
  int fftNx = 0, fftNy = 0, fftNplane = 0, fftPitch = 0, fftDist = 0;
  float *h_norm = nullptr;  // pinned: plane norms / reciprocals staged for the caller
  float *d_fft = nullptr;
  size_t fftCap = 0;
  float2 *d_tw = nullptr;          // exp(+2 pi i m / twN) for the FFT kernels
  int colsPerCU = 0, colsPerCU1024 = 0;  // resident k_fft_cols8 blocks per CU
  int twN = 0;
  float *d_corr = nullptr;  // xCorr[nx], yCorr[ny], 1/norm[nplane]
  size_t corrCap = 0;
  std::vector<float> h_corr;  // xCorr, yCorr of the last (nx, ny) (GridCorrFn)
  int corrNx = 0, corrNy = 0;
  // tables
  float *d_gcf = nullptr, *d_freq = nullptr, *d_sig = nullptr, *d_pf = nullptr, *d_pfreq = nullptr;
  int *d_pmono = nullptr;
  int *d_pchan = nullptr, *d_pcount = nullptr;
  Common cm{};
  bool ready = false;
  // staging for host buffers
  float *d_vis = nullptr;
  size_t visCap = 0;
};

struct GridCtx {
  std::vector<GpuState *> gpus;
};

std::mutex g_mu;
std::unordered_map<const CUDAGridInfo *, GridCtx *> g_ctx;

GridCtx *get_ctx(const CUDAGridInfo *in, bool create) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_ctx.find(in);
  if (it != g_ctx.end()) return it->second;
  if (!create) return nullptr;
  GridCtx *c = new GridCtx();
  g_ctx[in] = c;
  return c;
}

GpuState *get_gpu(CUDAGridInfo *in, long iGPU, bool create) {
  GridCtx *c = get_ctx(in, create);
  if (!c) return nullptr;
  if ((long)c->gpus.size() <= iGPU) {
    if (!create) return nullptr;
    c->gpus.resize(iGPU + 1, nullptr);
  }
  if (!c->gpus[iGPU] && create) c->gpus[iGPU] = new GpuState();
  return c->gpus[iGPU];
}

int device_of(const CUDAGridInfo *in, long iGPU) {
  if (in->h_gridInfo && in->h_gridInfo->GPU_device_no) return in->h_gridInfo->GPU_device_no[iGPU];
  if (in->cuda_device) return (int)in->cuda_device[iGPU];
  return (int)iGPU;
}

// Build the CPU gridding-order table from the tabulated GCF the host
// passes (h_convfn = ObitUVGrid convfn, 100 points per cell):
// gcf[7*it + c] = convfn[c*100 + it], it = 0..100 (src/ObitUVGrid.c:1398-1409).
void build_gcf(const float *convfn, int W, int N, std::vector<float> &gcf) {
  gcf.assign(kGcf, 0.0f);
  for (int it = 0; it <= N; ++it)
    for (int c = 0; c < W; ++c) {
      const int src = c * N + it;
      const int dst = W * it + c;
      if (dst < kGcf && src < W * N + 1) gcf[dst] = convfn[src];
    }
}

// Upload tables and facet parameters from the host descriptors.
int refresh(CUDAGridInfo *in, long iGPU, GpuState *gs) {
  gs->listsValid = false;
  GridInfo *gi = in->h_gridInfo;
  if (!gi) return fail(kState, "h_gridInfo not allocated");
  OBIT_HIP_TRY(hipSetDevice(gs->device));
  const int nchan = gi->nchan, nplane = gi->nplane;
  const int W = gi->convWidth, N = gi->convNperCell;
  if (W != 7 || N != 100) return fail(kInvalid, "GCF must be 7 cells x 100 points (got %d x %d)", W, N);
  if (nchan <= 0 || nplane <= 0) return fail(kInvalid, "nchan/nplane must be positive");
  if (gi->lenvis < gi->nrparm + 3 * nchan || gi->nrparm < 3)
    return fail(kInvalid, "rows of %d floats cannot hold %d random parameters and %d channels",
                gi->lenvis, gi->nrparm, nchan);
  // facets on this GPU
  gs->facets.clear();
  for (long i = 0; i < in->nfacet; ++i)
    if (!in->FacetGPU || in->FacetGPU[i] == iGPU) gs->facets.push_back((int)i);
  const int nf = (int)gs->facets.size();
  gs->h_facets.assign(nf, FacetDev{});
  long long listBase = 0;
  std::vector<int> listFacet;
  for (int k = 0; k < nf; ++k) {
    const FacetInfo *fi = in->h_facetInfo[gs->facets[k]];
    FacetDev &f = gs->h_facets[k];
    f.grid = fi->d_grid;
    f.sizeGrid = fi->sizeGrid;
    f.nx = fi->nx; f.ny = fi->ny; f.hw = W / 2;
    f.lrow = 2 * (1 + fi->nx / 2 + f.hw);
    if (fi->sizeGrid < f.lrow * fi->ny) return fail(kInvalid, "facet %d: sizeGrid too small", gs->facets[k]);
    // base cells are packed as 16-bit column | 16-bit signed row
    if (fi->nx < 2 || fi->ny < 2 || fi->nx > 65536 || fi->ny > 32768)
      return fail(kInvalid, "facet %d: %d x %d grid outside 2..65536 x 2..32768", gs->facets[k], fi->nx,
                  fi->ny);
    f.uscale = fi->uscale; f.vscale = fi->vscale;
    for (int j = 0; j < 3; ++j) {
      f.shift[j] = fi->shift[j];
      f.sturn[j] = (double)fi->shift[j] / (2.0 * 3.14159265358979323846);
    }
    for (int j = 0; j < 9; ++j) f.rot[j] = fi->rotUV[j];
    // guard band in wavelengths as src/ObitThreadGrid.c:1290-1291 forms it
    // from guardband*nx (src/ObitThreadGrid.c:328-329), in float
    const volatile float gcu = gi->guardu * (float)fi->nx;
    const volatile float gcv = gi->guardv * (float)fi->ny;
    const volatile float gqu = gcu / fi->uscale;
    const volatile float gqv = gcv / fi->vscale;
    f.guardu = std::fabs((float)gqu);
    f.guardv = std::fabs((float)gqv);
    const volatile float mx = fi->maxBL * fi->maxBL, mn = fi->minBL * fi->minBL;
    f.maxBL2 = mx; f.minBL2 = mn;
    f.bmTaper = fi->bmTaper;
    f.doBeam = fi->doBeam;
    f.doTape = (std::fabs(fi->bmTaper) > 0.0f) || (gi->h_sigma1 && gi->h_sigma1[0] != 0.0f);
    f.ntu = (f.lrow / 2 + TU - 1) / TU;
    f.ntv = (f.ny + TV - 1) / TV;
    f.listBase = listBase;
    const long long nl = (long long)nplane * f.ntu * f.ntv;
    listFacet.insert(listFacet.end(), (size_t)nl, k);
    listBase += nl;
  }
  gs->nlist = listBase;
  // tables
  std::vector<float> gcf;
  build_gcf(gi->h_convfn, W, N, gcf);
  std::vector<std::vector<int>> pch(nplane);
  std::vector<float> pf(2 * nplane, 0.0f);
  for (int c = 0; c < nchan; ++c) {
    const int p = gi->h_freqPlane[c];
    if (p >= 0 && p < nplane) pch[p].push_back(c);
  }
  int maxpc = 1;
  for (int p = 0; p < nplane; ++p) maxpc = std::max(maxpc, (int)pch[p].size());
  if (maxpc > 32767) return fail(kInvalid, "%d channels in one plane: at most 32767", maxpc);
  std::vector<int> hpch((size_t)nplane * maxpc, 0), hpcount(nplane, 0), hmono(nplane, 0);
  std::vector<float> hpfreq((size_t)nplane * maxpc, 0.0f);
  for (int p = 0; p < nplane; ++p) {
    hpcount[p] = (int)pch[p].size();
    float lo = 0.0f, hi = 0.0f;
    for (size_t k = 0; k < pch[p].size(); ++k) {
      hpch[(size_t)p * maxpc + k] = pch[p][k];
      const float fr = gi->h_freqArr[pch[p][k]];
      if (k == 0 || fr < lo) lo = fr;
      if (k == 0 || fr > hi) hi = fr;
    }
    pf[p] = lo; pf[nplane + p] = hi;
    // the plane's frequencies along its channel list, and whether they
    // are monotonic there (then so are a track's cells)
    bool up = true, down = true;
    for (size_t k = 0; k < pch[p].size(); ++k) {
      const float fr = gi->h_freqArr[pch[p][k]];
      hpfreq[(size_t)p * maxpc + k] = fr;
      if (k) {
        const float prev = gi->h_freqArr[pch[p][k - 1]];
        up = up && fr >= prev;
        down = down && fr <= prev;
      }
    }
    hmono[p] = (up || down) ? 1 : 0;
  }
  bool hasSigma = false;
  if (gi->h_sigma1)
    for (int c = 0; c < nchan; ++c) hasSigma = hasSigma || gi->h_sigma1[c] != 0.0f;
  // (re)allocate device tables
  (void)hipFree(gs->d_gcf); (void)hipFree(gs->d_freq); (void)hipFree(gs->d_sig); (void)hipFree(gs->d_pf);
  (void)hipFree(gs->d_pchan); (void)hipFree(gs->d_pcount); (void)hipFree(gs->d_facets); (void)hipFree(gs->d_listFacet);
  (void)hipFree(gs->d_pfreq); (void)hipFree(gs->d_pmono);
  OBIT_HIP_TRY(hipMalloc(&gs->d_pfreq, sizeof(float) * hpfreq.size()));
  OBIT_HIP_TRY(hipMalloc(&gs->d_pmono, sizeof(int) * nplane));
  OBIT_HIP_TRY(hipMemcpy(gs->d_pfreq, hpfreq.data(), sizeof(float) * hpfreq.size(), hipMemcpyHostToDevice));
  OBIT_HIP_TRY(hipMemcpy(gs->d_pmono, hmono.data(), sizeof(int) * nplane, hipMemcpyHostToDevice));
  OBIT_HIP_TRY(hipMalloc(&gs->d_gcf, sizeof(float) * kGcf8));
  OBIT_HIP_TRY(hipMalloc(&gs->d_freq, sizeof(float) * nchan));
  OBIT_HIP_TRY(hipMalloc(&gs->d_sig, sizeof(float) * 3 * nchan));
  OBIT_HIP_TRY(hipMalloc(&gs->d_pf, sizeof(float) * 2 * nplane));
  OBIT_HIP_TRY(hipMalloc(&gs->d_pchan, sizeof(int) * hpch.size()));
  OBIT_HIP_TRY(hipMalloc(&gs->d_pcount, sizeof(int) * nplane));
  OBIT_HIP_TRY(hipMalloc(&gs->d_facets, sizeof(FacetDev) * std::max(nf, 1)));
  OBIT_HIP_TRY(hipMalloc(&gs->d_listFacet, sizeof(int) * std::max<long long>(gs->nlist, 1)));
  // rows of 8 (seven taps and a zero) so a row loads as two float4
  std::vector<float> gcf8(kGcf8, 0.0f);
  for (int it = 0; it <= N; ++it)
    for (int c = 0; c < W; ++c) gcf8[8 * it + c] = gcf[(size_t)W * it + c];
  OBIT_HIP_TRY(hipMemcpy(gs->d_gcf, gcf8.data(), sizeof(float) * kGcf8, hipMemcpyHostToDevice));
  OBIT_HIP_TRY(hipMemcpy(gs->d_freq, gi->h_freqArr, sizeof(float) * nchan, hipMemcpyHostToDevice));
  std::vector<float> sig(3 * (size_t)nchan, 0.0f);
  if (gi->h_sigma1) std::memcpy(sig.data(), gi->h_sigma1, sizeof(float) * nchan);
  if (gi->h_sigma2) std::memcpy(sig.data() + nchan, gi->h_sigma2, sizeof(float) * nchan);
  if (gi->h_sigma3) std::memcpy(sig.data() + 2 * nchan, gi->h_sigma3, sizeof(float) * nchan);
  OBIT_HIP_TRY(hipMemcpy(gs->d_sig, sig.data(), sizeof(float) * 3 * nchan, hipMemcpyHostToDevice));
  OBIT_HIP_TRY(hipMemcpy(gs->d_pf, pf.data(), sizeof(float) * 2 * nplane, hipMemcpyHostToDevice));
  OBIT_HIP_TRY(hipMemcpy(gs->d_pchan, hpch.data(), sizeof(int) * hpch.size(), hipMemcpyHostToDevice));
  OBIT_HIP_TRY(hipMemcpy(gs->d_pcount, hpcount.data(), sizeof(int) * nplane, hipMemcpyHostToDevice));
  if (nf) OBIT_HIP_TRY(hipMemcpy(gs->d_facets, gs->h_facets.data(), sizeof(FacetDev) * nf, hipMemcpyHostToDevice));
  if (gs->nlist)
    OBIT_HIP_TRY(hipMemcpy(gs->d_listFacet, listFacet.data(), sizeof(int) * gs->nlist, hipMemcpyHostToDevice));
  Common &cm = gs->cm;
  cm.nrparm = gi->nrparm; cm.lenvis = gi->lenvis; cm.nchan = nchan; cm.nplane = nplane;
  cm.convWidth = W; cm.convNperCell = N;
  cm.freq = gs->d_freq;
  cm.sig1 = gs->d_sig; cm.sig2 = gs->d_sig + nchan; cm.sig3 = gs->d_sig + 2 * nchan;
  cm.hasSigma = hasSigma ? 1 : 0;
  cm.pchan = gs->d_pchan; cm.pcount = gs->d_pcount;
  cm.pfmin = gs->d_pf; cm.pfmax = gs->d_pf + nplane;
  cm.maxpc = maxpc;
  cm.pfreq = gs->d_pfreq;
  cm.pmono = gs->d_pmono;
  cm.gcf = gs->d_gcf;
  // list buffers sized for the lists; sort keys hold the list index in
  // 25 bits above the 7-bit window block
  if (gs->nlist >= (1LL << 25))
    return fail(kInvalid, "%lld tile lists (facets x planes x tiles) exceed 2^25", gs->nlist);
  (void)hipFree(gs->d_counts); (void)hipFree(gs->d_lend); (void)hipFree(gs->d_offsets);
  gs->d_counts = nullptr; gs->d_lend = nullptr; gs->d_offsets = nullptr;
  OBIT_HIP_TRY(hipMalloc(&gs->d_counts, sizeof(int) * (gs->nlist + 1)));
  OBIT_HIP_TRY(hipMalloc(&gs->d_lend, sizeof(long long) * (gs->nlist + 1)));
  OBIT_HIP_TRY(hipMalloc(&gs->d_offsets, sizeof(long long) * (gs->nlist + 1)));
  (void)hipFree(gs->d_itemOff);
  OBIT_HIP_TRY(hipMalloc(&gs->d_itemOff, sizeof(long long) * (gs->nlist + 1)));
  if (!gs->h_tot) OBIT_HIP_TRY(hipHostMalloc(&gs->h_tot, 2 * sizeof(long long), hipHostMallocDefault));
  gs->ready = true;
  return kOK;
}

// Bin and grid nrow rows at a device pointer into every facet on gs.
// List sizes as 64-bit entry counts and as kChunk-track work-item counts.
struct ToEntries {
  __host__ __device__ long long operator()(int n) const { return (long long)n; }
};
struct ToEntriesU {
  __host__ __device__ unsigned long long operator()(unsigned n) const { return (unsigned long long)n; }
};
struct ToChunks {
  __host__ __device__ long long operator()(int n) const { return (long long)((n + kChunk - 1) / kChunk); }
};

// Work items (list, chunk) of every list, in list order, at the list's
// scanned item offset.
__global__ __launch_bounds__(256) void k_items(const int *counts, const long long *itemOff, long long nl,
                                               int2 *items) {
  for (long long L = (long long)blockIdx.x * blockDim.x + threadIdx.x; L < nl;
       L += (long long)gridDim.x * blockDim.x) {
    const int n = counts[L];
    const long long base = itemOff[L];
    for (int c = 0; c * kChunk < n; ++c) items[base + c] = make_int2((int)L, c);
  }
}

int launch_taps(GpuState *gs, const float *vis);

// Scan / sort scratch of at least need bytes.
hipError_t ensure_tmp(GpuState *gs, size_t need) {
  if (need <= gs->scanTmpBytes) return hipSuccess;
  (void)hipFree(gs->d_scanTmp);
  gs->d_scanTmp = nullptr;
  gs->scanTmpBytes = 0;
  hipError_t e = hipMalloc(&gs->d_scanTmp, need);
  if (e == hipSuccess) gs->scanTmpBytes = need;
  return e;
}

// Checksum of the uvw of nrow rows (k_uvw_hash), read back.
int uvw_hash(GpuState *gs, const float *vis, long long nrow, unsigned long long &h) {
  if (!gs->d_hash) OBIT_HIP_TRY(hipMalloc(&gs->d_hash, sizeof(unsigned long long)));
  hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, gs->stream, (int *)gs->d_hash, (int *)gs->d_hash + 1, 1LL);
  const int nb = (int)std::min<long long>((nrow + 255) / 256, (long long)gs->nCU * 4);
  hipLaunchKernelGGL(k_uvw_hash, dim3(std::max(nb, 1)), dim3(256), 0, gs->stream, vis, nrow, gs->cm.lenvis,
                     gs->d_hash);
  OBIT_HIP_TRY(hipGetLastError());
  OBIT_HIP_TRY(hipMemcpyAsync(&h, gs->d_hash, sizeof(h), hipMemcpyDeviceToHost, gs->stream));
  OBIT_HIP_TRY(hipStreamSynchronize(gs->stream));
  return kOK;
}

int grid_device(GpuState *gs, const float *vis, long long nrow, bool reuse) {
  if (!gs->ready) return fail(kState, "gridder not initialised");
  if (nrow > (long long)INT_MAX)
    return fail(kInvalid, "%lld rows: the track lists index at most 2^31 - 1 rows a pass", nrow);
  if (nrow <= 0 || gs->nlist == 0) return kOK;
  if (reuse && !(gs->listsValid && gs->listsRows == nrow))
    return fail(kState, "no track lists of a %lld-row pass to reuse", nrow);
  const int nf = (int)gs->facets.size();
  const long long nl = gs->nlist;
  if (reuse) {
    unsigned long long h = 0;
    int rc = uvw_hash(gs, vis, nrow, h);
    if (rc) return rc;
    if (h != gs->listsHash)
      return fail(kState, "cannot reuse the track lists: these rows' u, v, w differ from the last pass's");
    // the lists, offsets and work items of the last pass stand: only the
    // walk statistics restart
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, gs->stream, (int *)gs->d_walk, (int *)gs->d_walk + 2, 2LL);
    OBIT_HIP_TRY(hipGetLastError());
    return launch_taps(gs, vis);
  }
  gs->listsValid = false;
  if ((long long)nrow * gs->cm.nplane * nf > (1LL << 40)) return fail(kInvalid, "too many tracks");
  if (!gs->d_walk) OBIT_HIP_TRY(hipMalloc(&gs->d_walk, 2 * sizeof(unsigned long long)));
  hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, gs->stream, (int *)gs->d_walk,
                     (int *)gs->d_walk + 2, 2LL);
  OBIT_HIP_TRY(hipGetLastError());
  const long long ntr = nrow * (long long)nf;
  if (ntr + 1 > gs->tcap) {
    (void)hipFree(gs->d_tcnt); (void)hipFree(gs->d_toff);
    gs->d_tcnt = nullptr; gs->d_toff = nullptr;
    OBIT_HIP_TRY(hipMalloc(&gs->d_tcnt, sizeof(unsigned) * (size_t)(ntr + 1)));
    OBIT_HIP_TRY(hipMalloc(&gs->d_toff, sizeof(unsigned long long) * (size_t)(ntr + 1)));
    gs->tcap = ntr + 1;
  }
  const int nbB = (int)std::min<long long>((ntr + 255) / 256, (long long)gs->nCU * 32);
  // pass 1: entries per track; the slot after the last track is zero so
  // the scan's last element is the total
  hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, gs->stream, (int *)gs->d_tcnt + ntr,
                     (int *)gs->d_tcnt + ntr, 1LL);
  hipLaunchKernelGGL(k_bin_count, dim3(nbB), dim3(256), 0, gs->stream, vis, nrow, gs->d_facets, nf, gs->cm,
                     gs->d_tcnt);
  OBIT_HIP_TRY(hipGetLastError());
  {
    size_t need = 0;
    const auto cnt64 = rocprim::make_transform_iterator(gs->d_tcnt, ToEntriesU());
    OBIT_HIP_TRY(rocprim::exclusive_scan(nullptr, need, cnt64, gs->d_toff, 0ULL, (size_t)ntr + 1,
                                         rocprim::plus<unsigned long long>(), gs->stream));
    OBIT_HIP_TRY(ensure_tmp(gs, need));
    OBIT_HIP_TRY(rocprim::exclusive_scan(gs->d_scanTmp, need, cnt64, gs->d_toff, 0ULL, (size_t)ntr + 1,
                                         rocprim::plus<unsigned long long>(), gs->stream));
    OBIT_HIP_TRY(hipMemcpyAsync(gs->h_tot, gs->d_toff + ntr, sizeof(long long), hipMemcpyDeviceToHost,
                                gs->stream));
    OBIT_HIP_TRY(hipStreamSynchronize(gs->stream));
  }
  const long long total = gs->h_tot[0];
  if (total < 0) return fail(kHipError, "binning scan gave %lld entries", total);
  if (total > gs->listCap) {
    (void)hipFree(gs->d_lists); (void)hipFree(gs->d_keys); (void)hipFree(gs->d_keys2); (void)hipFree(gs->d_rows);
    gs->d_lists = nullptr; gs->d_keys = nullptr; gs->d_keys2 = nullptr; gs->d_rows = nullptr;
    gs->listCap = 0;
    const size_t m = (size_t)std::max<long long>(total, 1);
    hipError_t e = hipMalloc(&gs->d_lists, sizeof(int2) * m);
    if (e == hipSuccess) e = hipMalloc(&gs->d_rows, sizeof(int2) * m);
    if (e == hipSuccess) e = hipMalloc(&gs->d_keys, sizeof(unsigned) * m);
    if (e == hipSuccess) e = hipMalloc(&gs->d_keys2, sizeof(unsigned) * m);
    if (e != hipSuccess) return fail(kNoMemory, "track lists (%lld entries): %s", total, hipGetErrorString(e));
    gs->listCap = total;
  }
  // pass 2: keys and rows, then the stable sort by (list, window block)
  hipLaunchKernelGGL(k_bin_emit, dim3(nbB), dim3(256), 0, gs->stream, vis, nrow, gs->d_facets, nf, gs->cm,
                     gs->d_toff, gs->d_keys, gs->d_rows);
  OBIT_HIP_TRY(hipGetLastError());
  int endBit = 7;
  while ((1LL << (endBit - 7)) < nl) ++endBit;
  if (total > 0) {
    size_t need = 0;
    OBIT_HIP_TRY(rocprim::radix_sort_pairs(nullptr, need, gs->d_keys, gs->d_keys2, gs->d_rows, gs->d_lists,
                                           (size_t)total, 0, endBit, gs->stream));
    OBIT_HIP_TRY(ensure_tmp(gs, need));
    OBIT_HIP_TRY(rocprim::radix_sort_pairs(gs->d_scanTmp, need, gs->d_keys, gs->d_keys2, gs->d_rows,
                                           gs->d_lists, (size_t)total, 0, endBit, gs->stream));
  }
  // list extents and sizes, then the work items: an exclusive scan of the
  // lists' kChunk-track chunk counts (the entry after the last list is
  // zero, so the scan's last element is the number of items)
  {
    const int nbz = (int)std::min<long long>((nl + 1 + 255) / 256, 65535);
    hipLaunchKernelGGL(k_zero64, dim3(nbz), dim3(256), 0, gs->stream, gs->d_offsets, gs->d_lend, nl + 1);
    const int nbk = (int)std::min<long long>((total + 255) / 256, (long long)gs->nCU * 32);
    if (total > 0)
      hipLaunchKernelGGL(k_list_bounds, dim3(nbk), dim3(256), 0, gs->stream, gs->d_keys2, total, gs->d_offsets,
                         gs->d_lend);
    hipLaunchKernelGGL(k_list_counts, dim3(nbz), dim3(256), 0, gs->stream, gs->d_offsets, gs->d_lend, nl + 1,
                       gs->d_counts);
    OBIT_HIP_TRY(hipGetLastError());
    const auto chunks = rocprim::make_transform_iterator(gs->d_counts, ToChunks());
    size_t need = 0;
    OBIT_HIP_TRY(rocprim::exclusive_scan(nullptr, need, chunks, gs->d_itemOff, 0LL, (size_t)nl + 1,
                                         rocprim::plus<long long>(), gs->stream));
    OBIT_HIP_TRY(ensure_tmp(gs, need));
    OBIT_HIP_TRY(rocprim::exclusive_scan(gs->d_scanTmp, need, chunks, gs->d_itemOff, 0LL, (size_t)nl + 1,
                                         rocprim::plus<long long>(), gs->stream));
    OBIT_HIP_TRY(hipMemcpyAsync(gs->h_tot + 1, gs->d_itemOff + nl, sizeof(long long), hipMemcpyDeviceToHost,
                                gs->stream));
    OBIT_HIP_TRY(hipStreamSynchronize(gs->stream));
  }
  const long long nitems = gs->h_tot[1];
  gs->nitems = nitems;
  if (nitems < 0 || nitems > total) return fail(kHipError, "list scan gave %lld entries, %lld items", total, nitems);
  if (nitems > gs->itemCap) {
    (void)hipFree(gs->d_items);
    gs->d_items = nullptr;
    OBIT_HIP_TRY(hipMalloc(&gs->d_items, sizeof(int2) * (size_t)std::max<long long>(nitems, 1)));
    gs->itemCap = nitems;
  }
  if (nitems) {
    const int nbI = (int)std::min<long long>((nl + 255) / 256, (long long)gs->nCU * 16);
    hipLaunchKernelGGL(k_items, dim3(nbI), dim3(256), 0, gs->stream, gs->d_counts, gs->d_itemOff, nl,
                       gs->d_items);
    OBIT_HIP_TRY(hipGetLastError());
  }
  {
    int rc = uvw_hash(gs, vis, nrow, gs->listsHash);
    if (rc) return rc;
  }
  gs->listsValid = true;
  gs->listsRows = nrow;
  return launch_taps(gs, vis);
}

// The tile-gridding launch over the current work items; the pass is
// complete when it returns.
int launch_taps(GpuState *gs, const float *vis) {
  const long long nitems = gs->nitems;
  if (nitems) {
    // persistent grid of exactly the resident blocks: more would queue a
    // second round of blocks behind the first and idle CUs at the end
    if (gs->tapsPerCU <= 0) {
      int nb = 0;
      OBIT_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_grid_taps, 64 * kWaves, 0));
      gs->tapsPerCU = std::max(nb, 1);
    }
    const long long nbG =
        std::min<long long>((nitems + kWaves - 1) / kWaves, (long long)gs->nCU * gs->tapsPerCU);
    hipLaunchKernelGGL(k_grid_taps, dim3((unsigned)nbG), dim3(64 * kWaves), 0, gs->stream, vis,
                       gs->d_facets, gs->d_listFacet, (const int2 *)gs->d_items, nitems, gs->cm,
                       gs->d_counts, gs->d_offsets, gs->d_lists, gs->d_walk);
    OBIT_HIP_TRY(hipGetLastError());
  }
  // the pass is complete when the call returns
  OBIT_HIP_TRY(hipStreamSynchronize(gs->stream));
  return kOK;
}

}  // namespace grid
}  // namespace obit

using obit::grid::GpuState;

static void grid_abort(const char *where) {
  std::fprintf(stderr, "obit_hip: %s: %s\n", where, ObitHIPLastError());
  std::abort();
}

extern "C" {

void ObitCUDAGridAlloc(CUDAGridInfo *cudaInfo, long nfacet, long nchan, long nif, long nVisPIO,
                       long lenvis, long nplane, long iGPU, int doBuff) {
  if (!cudaInfo || nfacet <= 0 || iGPU < 0 || iGPU >= (cudaInfo->nGPU > 0 ? cudaInfo->nGPU : 1)) {
    std::fprintf(stderr, "obit_hip: ObitCUDAGridAlloc: bad arguments\n");
    std::abort();
  }
  const long nGPU = cudaInfo->nGPU > 0 ? cudaInfo->nGPU : 1;
  cudaInfo->nVisPIO = nVisPIO;
  cudaInfo->nfacet = nfacet;
  if (!cudaInfo->h_gridInfo) {
    cudaInfo->h_gridInfo = (GridInfo *)ObitCUDAUtilAllocHost(sizeof(GridInfo));
    std::memset(cudaInfo->h_gridInfo, 0, sizeof(GridInfo));
    cudaInfo->h_gridInfo->nStream = GRID_STREAM_COUNT;
    cudaInfo->h_gridInfo->stream = (int **)std::calloc(nGPU, sizeof(int *));
    cudaInfo->h_gridInfo->cycleDone = (int **)std::calloc(nGPU, sizeof(int *));
  }
  GridInfo *gi = cudaInfo->h_gridInfo;
  gi->nfacet = (int)nfacet;
  gi->nchan = (int)(nchan * nif);
  gi->nif = (int)nif;
  gi->nvis = (int)nVisPIO;
  gi->lenvis = (int)lenvis;
  gi->nplane = (int)nplane;
  gi->nGPU = (int)nGPU;
  const size_t nci = (size_t)(nchan * nif);
  // the reference sizes the plane table with sizeof(long) per entry
  if (!gi->h_freqPlane) { gi->h_freqPlane = (int *)ObitCUDAUtilAllocHost((int)(nci * sizeof(long))); std::memset(gi->h_freqPlane, 0, nci * sizeof(long)); }
  if (!gi->h_sigma1) {
    gi->h_sigma1 = ObitCUDAUtilAllocHost((int)(nci * sizeof(float)));
    gi->h_sigma2 = ObitCUDAUtilAllocHost((int)(nci * sizeof(float)));
    gi->h_sigma3 = ObitCUDAUtilAllocHost((int)(nci * sizeof(float)));
    std::memset(gi->h_sigma1, 0, nci * sizeof(float));
    std::memset(gi->h_sigma2, 0, nci * sizeof(float));
    std::memset(gi->h_sigma3, 0, nci * sizeof(float));
  }
  if (!gi->h_freqArr) { gi->h_freqArr = ObitCUDAUtilAllocHost((int)(nci * sizeof(float))); std::memset(gi->h_freqArr, 0, nci * sizeof(float)); }
  if (!gi->h_convfn) { gi->h_convfn = ObitCUDAUtilAllocHost(2000 * sizeof(float)); std::memset(gi->h_convfn, 0, 2000 * sizeof(float)); }
  if (doBuff && !gi->h_data_in) {
    const size_t b = (size_t)nVisPIO * lenvis * sizeof(float);
    void *p = nullptr;
    OBIT_HIP_CHECK(hipHostMalloc(&p, b ? b : 4, hipHostMallocDefault));
    gi->h_data_in = (float *)p;
    cudaInfo->h_data_in = gi->h_data_in;
  }
  if (!cudaInfo->h_facetInfo) {
    cudaInfo->h_facetInfo = (FacetInfo **)ObitCUDAUtilAllocHost((int)(nfacet * sizeof(FacetInfo *)));
    for (long i = 0; i < nfacet; ++i) {
      cudaInfo->h_facetInfo[i] = (FacetInfo *)ObitCUDAUtilAllocHost(sizeof(FacetInfo));
      std::memset(cudaInfo->h_facetInfo[i], 0, sizeof(FacetInfo));
      cudaInfo->h_facetInfo[i]->GPU_num = (int)iGPU;
      cudaInfo->h_facetInfo[i]->rotUV[0] = 1.0f;
      cudaInfo->h_facetInfo[i]->rotUV[4] = 1.0f;
      cudaInfo->h_facetInfo[i]->rotUV[8] = 1.0f;
    }
  }
  GpuState *gs = obit::grid::get_gpu(cudaInfo, iGPU, true);
  gs->device = obit::grid::device_of(cudaInfo, iGPU);
  OBIT_HIP_CHECK(hipSetDevice(gs->device));
  if (!gs->stream) OBIT_HIP_CHECK(hipStreamCreateWithFlags(&gs->stream, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  OBIT_HIP_CHECK(hipGetDeviceProperties(&prop, gs->device));
  gs->nCU = prop.multiProcessorCount;
  if (!gi->stream[iGPU]) gi->stream[iGPU] = (int *)gs->stream;
}

void ObitCUDAGridInit(CUDAGridInfo *cudaInfo, long iGPU) {
  GpuState *gs = obit::grid::get_gpu(cudaInfo, iGPU, false);
  if (!gs) { std::fprintf(stderr, "obit_hip: ObitCUDAGridInit before Alloc\n"); std::abort(); }
  gs->device = obit::grid::device_of(cudaInfo, iGPU);
  OBIT_HIP_CHECK(hipSetDevice(gs->device));
  const int nplane = cudaInfo->h_gridInfo->nplane;
  for (long i = 0; i < cudaInfo->nfacet; ++i) {
    if (cudaInfo->FacetGPU && cudaInfo->FacetGPU[i] != iGPU) continue;
    FacetInfo *fi = cudaInfo->h_facetInfo[i];
    const size_t bytes = (size_t)fi->sizeGrid * nplane * sizeof(float);
    if (!fi->d_grid) OBIT_HIP_CHECK(hipMalloc(&fi->d_grid, bytes));
    // the pinned host copy (h_grid) is allocated by ObitCUDAGrid2CPU on
    // first use, so grids that stay on the device cost no host memory
    OBIT_HIP_CHECK(hipMemsetAsync(fi->d_grid, 0, bytes, gs->stream));
  }
  OBIT_HIP_CHECK(hipStreamSynchronize(gs->stream));
  if (obit::grid::refresh(cudaInfo, iGPU, gs)) grid_abort("ObitCUDAGridInit");
}

void ObitCUDAGridGrid(CUDAGridInfo *cudaInfo, long iGPU, int init) {
  GpuState *gs = obit::grid::get_gpu(cudaInfo, iGPU, false);
  if (!gs) { std::fprintf(stderr, "obit_hip: ObitCUDAGridGrid before Alloc\n"); std::abort(); }
  OBIT_HIP_CHECK(hipSetDevice(gs->device));
  if (init || !gs->ready)
    if (obit::grid::refresh(cudaInfo, iGPU, gs)) grid_abort("ObitCUDAGridGrid");
  GridInfo *gi = cudaInfo->h_gridInfo;
  const long long nvis = gi->nvis;
  if (nvis <= 0) return;
  const size_t bytes = (size_t)nvis * gi->lenvis * sizeof(float);
  if (bytes > gs->visCap) {
    (void)hipFree(gs->d_vis);
    OBIT_HIP_CHECK(hipMalloc(&gs->d_vis, bytes));
    gs->visCap = bytes;
  }
  OBIT_HIP_CHECK(hipMemcpyAsync(gs->d_vis, gi->h_data_in, bytes, hipMemcpyHostToDevice, gs->stream));
  if (obit::grid::grid_device(gs, gs->d_vis, nvis, false)) grid_abort("ObitCUDAGridGrid");
  OBIT_HIP_CHECK(hipStreamSynchronize(gs->stream));
}

void ObitCUDAGridFlip(CUDAGridInfo *cudaInfo, long iGPU) {
  GpuState *gs = obit::grid::get_gpu(cudaInfo, iGPU, false);
  if (!gs || !gs->ready) { std::fprintf(stderr, "obit_hip: ObitCUDAGridFlip before Init\n"); std::abort(); }
  OBIT_HIP_CHECK(hipSetDevice(gs->device));
  const int nf = (int)gs->facets.size();
  const int nplane = cudaInfo->h_gridInfo->nplane;
  int maxny = 0, hw = gs->cm.convWidth / 2;
  for (auto &f : gs->h_facets) maxny = std::max(maxny, f.ny);
  const long long n = (long long)(maxny / 2) * (hw + 1);
  if (nf && n > 0) {
    dim3 grid((unsigned)((n + 255) / 256), 1, (unsigned)(nf * nplane));
    hipLaunchKernelGGL(obit::grid::k_flip, grid, dim3(256), 0, gs->stream, gs->d_facets, nf, nplane, hw);
    OBIT_HIP_CHECK(hipGetLastError());
  }
  OBIT_HIP_CHECK(hipStreamSynchronize(gs->stream));
}

void ObitCUDAGrid2CPU(CUDAGridInfo *cudaInfo, long ifacet) {
  FacetInfo *fi = cudaInfo->h_facetInfo[ifacet];
  const long iGPU = cudaInfo->FacetGPU ? cudaInfo->FacetGPU[ifacet] : 0;
  GpuState *gs = obit::grid::get_gpu(cudaInfo, iGPU, false);
  if (gs) OBIT_HIP_CHECK(hipSetDevice(gs->device));
  const size_t bytes = (size_t)fi->sizeGrid * cudaInfo->h_gridInfo->nplane * sizeof(float);
  if (!fi->h_grid) {
    void *p = nullptr;
    OBIT_HIP_CHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    fi->h_grid = (float *)p;
  }
  if (gs) OBIT_HIP_CHECK(hipStreamSynchronize(gs->stream));
  OBIT_HIP_CHECK(hipMemcpy(fi->h_grid, fi->d_grid, bytes, hipMemcpyDeviceToHost));
}

void ObitCUDAGridShutdown(CUDAGridInfo *cudaInfo, long iGPU) {
  GpuState *gs = obit::grid::get_gpu(cudaInfo, iGPU, false);
  GridInfo *gi = cudaInfo->h_gridInfo;
  const int nGPU = gi ? gi->nGPU : 1;
  const bool last = (iGPU == nGPU - 1);
  if (gs) {
    (void)hipSetDevice(gs->device);
    (void)hipStreamSynchronize(gs->stream);
    for (long i = 0; i < cudaInfo->nfacet; ++i) {
      if (cudaInfo->FacetGPU && cudaInfo->FacetGPU[i] != iGPU) continue;
      FacetInfo *fi = cudaInfo->h_facetInfo ? cudaInfo->h_facetInfo[i] : nullptr;
      if (fi && fi->d_grid) { (void)hipFree(fi->d_grid); fi->d_grid = nullptr; }
    }
    (void)hipFree(gs->d_gcf); (void)hipFree(gs->d_freq); (void)hipFree(gs->d_sig); (void)hipFree(gs->d_pf);
    (void)hipFree(gs->d_pfreq); (void)hipFree(gs->d_pmono);
    (void)hipFree(gs->d_pchan); (void)hipFree(gs->d_pcount); (void)hipFree(gs->d_facets); (void)hipFree(gs->d_listFacet);
    (void)hipFree(gs->d_counts); (void)hipFree(gs->d_lend); (void)hipFree(gs->d_offsets); (void)hipFree(gs->d_lists);
    (void)hipFree(gs->d_tcnt); (void)hipFree(gs->d_toff); (void)hipFree(gs->d_keys); (void)hipFree(gs->d_keys2);
    (void)hipFree(gs->d_rows);
    (void)hipFree(gs->d_items); (void)hipFree(gs->d_vis); (void)hipFree(gs->d_walk); (void)hipFree(gs->d_hash);
    (void)hipFree(gs->d_itemOff); (void)hipFree(gs->d_scanTmp);
    if (gs->h_tot) (void)hipHostFree(gs->h_tot);
    if (gs->fftPlan) (void)hipfftDestroy(gs->fftPlan);
    (void)hipFree(gs->d_fft); (void)hipFree(gs->d_corr); (void)hipFree(gs->d_tw);
    if (gs->h_norm) (void)hipHostFree(gs->h_norm);
    if (gi && gi->stream && gi->stream[iGPU] == (int *)gs->stream) gi->stream[iGPU] = nullptr;
    (void)hipStreamDestroy(gs->stream);
    obit::grid::GridCtx *c = obit::grid::get_ctx(cudaInfo, false);
    if (c) c->gpus[iGPU] = nullptr;
    delete gs;
  }
  if (last && gi) {
    if (gi->h_data_in) { (void)hipHostFree(gi->h_data_in); gi->h_data_in = nullptr; cudaInfo->h_data_in = nullptr; }
    ObitCUDAUtilFreeHost(gi->h_freqArr);
    ObitCUDAUtilFreeHost((float *)gi->h_freqPlane);
    ObitCUDAUtilFreeHost(gi->h_sigma1);
    ObitCUDAUtilFreeHost(gi->h_sigma2);
    ObitCUDAUtilFreeHost(gi->h_sigma3);
    ObitCUDAUtilFreeHost(gi->h_convfn);
    if (gi->GPU_device_no) ObitCUDAUtilFreeHost((float *)gi->GPU_device_no);
    std::free(gi->stream);
    std::free(gi->cycleDone);
    if (cudaInfo->h_facetInfo) {
      for (long i = 0; i < cudaInfo->nfacet; ++i) {
        FacetInfo *fi = cudaInfo->h_facetInfo[i];
        if (!fi) continue;
        if (fi->h_grid) (void)hipHostFree(fi->h_grid);
        ObitCUDAUtilFreeHost((float *)fi);
      }
      ObitCUDAUtilFreeHost((float *)cudaInfo->h_facetInfo);
      cudaInfo->h_facetInfo = nullptr;
    }
    ObitCUDAUtilFreeHost((float *)gi);
    cudaInfo->h_gridInfo = nullptr;
    std::lock_guard<std::mutex> lk(obit::grid::g_mu);
    auto it = obit::grid::g_ctx.find(cudaInfo);
    if (it != obit::grid::g_ctx.end()) { delete it->second; obit::grid::g_ctx.erase(it); }
  }
}

// ---- extensions ----

int ObitHIPGridGridDevice(CUDAGridInfo *in, long iGPU, const float *d_vis, long long nvis) {
  if (nvis > (long long)INT_MAX)
    return obit::fail(obit::kInvalid, "%lld rows: the track lists index at most 2^31 - 1 rows a pass", nvis);
  GpuState *gs = obit::grid::get_gpu(in, iGPU, false);
  if (!gs) return obit::fail(obit::kState, "gridder not allocated on GPU index %ld", iGPU);
  OBIT_HIP_TRY(hipSetDevice(gs->device));
  if (!gs->ready) {
    int rc = obit::grid::refresh(in, iGPU, gs);
    if (rc) return rc;
  }
  return obit::grid::grid_device(gs, d_vis, nvis, false);
}

int ObitHIPGridGridDeviceReuse(CUDAGridInfo *in, long iGPU, const float *d_vis, long long nvis) {
  if (nvis > (long long)INT_MAX)
    return obit::fail(obit::kInvalid, "%lld rows: the track lists index at most 2^31 - 1 rows a pass", nvis);
  GpuState *gs = obit::grid::get_gpu(in, iGPU, false);
  if (!gs) return obit::fail(obit::kState, "gridder not allocated on GPU index %ld", iGPU);
  OBIT_HIP_TRY(hipSetDevice(gs->device));
  return obit::grid::grid_device(gs, d_vis, nvis, true);
}

int ObitHIPGrid2Device(CUDAGridInfo *in, long ifacet, float *d_out) {
  const long iGPU = in->FacetGPU ? in->FacetGPU[ifacet] : 0;
  GpuState *gs = obit::grid::get_gpu(in, iGPU, false);
  if (!gs || !gs->ready) return obit::fail(obit::kState, "gridder not initialised");
  OBIT_HIP_TRY(hipSetDevice(gs->device));
  int k = -1;
  for (size_t i = 0; i < gs->facets.size(); ++i) if (gs->facets[i] == ifacet) k = (int)i;
  if (k < 0) return obit::fail(obit::kInvalid, "facet %ld not on GPU index %ld", ifacet, iGPU);
  const auto &f = gs->h_facets[k];
  const long long total = (long long)in->h_gridInfo->nplane * f.ny * (f.nx / 2 + 1);
  const int nb = (int)std::min<long long>((total + 255) / 256, (long long)gs->nCU * 16);
  hipLaunchKernelGGL(obit::grid::k_trim_centre, dim3(std::max(nb, 1)), dim3(256), 0, gs->stream,
                     gs->d_facets, k, in->h_gridInfo->nplane, d_out);
  OBIT_HIP_TRY(hipGetLastError());
  return obit::kOK;
}

int ObitHIPGrid2HostPlanes(CUDAGridInfo *in, long ifacet, float *h_out) {
  FacetInfo *fi = in->h_facetInfo[ifacet];
  const long long n = (long long)in->h_gridInfo->nplane * fi->ny * (fi->nx / 2 + 1) * 2;
  float *d = nullptr;
  OBIT_HIP_TRY(hipMalloc(&d, sizeof(float) * n));
  int rc = ObitHIPGrid2Device(in, ifacet, d);
  if (!rc) {
    const long iGPU = in->FacetGPU ? in->FacetGPU[ifacet] : 0;
    GpuState *gs = obit::grid::get_gpu(in, iGPU, false);
    OBIT_HIP_TRY(hipStreamSynchronize(gs->stream));
    OBIT_HIP_TRY(hipMemcpy(h_out, d, sizeof(float) * n, hipMemcpyDeviceToHost));
  }
  (void)hipFree(d);
  return rc;
}

int ObitHIPGridFFT2Image(CUDAGridInfo *in, long ifacet, int isBeam, float *beamNorm, float *d_image) {
  if (!in || !in->h_gridInfo || !beamNorm || !d_image) return obit::fail(obit::kInvalid, "null argument");
  const long iGPU = in->FacetGPU ? in->FacetGPU[ifacet] : 0;
  GpuState *gs = obit::grid::get_gpu(in, iGPU, false);
  if (!gs || !gs->ready) return obit::fail(obit::kState, "gridder not initialised");
  OBIT_HIP_TRY(hipSetDevice(gs->device));
  int k = -1;
  for (size_t i = 0; i < gs->facets.size(); ++i) if (gs->facets[i] == ifacet) k = (int)i;
  if (k < 0) return obit::fail(obit::kInvalid, "facet %ld not on GPU index %ld", ifacet, iGPU);
  const auto &f = gs->h_facets[k];
  const int nx = f.nx, ny = f.ny, nplane = in->h_gridInfo->nplane;
  const GridInfo *gi = in->h_gridInfo;
  if (nx != ny || (nx % 8))
    return obit::fail(obit::kInvalid, "FFT2Image needs a square facet with a side a multiple of 8 (got %d x %d): "
                      "the reference's correction indexes rows and columns alike", nx, ny);
  if (!gi->h_convfn) return obit::fail(obit::kState, "no convolving function");
  if (nplane > 1024) return obit::fail(obit::kInvalid, "%d planes: at most 1024", nplane);
  // gridding corrections (GridCorrFn, computed once per size): xCorr[nx],
  // yCorr[ny], then the planes' norms and reciprocals
  const size_t ncorr = (size_t)nx + ny + 2 * 1024;
  if (gs->corrNx != nx || gs->corrNy != ny || gs->corrCap < ncorr) {
    gs->h_corr.assign((size_t)nx + ny, 0.0f);
    ObitHostGridCorrFn(nx, nx / 2 + 1, gi->h_convfn, gi->convWidth, gi->convNperCell, gs->h_corr.data());
    ObitHostGridCorrFn(ny, ny / 2 + 1, gi->h_convfn, gi->convWidth, gi->convNperCell,
                       gs->h_corr.data() + nx);
    if (gs->corrCap < ncorr) {
      (void)hipFree(gs->d_corr);
      gs->d_corr = nullptr;
      gs->corrCap = 0;
      OBIT_HIP_TRY(hipMalloc(&gs->d_corr, sizeof(float) * ncorr));
      gs->corrCap = ncorr;
    }
    OBIT_HIP_TRY(hipMemcpy(gs->d_corr, gs->h_corr.data(), sizeof(float) * (nx + ny), hipMemcpyHostToDevice));
    gs->corrNx = nx;
    gs->corrNy = ny;
  }
  // coherent, so the norm kernels' stores reach it without a copy
  if (!gs->h_norm) OBIT_HIP_TRY(hipHostMalloc(&gs->h_norm, sizeof(float) * 2 * 1024, hipHostMallocCoherent));
  float *d_norm = gs->d_corr + nx + ny, *d_inv = d_norm + 1024;
  const float *xc = gs->d_corr, *yc = gs->d_corr + nx;
  if (!isBeam) {
    // an image takes its beam's norms
    OBIT_HIP_TRY(hipStreamSynchronize(gs->stream));  // the pinned staging is free
    for (int p = 0; p < nplane; ++p) gs->h_norm[p] = (float)(1.0 / std::max(1.0e-20, (double)beamNorm[p]));
    OBIT_HIP_TRY(hipMemcpyAsync(d_inv, gs->h_norm, sizeof(float) * nplane, hipMemcpyHostToDevice, gs->stream));
  }
  // FFTW's unnormalised backward transform of the gridder's padded half
  // planes (row pitch lrow/2 complex, the u = 0 column hw complex into each
  // row, planes sizeGrid floats apart), then ObitFArray2DCenter's quadrant
  // swap and ObitFArrayMulColRow's correction.  For a side of 1024 or 2048:
  // k_fft_cols8 transforms the columns u = 0 .. nx/2 along v
  // in place (the grid is consumed), a beam's norm is summed from row 0 of
  // the result, and
  // k_fft_rows_c2r transforms every row along u, centres, corrects and
  // writes the image: two passes over the planes.  Other sizes: one batched
  // 2-D C2R (hipFFT) into scratch, then k_center_correct4.
  const int pitchC = f.lrow / 2;
  const bool own = (nx == 1024 || nx == 2048) && pitchC % 4 == 0 && f.sizeGrid % 8 == 0 && ny % 4 == 0;
  if (own) {
    // exp(+2 pi i m / n) for m < n, and for m < n / 2 at n / 2, in FP64
    if (gs->twN != ny) {
      std::vector<float2> tw((size_t)ny + ny / 2);
      for (int m = 0; m < ny; ++m) {
        const double a = 2.0 * M_PI * (double)m / (double)ny;
        tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
      }
      for (int m = 0; m < ny / 2; ++m) tw[ny + m] = tw[2 * m];
      (void)hipFree(gs->d_tw);
      gs->d_tw = nullptr;
      gs->twN = 0;
      OBIT_HIP_TRY(hipMalloc(&gs->d_tw, sizeof(float2) * tw.size()));
      OBIT_HIP_TRY(hipMemcpy(gs->d_tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
      gs->twN = ny;
    }
    const float2 *twN = gs->d_tw, *twM = gs->d_tw + ny;
    float2 *g2 = reinterpret_cast<float2 *>(f.grid);
    const long long planeC = (long long)f.sizeGrid / 2;
    // The nx/2 + 1 columns u = 0 .. nx/2 (hw .. hw + nx/2 of the padded
    // rows; the hw padding columns left of them are not read again): nx/2
    // of them in 8-column blocks starting at an even column, so the list
    // of nplane x nx/16 blocks divides evenly over the CUs, and the one
    // left over (u = 0 or u = nx/2) per plane after them.
    const int c0 = (f.hw % 2) ? f.hw + 1 : f.hw;
    const int side = (f.hw % 2) ? f.hw : f.hw + nx / 2;
    const int nfull = nx / 16;
    if (gs->colsPerCU <= 0) {
      int nb = 0;
      OBIT_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, obit::grid::k_fft_cols8<2048, 8, 16, 16>, 512, 0));
      gs->colsPerCU = std::max(nb, 1);
      OBIT_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, obit::grid::k_fft_cols8<1024, 4, 16, 16>, 512, 0));
      gs->colsPerCU1024 = std::max(nb, 1);
    }
    const long long resident = (long long)gs->nCU * (nx == 2048 ? gs->colsPerCU : gs->colsPerCU1024);
    const long long ncb = 8 * ((std::min<long long>((long long)nplane * nfull, resident) + 7) / 8);
    const dim3 gcol((unsigned)ncb);
    // rows a block: 1 at 2048 (measured 8.1 ms for 320 planes, against 8.2
    // with 2 rows, 8.4 with 4 and 9.35 with 8), 4 at 1024
    const dim3 grow((unsigned)(ny / 4), (unsigned)nplane), grow2048((unsigned)ny, (unsigned)nplane);
    if (nx == 2048)
      hipLaunchKernelGGL((obit::grid::k_fft_cols8<2048, 8, 16, 16>), gcol, dim3(512), 0, gs->stream, g2, pitchC,
                         planeC, c0, nplane, nfull, side, twN);
    else
      hipLaunchKernelGGL((obit::grid::k_fft_cols8<1024, 4, 16, 16>), gcol, dim3(512), 0, gs->stream, g2, pitchC,
                         planeC, c0, nplane, nfull, side, twN);
    OBIT_HIP_TRY(hipGetLastError());
    if (isBeam) {
      hipLaunchKernelGGL(obit::grid::k_beam_norm_rows, dim3((unsigned)nplane), dim3(256), 0, gs->stream,
                         (const float2 *)g2, planeC, f.hw, nx, xc, yc, ny, d_norm, d_inv, gs->h_norm);
      OBIT_HIP_TRY(hipGetLastError());
    }
    if (nx == 2048)
      hipLaunchKernelGGL((obit::grid::k_fft_rows_c2r<2048, 4, 16, 16, 1>), grow2048, dim3(256), 0, gs->stream,
                         (const float2 *)g2, pitchC, planeC, f.hw, ny, twN, twM, xc, yc, (const float *)d_inv,
                         d_image);
    else
      hipLaunchKernelGGL((obit::grid::k_fft_rows_c2r<1024, 8, 8, 8>), grow, dim3(256), 0, gs->stream,
                         (const float2 *)g2, pitchC, planeC, f.hw, ny, twN, twM, xc, yc, (const float *)d_inv,
                         d_image);
    OBIT_HIP_TRY(hipGetLastError());
  } else {
    // real output of the C2R for every plane
    const size_t need = (size_t)nplane * ny * nx;
    if (need > gs->fftCap) {
      (void)hipFree(gs->d_fft);
      gs->d_fft = nullptr;
      gs->fftCap = 0;
      OBIT_HIP_TRY(hipMalloc(&gs->d_fft, sizeof(float) * need));
      gs->fftCap = need;
    }
    if (!gs->fftPlan || gs->fftNx != nx || gs->fftNy != ny || gs->fftNplane != nplane ||
        gs->fftPitch != pitchC || gs->fftDist != f.sizeGrid) {
      if (gs->fftPlan) (void)hipfftDestroy(gs->fftPlan);
      gs->fftPlan = 0;
      int n[2] = {ny, nx}, inembed[2] = {ny, pitchC}, onembed[2] = {ny, nx};
      hipfftResult r = hipfftPlanMany(&gs->fftPlan, 2, n, inembed, 1, f.sizeGrid / 2, onembed, 1, ny * nx,
                                      HIPFFT_C2R, nplane);
      if (r != HIPFFT_SUCCESS) return obit::fail(obit::kHipError, "hipfftPlanMany: %d", (int)r);
      gs->fftNx = nx;
      gs->fftNy = ny;
      gs->fftNplane = nplane;
      gs->fftPitch = pitchC;
      gs->fftDist = f.sizeGrid;
    }
    if (hipfftSetStream(gs->fftPlan, gs->stream) != HIPFFT_SUCCESS)
      return obit::fail(obit::kHipError, "hipfftSetStream");
    hipfftComplex *src = reinterpret_cast<hipfftComplex *>(f.grid + 2 * f.hw);
    hipfftResult r = hipfftExecC2R(gs->fftPlan, src, reinterpret_cast<hipfftReal *>(gs->d_fft));
    if (r != HIPFFT_SUCCESS) return obit::fail(obit::kHipError, "hipfftExecC2R: %d", (int)r);
    if (isBeam) {
      hipLaunchKernelGGL(obit::grid::k_beam_norm, dim3((nplane + 63) / 64), dim3(64), 0, gs->stream, gs->d_fft, nx,
                         ny, nplane, xc, yc, d_norm, d_inv, gs->h_norm);
      OBIT_HIP_TRY(hipGetLastError());
    }
    dim3 grid((unsigned)((nx / 4 + 255) / 256), (unsigned)ny, (unsigned)nplane);
    hipLaunchKernelGGL(obit::grid::k_center_correct4, grid, dim3(256), 0, gs->stream, gs->d_fft, nx, ny, xc, yc,
                       d_inv, d_image);
    OBIT_HIP_TRY(hipGetLastError());
  }
  OBIT_HIP_TRY(hipStreamSynchronize(gs->stream));
  if (isBeam) {
    for (int p = 0; p < nplane; ++p) {
      beamNorm[p] = gs->h_norm[p];
      if (beamNorm[p] == 0.0f) return obit::fail(obit::kInvalid, "beam peak is zero in plane %d", p);
    }
  }
  return obit::kOK;
}

int ObitHIPGridPrepDebug(CUDAGridInfo *in, long ifacet, const float *h_vis, long long nvis, int *iu,
                         int *iv, int *itu, int *itv, float *vis) {
  const long iGPU = in->FacetGPU ? in->FacetGPU[ifacet] : 0;
  GpuState *gs = obit::grid::get_gpu(in, iGPU, false);
  if (!gs) return obit::fail(obit::kState, "gridder not allocated");
  OBIT_HIP_TRY(hipSetDevice(gs->device));
  if (!gs->ready) { int rc = obit::grid::refresh(in, iGPU, gs); if (rc) return rc; }
  int k = -1;
  for (size_t i = 0; i < gs->facets.size(); ++i) if (gs->facets[i] == ifacet) k = (int)i;
  if (k < 0) return obit::fail(obit::kInvalid, "facet %ld not on GPU index %ld", ifacet, iGPU);
  const int nch = gs->cm.nchan;
  const long long ns = nvis * nch;
  const size_t vb = (size_t)nvis * gs->cm.lenvis * sizeof(float);
  float *dv = nullptr, *dvo = nullptr;
  int *di = nullptr;
  OBIT_HIP_TRY(hipMalloc(&dv, vb ? vb : 4));
  OBIT_HIP_TRY(hipMalloc(&dvo, sizeof(float) * 2 * std::max<long long>(ns, 1)));
  OBIT_HIP_TRY(hipMalloc(&di, sizeof(int) * 4 * std::max<long long>(ns, 1)));
  OBIT_HIP_TRY(hipMemcpy(dv, h_vis, vb, hipMemcpyHostToDevice));
  if (ns > 0) {
    const int nb = (int)std::min<long long>((ns + 255) / 256, 4096);
    hipLaunchKernelGGL(obit::grid::k_prep_debug, dim3(nb), dim3(256), 0, gs->stream, dv, nvis,
                       gs->d_facets, k, gs->cm, di, di + ns, di + 2 * ns, di + 3 * ns, dvo);
    OBIT_HIP_TRY(hipGetLastError());
  }
  OBIT_HIP_TRY(hipStreamSynchronize(gs->stream));
  OBIT_HIP_TRY(hipMemcpy(iu, di, sizeof(int) * ns, hipMemcpyDeviceToHost));
  OBIT_HIP_TRY(hipMemcpy(iv, di + ns, sizeof(int) * ns, hipMemcpyDeviceToHost));
  OBIT_HIP_TRY(hipMemcpy(itu, di + 2 * ns, sizeof(int) * ns, hipMemcpyDeviceToHost));
  OBIT_HIP_TRY(hipMemcpy(itv, di + 3 * ns, sizeof(int) * ns, hipMemcpyDeviceToHost));
  OBIT_HIP_TRY(hipMemcpy(vis, dvo, sizeof(float) * 2 * ns, hipMemcpyDeviceToHost));
  (void)hipFree(dv); (void)hipFree(dvo); (void)hipFree(di);
  return obit::kOK;
}

int ObitHIPGridListStats(CUDAGridInfo *in, long iGPU, long long *stats) {
  GpuState *gs = obit::grid::get_gpu(in, iGPU, false);
  if (!gs || !stats) return obit::fail(obit::kState, "gridder not allocated");
  long long nonEmpty = 0, entries = 0, maxc = 0, multi = 0;
  std::vector<int> counts((size_t)gs->nlist);
  if (gs->nlist > 0 && gs->d_counts) {
    OBIT_HIP_TRY(hipStreamSynchronize(gs->stream));
    OBIT_HIP_TRY(hipMemcpy(counts.data(), gs->d_counts, sizeof(int) * gs->nlist, hipMemcpyDeviceToHost));
  }
  for (int c : counts) {
    if (c > 0) ++nonEmpty;
    entries += c;
    if (c > maxc) maxc = c;
    if (c > obit::grid::kChunk) ++multi;
  }
  stats[0] = gs->nlist;
  stats[1] = nonEmpty;
  stats[2] = entries;
  stats[3] = gs->nitems;
  stats[4] = maxc;
  stats[5] = multi;
  unsigned long long walk[2] = {0, 0};
  if (gs->d_walk) {
    OBIT_HIP_TRY(hipMemcpy(walk, gs->d_walk, sizeof(walk), hipMemcpyDeviceToHost));
  }
  stats[6] = (long long)walk[0];
  stats[7] = (long long)walk[1];
  return obit::kOK;
}

void *ObitHIPGridStream(CUDAGridInfo *in, long iGPU) {
  GpuState *gs = obit::grid::get_gpu(in, iGPU, false);
  return gs ? (void *)gs->stream : nullptr;
}

}  // extern "C"
