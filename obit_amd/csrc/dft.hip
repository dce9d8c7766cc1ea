// dft.hip -- DFT sky-model evaluate/subtract/divide/replace for gfx950.
//
// Replaces Obit's CUDA DFT layer (reference src/ObitCUDASkyModel.cu:73-963)
// behind the same C entry points (ObitCUDASkyModel.h) plus resident/batched
// extensions (ObitHIPVis.h).  Semantics follow Obit's CPU DFT
// (src/ObitSkyModel.c:2354-2857, src/ObitSkyModelMF.c:1971-2440).
//
// Numerics ("split precision"):
//   * phase psi_k = u x_k + v y_k + w z_k is formed in FP64 in turns per
//     reference-frequency wavelength (x_k = 2 pi l_k / 2 pi, exact from the
//     FP32 table), multiplied by nu/nu0 in FP64 and range-reduced to
//     [-1/2, 1/2] turns in FP64;
//   * one FP32 sincos for the phasor at the centre channel of the thread's
//     channel chunk and one for half the per-channel step angle
//     2 pi psi dnu/nu0, after the step is reduced to at most a quarter turn
//     (the half turns taken out flip the sign of odd channel offsets);
//   * the phasor walks out from the centre in both directions (at most
//     CPT/2 steps), each step a rotation written as three shears (Paeth:
//     x -= tan(d/2) y, y += sin(d) x, x -= tan(d/2) y), one FMA each, so
//     every channel costs three FMAs for the step and two for the
//     amplitude multiply-add; with |d| <= pi/2 the shear factors stay in
//     [-1, 1] and the rounding of tan and sin perturbs the angle, not the
//     modulus, by a relative few ulp;
//   * sums run in FP32 over a component tile and are folded into FP64
//     accumulators after every tile (two-level compensated summation).
//
// Work decomposition: a workgroup is 256 consecutive rows x one (IF,
// channel chunk); each thread owns one row and CPT channels in registers.
// Component tiles (TILE components) are staged in LDS by the workgroup
// and read as broadcasts.  The grid is persistent: gridDim.x workgroups
// stride over all (row block, IF, chunk) items.
#include "common.h"
#include "ObitCUDASkyModel.h"
#include "ObitCUDAUtil.h"
#include "ObitHIPVis.h"
#include "phasor.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace obit {
namespace dft {

constexpr int kRows = 256;  // rows (threads) per workgroup

enum Op { kSub = 0, kDiv = 1, kRepl = 2 };

// Kernel descriptor: plain values and device pointers only.
struct Desc {
  float *vis;
  long long nrow;
  int lrec, nrparm, ilocu, ilocv, ilocw;
  int incs, incf, incif;  // floats
  int chanb, nchsel, IFb, nIFsel, stokb, nstoksel;
  int nchunk, nrowblk;
  long long nitems;
  const double *fbase, *fdel;  // per selected IF: nu/nu0 at chanb and per channel
  const double *fchan;         // nu/nu0 of every selected channel [nIFsel][nchsel]
  int ncomp;
  const double *cx, *cy, *cz;  // turns per reference wavelength
  const float *camp;           // amplitude
  const float *cg;             // Gaussian coefficients [3][ncomp]
  const float *tab;            // spectral amplitudes [ncomp][nIFsel][nchpad]
  int tab_ldk, tab_nchpad;
  int op, doSwap;
  float stokFactor;
  // component parts: with nsplit > 1 an item takes one part of kpart
  // components and writes its FP64 sums to part ([nsplit][nrow][nIFsel]
  // [nchsel] complex) for k_dft_combine, instead of applying them
  int nsplit, kpart;
  double2 *part;
};

template <int PREC>
__device__ __forceinline__ void sincos_turns(float t, float *s, float *c) {
  if (PREC == 0) sincos_turns_hw(t, s, c);
  else sincos_turns_poly(t, s, c);
}

// Apply the model of one (row, channel) to the selected Stokes, following
// src/ObitSkyModel.c:2794-2844.
__device__ __forceinline__ void apply_op(const Desc &d, float *p, float mr, float mi) {
  if (d.doSwap) { const float t = mr; mr = -mi; mi = t; }
  float wm = 1.0f;
  if (d.op == kDiv) {
    wm = mr * mr + mi * mi;
    mr = mr / wm;
    mi = mi / wm;
    wm = sqrtf(wm);
  }
  for (int s = 0; s < d.nstoksel; ++s) {
    float *pp = p + s * d.incs;
    const float wt = pp[2];
    if (wt <= 0.0f && d.op != kRepl) continue;
    if (d.op == kDiv) {
      const float re = pp[0], im = pp[1];
      pp[0] = mr * re + mi * im;
      pp[1] = mr * im - mi * re;
      pp[2] = wt * wm;
    } else if (d.op == kRepl) {
      pp[0] = mr;
      pp[1] = mi;
    } else {
      pp[0] -= mr;
      pp[1] -= mi;
    }
    mr *= d.stokFactor;
    mi *= d.stokFactor;
  }
}

template <int CPT, int TILE, bool GAUSS, bool SPEC, int PREC>
__global__ __launch_bounds__(kRows) void k_dft_split(Desc d) {
  static_assert(CPT % 2 == 0, "CPT must be even");
  __shared__ double s_x[TILE], s_y[TILE], s_z[TILE];
  __shared__ float s_a[TILE];
  __shared__ float s_g[GAUSS ? 3 * TILE : 1];
  __shared__ float s_t[SPEC ? TILE * CPT : 1];
  const int tid = threadIdx.x;

  for (long long item = blockIdx.x; item < d.nitems; item += gridDim.x) {
    const int rb = (int)(item % d.nrowblk);
    const long long q = item / d.nrowblk;
    const int nic = d.nIFsel * d.nchunk;
    const int ic = (int)(q % nic);
    const int part = (int)(q / nic);
    const int kBeg = d.nsplit > 1 ? part * d.kpart : 0;
    const int kEnd = d.nsplit > 1 ? min(d.ncomp, kBeg + d.kpart) : d.ncomp;
    const int isel = ic / d.nchunk;
    const int c0 = (ic - isel * d.nchunk) * CPT;
    const long long row = (long long)rb * kRows + tid;
    const bool valid = row < d.nrow;
    const float *rp = d.vis + (valid ? row : 0LL) * (long long)d.lrec;
    const float uf = valid ? rp[d.ilocu] : 0.0f;
    const float vf = valid ? rp[d.ilocv] : 0.0f;
    const float wf = valid ? rp[d.ilocw] : 0.0f;
    const double u = uf, v = vf, w = wf;
    const double fd = d.fdel[isel];
    const double fm = d.fbase[isel] + (double)(c0 + CPT / 2) * fd;
    const float fmf = (float)fm, fdf = (float)fd;
    const float uu = uf * uf, vv = vf * vf, uv = uf * vf;

    float accr[CPT], acci[CPT];
    double outr[CPT], outi[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) { accr[j] = acci[j] = 0.0f; outr[j] = outi[j] = 0.0; }

    for (int k0 = kBeg; k0 < kEnd; k0 += TILE) {
      __syncthreads();
      for (int t = tid; t < TILE; t += kRows) {
        const int k = k0 + t;
        const bool ok = k < kEnd;
        s_x[t] = ok ? d.cx[k] : 0.0;
        s_y[t] = ok ? d.cy[k] : 0.0;
        s_z[t] = ok ? d.cz[k] : 0.0;
        s_a[t] = ok ? d.camp[k] : 0.0f;
        if (GAUSS) {
          s_g[3 * t + 0] = ok ? d.cg[k] : 0.0f;
          s_g[3 * t + 1] = ok ? d.cg[d.ncomp + k] : 0.0f;
          s_g[3 * t + 2] = ok ? d.cg[2 * d.ncomp + k] : 0.0f;
        }
      }
      if (SPEC) {
        for (int t = tid; t < TILE * CPT; t += kRows) {
          const int kk = t / CPT, cc = t - kk * CPT;
          const int k = k0 + kk;
          s_t[t] = (k < kEnd)
                       ? d.tab[(size_t)k * d.tab_ldk + (size_t)isel * d.tab_nchpad + c0 + cc]
                       : 0.0f;
        }
      }
      __syncthreads();
      const int kn = min(TILE, kEnd - k0);
      for (int kk = 0; kk < kn; ++kk) {
        const double psi = fma(w, s_z[kk], fma(v, s_y[kk], u * s_x[kk]));
        double pm = psi * fm, pd = psi * fd;
        pm -= rint(pm);
        // step reduced to |pd| <= 1/4 turn; an odd count of half turns
        // taken out makes every odd channel offset change sign
        const double h2 = rint(2.0 * pd);
        pd -= 0.5 * h2;
        const bool flip = ((int)h2) & 1;
        float sm, cm, sh, ch;
        sincos_turns<PREC == 2 ? 1 : 0>((float)pm, &sm, &cm);
        if (PREC == 0) sincos_turns_hw((float)(0.5 * pd), &sh, &ch);
        else sincos_halfturns_small((float)pd, &sh, &ch);
        // rotation by the step as three shears: tan and sin of the step
        const float tau = sh * __builtin_amdgcn_rcpf(ch);
        const float sig = 2.0f * sh * ch;
        const float a0 = s_a[kk];
        const float sgn = flip ? -1.0f : 1.0f;
        const float ae = a0, ao = flip ? -a0 : a0;  // amplitude at even / odd offsets
        float gf = a0, rf = 0.0f, rbk = 0.0f, qg = 0.0f;
        if (GAUSS) {
          const float G = fmaf(s_g[3 * kk + 2], uv, fmaf(s_g[3 * kk + 1], vv, s_g[3 * kk] * uu));
          const float e0 = G * fmf * fmf;
          const float Gd = G * fdf;
          const bool live = e0 > -60.0f;
          gf = live ? (SPEC ? 1.0f : a0) * __expf(e0) : 0.0f;
          // the per-step ratios carry the sign of the odd offsets
          rf = live ? sgn * __expf(Gd * (2.0f * fmf + fdf)) : 0.0f;
          rbk = live ? sgn * __expf(Gd * (fdf - 2.0f * fmf)) : 0.0f;
          qg = live ? __expf(2.0f * Gd * fdf) : 0.0f;
        }
        // forward from the centre channel
        float x = cm, y = sm, g = gf, r = rf;
#pragma unroll
        for (int j = CPT / 2; j < CPT; ++j) {
          const bool odd = (j - CPT / 2) & 1;
          float amp;
          if (GAUSS) amp = SPEC ? s_t[kk * CPT + j] * g : g;
          else if (SPEC) amp = odd ? s_t[kk * CPT + j] * sgn : s_t[kk * CPT + j];
          else amp = odd ? ao : ae;
          accr[j] = fmaf(amp, x, accr[j]);
          acci[j] = fmaf(amp, y, acci[j]);
          if (j + 1 < CPT) {
            const float x1 = fmaf(-tau, y, x);
            y = fmaf(sig, x1, y);
            x = fmaf(-tau, y, x1);
            if (GAUSS) { g *= r; r *= qg; }
          }
        }
        // backward from the centre channel
        x = cm; y = sm;
        float gb = gf * rbk, rb2 = rbk * qg;
#pragma unroll
        for (int j = CPT / 2 - 1; j >= 0; --j) {
          const bool odd = (CPT / 2 - j) & 1;
          const float x1 = fmaf(tau, y, x);
          y = fmaf(-sig, x1, y);
          x = fmaf(tau, y, x1);
          float amp;
          if (GAUSS) amp = SPEC ? s_t[kk * CPT + j] * gb : gb;
          else if (SPEC) amp = odd ? s_t[kk * CPT + j] * sgn : s_t[kk * CPT + j];
          else amp = odd ? ao : ae;
          accr[j] = fmaf(amp, x, accr[j]);
          acci[j] = fmaf(amp, y, acci[j]);
          if (GAUSS && j > 0) { gb *= rb2; rb2 *= qg; }
        }
      }
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        outr[j] += (double)accr[j];
        outi[j] += (double)acci[j];
        accr[j] = acci[j] = 0.0f;
      }
    }

    if (valid && d.part) {
      double2 *pp = d.part + (((long long)part * d.nrow + row) * d.nIFsel + isel) * d.nchsel + c0;
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        if (c0 + j < d.nchsel) pp[j] = make_double2(outr[j], outi[j]);
      }
    } else if (valid) {
      float *p = d.vis + row * (long long)d.lrec + d.nrparm + (d.IFb + isel) * d.incif +
                 (d.chanb + c0) * d.incf + d.stokb * d.incs;
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        if (c0 + j < d.nchsel) apply_op(d, p + j * d.incf, (float)outr[j], (float)outi[j]);
      }
    }
  }
}

// ---- the matrix-core DFT (point components) ----
//
// Channel c = c1 + 8 c2 (c1, c2 < 8) of a 64-channel group g (an IF's
// chunk of 64 selected channels) is at frequency f_g(36) + df (c1 - 4) +
// 8 df (c2 - 4), so a point component's term factors as
//   a e(psi f) = X[c1] Y_g[c2],  X[c1] = a W^(c1-4),  Y_g[c2] = P_g V^(c2-4),
//   W = e(psi df), V = e(8 psi df), P_g = e(psi f_g(36)),  e(t) = exp(2 pi i t),
// and one row's model over two groups is a complex product summed over
// components, M[c1][g, c2] = sum_k X_k[c1] Y_kg[c2]: a 16 x 16 real tile
// with rows (Re c1, Im c1), columns (2 groups x c2) and the sum over
// (component, Re / Im of Y):
//   A[c1][(k,0)] = Xr,  A[c1][(k,1)] = -Xi,  A[8+c1][(k,0)] = Xi,  A[8+c1][(k,1)] = Xr,
//   B[(k,0)][n] = Yr,  B[(k,1)][n] = Yi   ->   D[c1][n] = Re XY,  D[8+c1][n] = Im XY.
// It runs on v_mfma_f32_16x16x32_f16 (16 components a step) with every
// operand split into FP16 parts, x = x_h + x_l (x_h = fp16(x), x_l =
// fp16(x - x_h)), and the three products A_h B_h + A_h B_l + A_l B_h summed
// in FP32: the dropped A_l B_l and the rounding of the parts are below
// 2^-21 of a term.  X is scaled by 2^ashift (the table's largest
// amplitude to [2^13, 2^14), set when the model is loaded), so that the
// parts stay in the FP16 range; the scale is taken
// out exactly when the sums are written.
// The phases psi df, 8 psi df and psi f_g(36) are formed and reduced to
// [-1/2, 1/2] turns in FP64 (as k_dft_split forms its phases), their sines
// and cosines in FP32 (hardware or polynomial, by the precision setting),
// and the powers W^1..4, V^1..4 by at most two products (W^-k = conj W^k).
// Every IF must have the same channel step (the host checks it).
//
// A block is 4 waves, each with its own row and working on its own: lane
// j reads component j mod 32 of a tile straight from the table, a tile
// ahead.  For a tile, lane j of a wave forms, for component j mod 32, half
// of X (c1 = 4 (j / 32) ..) and Y of two groups (lanes 0..31 groups 0, 1;
// lanes 32..63 groups 2, 3), splits them and writes the parts to the
// wave's operand area.  The first half-wave forms W and its powers, the
// second V and its, and one lane exchange gives each lane both.  The wave
// then runs 2 steps x 2 output tiles x 3 products, each lane reading its 8
// FP16 values of A and B as one 16-byte load; the steps are interleaved
// with the next tile's set-up, and the FP32 sums are folded into FP64
// every 4 tiles.  3 waves per SIMD (168 registers) with 3 blocks per CU
// (LDS).
constexpr int kMfTile = 32;  // components a tile
// Operand area of a wave, in 16-byte slots (4 components x 2 FP16 values
// of one row): slot (b, r) of component block b = k / 4 and row r is
// b * R + r, with R = 17 (A: 16 rows + 1) or 33 (B: 2 tiles x 16 columns +
// 1); the odd block stride keeps the 4-byte writes of 32 consecutive
// components and the 16-byte reads of 16 consecutive rows on distinct
// banks.
constexpr int kMfRA = 17, kMfRB = 33;
constexpr int kMfSlotsA = 8 * kMfRA, kMfSlotsB = 8 * kMfRB;
constexpr int kMfSlots = 2 * kMfSlotsA + 2 * kMfSlotsB;  // A_h, A_l, B_h, B_l

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// e(t) for a phase in turns formed in FP64: reduced to [-1/2, 1/2] first,
// then the hardware sin / cos (HW) or the Taylor polynomial
template <bool HW>
__device__ __forceinline__ float2 eturn(double t) {
  float sn, cs;
  if (HW) sincos_turns_hw((float)(t - rint(t)), &sn, &cs);
  else sincos_turns_poly((float)(t - rint(t)), &sn, &cs);
  return make_float2(cs, sn);
}

// z^(k - 4) for k = 0..7 from z, by at most two products
__device__ __forceinline__ void powers_m4(float2 z, float2 (&o)[8]) {
  const float2 z2 = cmul(z, z), z3 = cmul(z2, z), z4 = cmul(z2, z2);
  o[0] = make_float2(z4.x, -z4.y);
  o[1] = make_float2(z3.x, -z3.y);
  o[2] = make_float2(z2.x, -z2.y);
  o[3] = make_float2(z.x, -z.y);
  o[4] = make_float2(1.0f, 0.0f);
  o[5] = z;
  o[6] = z2;
  o[7] = z3;
}

// a value of the first half-wave (lanes 0..31) and of the second
// (32..63) in every lane: lo[l] = x[l mod 32], hi[l] = x[32 + l mod 32]
// (v_permlane32_swap exchanges lanes 32..63 of its first operand with
// lanes 0..31 of its second)
__device__ __forceinline__ void swap_halves(float2 x, float2 &lo, float2 &hi) {
  const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(x.x), __float_as_uint(x.x), false, false);
  const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(x.y), __float_as_uint(x.y), false, false);
  lo = make_float2(__uint_as_float(a[0]), __uint_as_float(b[0]));
  hi = make_float2(__uint_as_float(a[1]), __uint_as_float(b[1]));
}

// FP16 parts of (x, y), each pair packed in a 32-bit word (x in the low
// half): hi = fp16(x), fp16(y) rounded to nearest; lo = fp16(x - hi.x),
// fp16(y - hi.y), the differences formed in FP32 from the FP16 values.
// Three instructions (the compiler's own sequence re-forms x - hi in
// FP32 with two conversions and two subtractions).
__device__ __forceinline__ void split_pair(float x, float y, unsigned &hi, unsigned &lo) {
  asm("v_cvt_pk_f16_f32 %0, %2, %3\n\t"
      "v_fma_mixlo_f16 %1, -%0, 1.0, %2 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %1, -%0, 1.0, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
      : "=&v"(hi), "=&v"(lo)
      : "v"(x), "v"(y));
}

// ngrp groups of 64 channels: group g = IF g / nchunk64, channels
// 64 (g mod nchunk64) .. (those past the selection formed, not written);
// d.fdel the same for every IF.
// PREC as k_dft_split's: 0 hardware sin / cos for every phasor, 1 hardware
// for the centre phasors P and the polynomial for the steps W and V, 2 the
// polynomial for all.
template <int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_dft_mfma(Desc d, int ngrp, int nchunk64, int ashift) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  typedef _Float16 h8 __attribute__((ext_vector_type(8)));
  __shared__ uint4 s_op[4][kMfSlots];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  uint4 *sAh = s_op[wv], *sAl = sAh + kMfSlotsA, *sBh = sAl + kMfSlotsA, *sBl = sBh + kMfSlotsB;
  // matrix-step roles: A row / B column i, component block qd of a step
  const int i = lane & 15, qd = lane >> 4;
  const int rdA = qd * kMfRA + i, rdB = qd * kMfRB + i;
  // writer roles: component kc (word kc & 3 of block kc / 4), half h
  const int kc = lane & 31, h = lane >> 5;
  unsigned *wA = reinterpret_cast<unsigned *>(sAh + (kc >> 2) * kMfRA + 4 * h) + (kc & 3);
  unsigned *wB = reinterpret_cast<unsigned *>(sBh + (kc >> 2) * kMfRB + 16 * h) + (kc & 3);
  constexpr int kLoA = 4 * kMfSlotsA, kLoB = 4 * kMfSlotsB;  // words from the _h to the _l array
  const double df = d.fdel[0];
  const long long nquad = (d.nrow + 3) / 4;
  const int ntg = (ngrp + 3) / 4;  // tile groups of up to 4 channel groups
  // items: (component part, row quad); with d.part set, a part's FP64 sums
  // go to the part scratch for k_dft_combine
  const int nsplit = max(d.nsplit, 1);
  for (long long it = blockIdx.x; it < nquad * nsplit; it += gridDim.x) {
    const int part = (int)(it / nquad);
    const long long rq = it - (long long)part * nquad;
    const int kBeg = part * d.kpart, kEnd = nsplit > 1 ? min(d.ncomp, kBeg + d.kpart) : d.ncomp;
    const long long row = rq * 4 + wv;
    const bool valid = row < d.nrow;
    const float *rp = d.vis + (valid ? row : 0LL) * (long long)d.lrec;
    const double u = valid ? rp[d.ilocu] : 0.0f, v = valid ? rp[d.ilocv] : 0.0f, w = valid ? rp[d.ilocw] : 0.0f;
    for (int tg = 0; tg < ntg; ++tg) {
      // this lane's two groups of the tile group and their centre
      // frequencies (a group past the last is worked on as the last one:
      // its columns are not written)
      const int gl0 = 4 * tg + 2 * h;
      double fc[2];
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int g = min(gl0 + hh, ngrp - 1);
        const int isel = g / nchunk64, ch = (g - isel * nchunk64) * 64 + 36;
        fc[hh] = d.fbase[isel] + (double)ch * d.fdel[isel];
      }
      double o0[4] = {0.0, 0.0, 0.0, 0.0}, o1[4] = {0.0, 0.0, 0.0, 0.0};
      // this lane's component of the next tile, loaded a tile ahead (index
      // clamped into the table, zero amplitude past its end)
      double nx = 0.0, ny = 0.0, nz = 0.0;
      float na = 0.0f;
      auto fetch = [&](int k0) {
        const int k = k0 + kc;
        const int kk = min(k, kEnd - 1);
        nx = d.cx[kk];
        ny = d.cy[kk];
        nz = d.cz[kk];
        const float am = d.camp[kk];
        na = k < kEnd ? am : 0.0f;
      };
      // a tile's set-up (the X and Y parts of its components into the
      // operand area)
      auto setup = [&](double cxk, double cyk, double czk, float a) {
        const float as = ldexpf(a, ashift);
        {
          const double psi = fma(w, czk, fma(v, cyk, u * cxk));
          // the first half forms W = e(psi df) and its powers, the second
          // V = e(8 psi df) and its; one lane exchange gives every lane both
          const float2 s1 = eturn<PREC == 0>(psi * (h ? 8.0 * df : df)), s2 = cmul(s1, s1), s3 = cmul(s2, s1),
                       s4 = cmul(s2, s2);
          float2 z1, z2, z3, z4, q1, q2, q3, q4;
          swap_halves(s1, z1, q1);
          swap_halves(s2, z2, q2);
          swap_halves(s3, z3, q3);
          swap_halves(s4, z4, q4);
          float2 vp[8];
          vp[0] = make_float2(q4.x, -q4.y);
          vp[1] = make_float2(q3.x, -q3.y);
          vp[2] = make_float2(q2.x, -q2.y);
          vp[3] = make_float2(q1.x, -q1.y);
          vp[4] = make_float2(1.0f, 0.0f);
          vp[5] = q1;
          vp[6] = q2;
          vp[7] = q3;
          // W^(c1 - 4) for this half's c1 (selects, not an indexed array)
          float2 wc[4];
          wc[0] = h ? make_float2(1.0f, 0.0f) : make_float2(z4.x, -z4.y);
          wc[1] = h ? z1 : make_float2(z3.x, -z3.y);
          wc[2] = h ? z2 : make_float2(z2.x, -z2.y);
          wc[3] = h ? z3 : make_float2(z1.x, -z1.y);
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            unsigned xh, xl;
            split_pair(as * wc[c].x, as * wc[c].y, xh, xl);
            // row c1 = 4h + c: (Xr, -Xi); row 8 + c1: (Xi, Xr)
            wA[4 * c] = xh ^ 0x80000000u;
            wA[4 * c + kLoA] = xl ^ 0x80000000u;
            wA[4 * (c + 8)] = __builtin_amdgcn_alignbit(xh, xh, 16);
            wA[4 * (c + 8) + kLoA] = __builtin_amdgcn_alignbit(xl, xl, 16);
          }
#pragma unroll
          for (int hh = 0; hh < 2; ++hh) {
            const float2 pg = eturn<PREC != 2>(psi * fc[hh]);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
              const float2 z = cmul(pg, vp[c]);
              unsigned yh, yl;
              split_pair(z.x, z.y, yh, yl);
              wB[4 * (8 * hh + c)] = yh;
              wB[4 * (8 * hh + c) + kLoB] = yl;
            }
          }
        }
      };
      if (kEnd > kBeg) {
        fetch(kBeg);
        const double cxk = nx, cyk = ny, czk = nz;
        const float a = na;
        fetch(kBeg + kMfTile);
        setup(cxk, cyk, czk, a);
      }
      // 4 tiles between folds into FP64 (tiles past the table have zero
      // amplitude); a tile's matrix steps run interleaved with the next
      // tile's set-up, whose writes follow the reads of this tile's
      // operands (a wave's LDS accesses complete in order)
      for (int k4 = kBeg; k4 < kEnd; k4 += 4 * kMfTile) {
        f4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
          const int k0 = k4 + kt * kMfTile;
          asm volatile("" ::: "memory");
          h8 ah[2], al[2], b0h[2], b0l[2], b1h[2], b1l[2];
#pragma unroll
          for (int s = 0; s < 2; ++s) {
            ah[s] = __builtin_bit_cast(h8, sAh[rdA + 4 * s * kMfRA]);
            al[s] = __builtin_bit_cast(h8, sAl[rdA + 4 * s * kMfRA]);
            b0h[s] = __builtin_bit_cast(h8, sBh[rdB + 4 * s * kMfRB]);
            b0l[s] = __builtin_bit_cast(h8, sBl[rdB + 4 * s * kMfRB]);
            b1h[s] = __builtin_bit_cast(h8, sBh[rdB + 4 * s * kMfRB + 16]);
            b1l[s] = __builtin_bit_cast(h8, sBl[rdB + 4 * s * kMfRB + 16]);
          }
          asm volatile("" ::: "memory");
          {
            const double cxk = nx, cyk = ny, czk = nz;
            const float a = na;
            fetch(k0 + 2 * kMfTile);
            setup(cxk, cyk, czk, a);
          }
#pragma unroll
          for (int s = 0; s < 2; ++s) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[s], b0h[s], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[s], b1h[s], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[s], b0l[s], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[s], b1l[s], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[s], b0h[s], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[s], b1h[s], acc1, 0, 0, 0);
          }
#pragma unroll
          for (int g = 0; g < 12; ++g) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 18, 0);
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          o0[r] += (double)acc0[r];
          o1[r] += (double)acc1[r];
        }
      }
      // lane l < 32 holds Re of rows 4 (l / 16) + r, lane l + 32 the Im of
      // the same c1; column n = l mod 16 is (group 2t + n / 8, c2 = n mod 8)
      double im0[4], im1[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        im0[r] = __shfl_xor(o0[r], 32);
        im1[r] = __shfl_xor(o1[r], 32);
      }
      if (valid && lane < 32) {
        const double unscale = ldexp(1.0, -ashift);
        const int n = lane & 15, c2 = n & 7;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int g = 4 * tg + 2 * t + (n >> 3);
          if (g >= ngrp) continue;
          const int isel = g / nchunk64, cbase = (g - isel * nchunk64) * 64;
          float *p = d.vis + row * (long long)d.lrec + d.nrparm + (d.IFb + isel) * d.incif +
                     (d.chanb + cbase) * d.incf + d.stokb * d.incs;
          double2 *pp = d.part ? d.part + (((long long)part * d.nrow + row) * d.nIFsel + isel) * d.nchsel + cbase
                               : nullptr;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int c1 = 4 * (lane >> 4) + r;
            const int c = c1 + 8 * c2;
            if (cbase + c >= d.nchsel) continue;  // past a ragged selection's last channel
            const double re = t ? o1[r] : o0[r], imv = t ? im1[r] : im0[r];
            if (pp) pp[c] = make_double2(re * unscale, imv * unscale);
            else apply_op(d, p + c * d.incf, (float)(re * unscale), (float)(imv * unscale));
          }
        }
      }
    }
  }
}

// The sums of the nsplit component parts of every (row, IF, channel),
// added in part order in FP64, applied to the row as k_dft_split applies
// its own.
__global__ __launch_bounds__(256) void k_dft_combine(Desc d) {
  const long long n = d.nrow * d.nIFsel * d.nchsel;
  const long long slab = n;  // entries of one part
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int ch = (int)(i % d.nchsel);
    const long long q = i / d.nchsel;
    const int isel = (int)(q % d.nIFsel);
    const long long row = q / d.nIFsel;
    double mr = 0.0, mi = 0.0;
    for (int s = 0; s < d.nsplit; ++s) {
      const double2 v = d.part[(long long)s * slab + i];
      mr += v.x;
      mi += v.y;
    }
    float *p = d.vis + row * (long long)d.lrec + d.nrparm + (d.IFb + isel) * d.incif +
               (d.chanb + ch) * d.incf + d.stokb * d.incs;
    apply_op(d, p, (float)mr, (float)mi);
  }
}

// Plain all-FP64 comparator: one thread per (row, channel); FP64 phase in
// radians, FP64 sincos and FP64 sums per evaluation.
template <bool GAUSS, bool SPEC>
__global__ __launch_bounds__(kRows) void k_dft_fp64(Desc d) {
  constexpr int T = 64;
  __shared__ double s_x[T], s_y[T], s_z[T], s_a[T], s_g[3 * T];
  const int tid = threadIdx.x;
  const long long nitems = (long long)d.nrowblk * d.nIFsel * d.nchsel;
  for (long long item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int rb = (int)(item % d.nrowblk);
    const int ic = (int)(item / d.nrowblk);
    const int isel = ic / d.nchsel;
    const int cs = ic - isel * d.nchsel;
    const long long row = (long long)rb * kRows + tid;
    const bool valid = row < d.nrow;
    const float *rp = d.vis + (valid ? row : 0LL) * (long long)d.lrec;
    // the channel's own frequency: this kernel also takes tables that are
    // not a line within an IF
    const double f = d.fchan[(size_t)isel * d.nchsel + cs];
    const double u = (double)(valid ? rp[d.ilocu] : 0.0f) * f;
    const double v = (double)(valid ? rp[d.ilocv] : 0.0f) * f;
    const double w = (double)(valid ? rp[d.ilocw] : 0.0f) * f;
    const double twopi = 6.283185307179586476925;
    double sr = 0.0, si = 0.0;
    for (int k0 = 0; k0 < d.ncomp; k0 += T) {
      __syncthreads();
      for (int t = tid; t < T; t += kRows) {
        const int k = k0 + t;
        const bool ok = k < d.ncomp;
        s_x[t] = ok ? d.cx[k] * twopi : 0.0;
        s_y[t] = ok ? d.cy[k] * twopi : 0.0;
        s_z[t] = ok ? d.cz[k] * twopi : 0.0;
        double a = ok ? (double)d.camp[k] : 0.0;
        if (SPEC && ok) a = d.tab[(size_t)k * d.tab_ldk + (size_t)isel * d.tab_nchpad + cs];
        s_a[t] = a;
        if (GAUSS) {
          s_g[3 * t + 0] = ok ? (double)d.cg[k] : 0.0;
          s_g[3 * t + 1] = ok ? (double)d.cg[d.ncomp + k] : 0.0;
          s_g[3 * t + 2] = ok ? (double)d.cg[2 * d.ncomp + k] : 0.0;
        }
      }
      __syncthreads();
      const int kn = min(T, d.ncomp - k0);
      for (int kk = 0; kk < kn; ++kk) {
        double amp = s_a[kk];
        if (GAUSS) amp *= exp(s_g[3 * kk] * u * u + s_g[3 * kk + 1] * v * v + s_g[3 * kk + 2] * u * v);
        const double ph = s_x[kk] * u + s_y[kk] * v + s_z[kk] * w;
        double s, c;
        sincos(ph, &s, &c);
        sr += amp * c;
        si += amp * s;
      }
    }
    if (valid) {
      float *p = d.vis + row * (long long)d.lrec + d.nrparm + (d.IFb + isel) * d.incif +
                 (d.chanb + cs) * d.incf + d.stokb * d.incs;
      apply_op(d, p, (float)sr, (float)si);
    }
  }
}

// Convert Obit's component table (rows of 'size' floats) into the
// resident layout: positions in turns (FP64), amplitude, Gaussian terms.
// Table layouts: src/ObitSkyModel.c:1965-2043, src/ObitSkyModelMF.c:1480-1514.
__global__ void k_prep_comps(const float *model, int size, int n, int gaussOff, double *cx,
                             double *cy, double *cz, float *camp, float *cg) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float *r = model + (size_t)k * size;
  const double inv2pi = 0.15915494309189533577;
  camp[k] = r[0];
  cx[k] = (double)r[1] * inv2pi;
  cy[k] = (double)r[2] * inv2pi;
  cz[k] = (double)r[3] * inv2pi;
  if (gaussOff >= 0) {
    cg[k] = r[gaussOff];
    cg[n + k] = r[gaussOff + 1];
    cg[2 * n + k] = r[gaussOff + 2];
  }
}

// Largest |amplitude| of n components into out[0] (one block).
__global__ __launch_bounds__(256) void k_amp_max(const float *camp, int n, float *out) {
  __shared__ float s[256];
  float m = 0.0f;
  for (int k = threadIdx.x; k < n; k += 256) m = fmaxf(m, fabsf(camp[k]));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// Spectral amplitude per (component, selected IF, selected channel) in
// FP64, stored FP32.  Parametrised (type Spec): a*exp(sum_t c_t L^(t+1)),
// L = ln(nu/nu_ref), coefficients at coefOff (src/ObitSkyModel.c:2528-2552,
// CPU sign).  Tabulated (TSpec): S_bin*exp(alpha*L), L = ln(nu/nu_bin),
// alpha at coefOff and S at coefOff+1+bin (src/ObitSkyModelMF.c:2143-2172).
__global__ void k_prep_spec(const float *model, int size, int n, int tabulated, int coefOff,
                            int nterm, const double *lnr, const int *bin, int nIFsel,
                            int nchsel, int nchpad, float *tab) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = (long long)nIFsel * nchpad;
  if (idx >= (long long)n * per) return;
  const int k = (int)(idx / per);
  const int rem = (int)(idx - (long long)k * per);
  const int isel = rem / nchpad, cs = rem - isel * nchpad;
  float out = 0.0f;
  if (cs < nchsel) {
    const float *r = model + (size_t)k * size;
    const double L = lnr[isel * nchsel + cs];
    if (tabulated) {
      const int b = bin[isel * nchsel + cs];
      out = (float)((double)r[coefOff + 1 + b] * exp((double)r[coefOff] * L));
    } else {
      double a = 0.0, Lp = L;
      for (int t = 0; t < nterm; ++t) { a += (double)r[coefOff + t] * Lp; Lp *= L; }
      out = (float)((double)r[0] * exp(a));
    }
  }
  tab[idx] = out;
}

// ---------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------

using KernelFn = void (*)(Desc);

template <int CPT, int TILE, bool G, bool S, int P>
KernelFn pick1() { return k_dft_split<CPT, TILE, G, S, P>; }

template <int CPT, int TILE, int P>
KernelFn pick_gs(bool g, bool s) {
  if (g) return s ? pick1<CPT, TILE, true, true, P>() : pick1<CPT, TILE, true, false, P>();
  return s ? pick1<CPT, TILE, false, true, P>() : pick1<CPT, TILE, false, false, P>();
}

template <int P>
KernelFn pick_shape(int cpt, int tile, bool g, bool s) {
  if (cpt == 8) return pick_gs<8, 64, P>(g, s);  // small buffers: more work items
  if (cpt == 16) {
    if (tile == 32) return pick_gs<16, 32, P>(g, s);
    if (tile == 128) return pick_gs<16, 128, P>(g, s);
    return pick_gs<16, 64, P>(g, s);
  }
  if (tile == 32) return pick_gs<32, 32, P>(g, s);
  if (tile == 128) return pick_gs<32, 128, P>(g, s);
  return pick_gs<32, 64, P>(g, s);
}

KernelFn pick_split(int cpt, int tile, bool g, bool s, int prec) {
  if (prec == 0) return pick_shape<0>(cpt, tile, g, s);
  if (prec == 2) return pick_shape<2>(cpt, tile, g, s);
  return pick_shape<1>(cpt, tile, g, s);
}

KernelFn pick_fp64(bool g, bool s) {
  if (g) return s ? k_dft_fp64<true, true> : k_dft_fp64<true, false>;
  return s ? k_dft_fp64<false, true> : k_dft_fp64<false, false>;
}

struct State {
  int device = 0;
  hipStream_t stream = nullptr;
  int nCU = 256;
  // model
  bool modelReady = false;
  int type = 0, nterm = 0, size = 0, ncomp = 0;
  double *cx = nullptr, *cy = nullptr, *cz = nullptr;
  float *camp = nullptr, *cg = nullptr;
  int capComp = 0;
  int ampShift = 0;          // k_dft_mfma's amplitude scale exponent
  // Gaussian tables: a copy ordered points first (mx, my, mz, mamp), its
  // npoint leading components without Gaussian terms, and the Gaussian
  // terms of the rest ([3][ncomp - npoint]), for the mixed launch
  int npoint = 0;
  double *mx = nullptr, *my = nullptr, *mz = nullptr;
  float *mamp = nullptr, *cgG = nullptr;
  void *mbuf = nullptr;
  size_t mbufCap = 0;
  float *dAmax = nullptr;    // largest |amplitude| of the table (device)
  const float *modelDev = nullptr;  // caller's table, for the spectral prep
  // spectral table
  float *tab = nullptr;
  size_t tabCap = 0;
  bool tabReady = false;
  long long tabKey = -1;
  // frequencies
  std::vector<double> freq64;
  double uvRef = 0.0;
  std::vector<double> specRef;
  std::vector<float> fscaleHost;  // cached copy of freqScale ...
  const float *fscalePtr = nullptr;        // ... read from this device table
  unsigned long long fscaleGen = ~0ULL;    // ... after this many table uploads
  bool fallbackNoted = false;     // the FP64 fallback message was printed
  int lastPath = 0;               // kernel of the last launch: 0 split, 1 FP64
  int lastCpt = 0, lastSplit = 0;  // shape of the last per-buffer launch
  double *dfreq = nullptr;        // [2*nIFsel]: fbase, fdel
  double *hfreq = nullptr;        // pinned staging
  int freqCap = 0;
  bool uniform = true;
  // resident data
  float *dvis = nullptr;
  long long nres = 0;
  bool ownVis = false;
  float *stage = nullptr;
  size_t stageCap = 0;
  double2 *part = nullptr;  // per-part sums of the per-buffer path
  size_t partCap = 0;
  std::vector<hipEvent_t> events;                 // per-buffer segment events
  std::unordered_map<const void *, int> occ;      // resident workgroups per CU of a kernel
  // kernel shape
  int cpt = 32, tile = 64, wgPerCU = 2, prec = 1;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long long graphKey = 0;
};

std::mutex g_mu;
std::unordered_map<const GPUInfo *, State *> g_states;

State *find_state(const GPUInfo *g) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_states.find(g);
  return it == g_states.end() ? nullptr : it->second;
}

int model_gauss_off(int type) {
  return (type == GPUModTypeGauss || type == GPUModTypeGaussSpec || type == GPUModTypeGaussTSpec) ? 4 : -1;
}
bool model_is_spec(int type) {
  return type == GPUModTypePointSpec || type == GPUModTypePointTSpec ||
         type == GPUModTypeGaussSpec || type == GPUModTypeGaussTSpec;
}
bool model_is_tab(int type) { return type == GPUModTypePointTSpec || type == GPUModTypeGaussTSpec; }
int model_coef_off(int type) {
  switch (type) {
    case GPUModTypePointSpec: return 4;
    case GPUModTypeGaussSpec: return 7;
    case GPUModTypePointTSpec: return 4;
    case GPUModTypeGaussTSpec: return 7;
    default: return -1;
  }
}

void destroy_graph(State *st) {
  if (st->gexec) { (void)hipGraphExecDestroy(st->gexec); st->gexec = nullptr; }
  if (st->graph) { (void)hipGraphDestroy(st->graph); st->graph = nullptr; }
  st->graphKey = 0;
}

// Host copy of the FP32 freqScale table: read again when it is another
// table, has been uploaded again through the copy utilities, is longer than
// the cached copy, or always.
int refresh_fscale(State *st, const GPUVisInfo *vi, bool always) {
  const int nfreq = vi->nchan * vi->nIF;
  const unsigned long long gen = upload_generation();
  if (always || st->fscaleHost.size() < (size_t)nfreq || st->fscalePtr != vi->freqScale ||
      st->fscaleGen != gen) {
    if (!vi->freqScale) return fail(kInvalid, "no frequency information (freqScale is NULL)");
    st->fscaleHost.resize(nfreq);
    OBIT_HIP_TRY(hipMemcpy(st->fscaleHost.data(), vi->freqScale, sizeof(float) * nfreq, hipMemcpyDeviceToHost));
    st->fscalePtr = vi->freqScale;
    st->fscaleGen = gen;
  }
  return kOK;
}

// Frequency of every selected channel relative to nu0, as a line per IF:
// f(c) = fbase + (c - chanb) * fdel.  Exact FP64 values when the caller
// gave them, else a least-squares line through the FP32 freqScale.
int frequency_lines(State *st, const GPUVisInfo *vi, std::vector<double> &fb,
                    std::vector<double> &fdl, std::vector<double> *perChan) {
  const int nchsel = vi->chane - vi->chanb + 1, nIFsel = vi->IFe - vi->IFb + 1;
  fb.assign(nIFsel, 0.0);
  fdl.assign(nIFsel, 0.0);
  if (perChan) perChan->assign((size_t)nIFsel * nchsel, 0.0);
  const int nfreq = vi->nchan * vi->nIF;
  std::vector<double> f(nfreq);
  if (!st->freq64.empty() && st->uvRef > 0.0) {
    if ((int)st->freq64.size() < nfreq)
      return fail(kInvalid, "frequency table has %zu values, need %d", st->freq64.size(), nfreq);
    for (int i = 0; i < nfreq; ++i) f[i] = st->freq64[i] / st->uvRef;
  } else {
    const int rc = refresh_fscale(st, vi, false);
    if (rc) return rc;
    for (int i = 0; i < nfreq; ++i) f[i] = st->fscaleHost[i];
  }
  st->uniform = true;
  for (int is = 0; is < nIFsel; ++is) {
    const int IF = vi->IFb + is;
    std::vector<double> y(nchsel);
    for (int c = 0; c < nchsel; ++c) y[c] = f[IF * vi->kincif + (vi->chanb + c) * vi->kincf];
    double a = y[0], b = 0.0;
    if (nchsel > 1) {
      if (!st->freq64.empty() && st->uvRef > 0.0) {
        b = (y[nchsel - 1] - y[0]) / (nchsel - 1);
      } else {  // least squares through FP32 values
        double sx = 0, sy = 0, sxx = 0, sxy = 0;
        for (int c = 0; c < nchsel; ++c) { sx += c; sy += y[c]; sxx += (double)c * c; sxy += c * y[c]; }
        const double n = nchsel;
        b = (n * sxy - sx * sy) / (n * sxx - sx * sx);
        a = (sy - b * sx) / n;
      }
      for (int c = 0; c < nchsel; ++c) {
        const double dev = std::fabs(y[c] - (a + b * c));
        const double tol = (!st->freq64.empty() && st->uvRef > 0.0) ? 1e-12 * std::fabs(y[c])
                                                                    : 4e-7 * std::fabs(y[c]);
        if (dev > tol) st->uniform = false;
      }
    }
    fb[is] = a;
    fdl[is] = b;
    if (perChan)
      for (int c = 0; c < nchsel; ++c) (*perChan)[(size_t)is * nchsel + c] = y[c];
  }
  return kOK;
}

// Build (or reuse) the spectral amplitude table for the current selection.
int ensure_spec_table(State *st, const GPUInfo *gi) {
  const GPUVisInfo *vi = gi->h_visInfo;
  const GPUModelInfo *mi = gi->h_modelInfo;
  const int nchsel = vi->chane - vi->chanb + 1, nIFsel = vi->IFe - vi->IFb + 1;
  const int nchpad = (nchsel + 63) / 64 * 64;
  const long long key = ((long long)vi->chanb << 40) ^ ((long long)vi->chane << 28) ^
                        ((long long)vi->IFb << 16) ^ (long long)vi->IFe;
  if (st->tabReady && st->tabKey == key) return kOK;
  std::vector<double> fb, fdl, fch;
  int rc = frequency_lines(st, vi, fb, fdl, &fch);
  if (rc) return rc;
  const bool tabulated = model_is_tab(st->type);
  const int coefOff = model_coef_off(st->type);
  if (!tabulated && coefOff + st->nterm > st->size)
    return fail(kInvalid, "%d spectral terms do not fit a %d-float component row", st->nterm, st->size);
  std::vector<double> lnr((size_t)nIFsel * nchsel);
  std::vector<int> bin((size_t)nIFsel * nchsel, 0);
  for (int is = 0; is < nIFsel; ++is) {
    for (int c = 0; c < nchsel; ++c) {
      const int IF = vi->IFb + is, ch = vi->chanb + c;
      const int ifq = IF * vi->kincif + ch * vi->kincf;
      const double fr = fch[(size_t)is * nchsel + c];  // nu/nu_uv
      double ratio;                                     // nu/nu_spectrum
      if (tabulated) {
        const int b = mi->specIndex ? mi->specIndex[ifq] : 0;
        if (b < 0 || coefOff + 1 + b >= st->size)
          return fail(kInvalid, "coarse bin %d outside the %d-float component row", b, st->size);
        bin[(size_t)is * nchsel + c] = b;
        if ((int)st->specRef.size() > b && st->uvRef > 0.0) ratio = fr * st->uvRef / st->specRef[b];
        else if (vi->freqRat) ratio = fr * (double)vi->freqRat[b];
        else return fail(kInvalid, "tabulated spectrum needs reference frequencies");
      } else {
        if (!st->specRef.empty() && st->uvRef > 0.0) ratio = fr * st->uvRef / st->specRef[0];
        else if (vi->freqRat) ratio = fr * (double)vi->freqRat[0];
        else ratio = fr;
      }
      lnr[(size_t)is * nchsel + c] = std::log(ratio);
    }
  }
  const size_t need = (size_t)st->ncomp * nIFsel * nchpad;
  if (need > st->tabCap) {
    destroy_graph(st);  // a captured launch reads the old table
    if (st->tab) OBIT_HIP_TRY(hipFree(st->tab));
    st->tab = nullptr;
    hipError_t e = hipMalloc(&st->tab, need * sizeof(float));
    if (e != hipSuccess) return fail(kNoMemory, "spectral table (%zu floats): %s", need, hipGetErrorString(e));
    st->tabCap = need;
  }
  double *dl = nullptr;
  int *db = nullptr;
  OBIT_HIP_TRY(hipMalloc(&dl, lnr.size() * sizeof(double)));
  OBIT_HIP_TRY(hipMalloc(&db, bin.size() * sizeof(int)));
  OBIT_HIP_TRY(hipMemcpy(dl, lnr.data(), lnr.size() * sizeof(double), hipMemcpyHostToDevice));
  OBIT_HIP_TRY(hipMemcpy(db, bin.data(), bin.size() * sizeof(int), hipMemcpyHostToDevice));
  const long long total = (long long)st->ncomp * nIFsel * nchpad;
  if (total > 0) {
    const int nb = (int)((total + 255) / 256);
    hipLaunchKernelGGL(k_prep_spec, dim3(nb), dim3(256), 0, st->stream, st->modelDev, st->size,
                       st->ncomp, tabulated ? 1 : 0, model_coef_off(st->type), st->nterm, dl, db,
                       nIFsel, nchsel, nchpad, st->tab);
    OBIT_HIP_TRY(hipGetLastError());
  }
  OBIT_HIP_TRY(hipStreamSynchronize(st->stream));
  OBIT_HIP_TRY(hipFree(dl));
  OBIT_HIP_TRY(hipFree(db));
  st->tabReady = true;
  st->tabKey = key;
  return kOK;
}

// Fill a kernel descriptor for nrow rows at vis; uploads the per-IF lines.
// The layout, selection and operation fields of a kernel descriptor for
// nrow rows at vis (no allocation, no frequencies, no model pointers).
int desc_layout(const GPUInfo *gi, float *vis, long long nrow, int cpt, Desc &d) {
  const GPUVisInfo *vi = gi->h_visInfo;
  const GPUModelInfo *mi = gi->h_modelInfo;
  std::memset(&d, 0, sizeof(d));
  d.vis = vis;
  d.nrow = nrow;
  d.lrec = vi->lenvis;
  d.nrparm = vi->nrparm;
  d.ilocu = vi->ilocu; d.ilocv = vi->ilocv; d.ilocw = vi->ilocw;
  d.incs = 3 * vi->incs; d.incf = 3 * vi->incf; d.incif = 3 * vi->incif;
  d.chanb = vi->chanb; d.nchsel = vi->chane - vi->chanb + 1;
  d.IFb = vi->IFb; d.nIFsel = vi->IFe - vi->IFb + 1;
  d.stokb = vi->stokb; d.nstoksel = vi->stoke - vi->stokb + 1;
  if (d.nchsel < 1 || d.nIFsel < 1 || d.nstoksel < 1)
    return fail(kInvalid, "empty channel/IF/Stokes selection");
  if (d.chanb < 0 || vi->chane >= vi->nchan || d.IFb < 0 || vi->IFe >= vi->nIF || d.stokb < 0 ||
      vi->stoke >= vi->nstok)
    return fail(kInvalid, "selection outside the data axes");
  if (vi->lenvis != gi->lenvis || d.lrec <= 0)
    return fail(kInvalid, "lenvis mismatch (%d vs %d)", vi->lenvis, gi->lenvis);
  {
    // the last correlation touched must lie inside the row
    const long long last = (long long)d.nrparm + (long long)vi->IFe * d.incif +
                           (long long)vi->chane * d.incf + (long long)vi->stoke * d.incs + 2;
    const long long uvwMax = std::max(d.ilocu, std::max(d.ilocv, d.ilocw));
    if (last >= d.lrec || uvwMax >= d.lrec || std::min(d.ilocu, std::min(d.ilocv, d.ilocw)) < 0)
      return fail(kInvalid, "uv layout exceeds the %d-float row", d.lrec);
  }
  d.nchunk = (d.nchsel + cpt - 1) / cpt;
  d.nrowblk = (int)((nrow + kRows - 1) / kRows);
  d.nitems = (long long)d.nrowblk * d.nIFsel * d.nchunk;
  d.op = (int)mi->opType;
  if (d.op < 0 || d.op > 2) return fail(kInvalid, "unknown operation %d", d.op);
  d.doSwap = mi->doSwap;
  d.stokFactor = mi->stokFact[1];
  return kOK;
}

// Fill a kernel descriptor for nrow rows at vis; uploads the per-IF lines.
int make_desc(State *st, const GPUInfo *gi, float *vis, long long nrow, int cpt, Desc &d) {
  const GPUVisInfo *vi = gi->h_visInfo;
  const GPUModelInfo *mi = gi->h_modelInfo;
  if (!vi || !mi) return fail(kState, "h_visInfo / h_modelInfo not set");
  if (!st->modelReady) return fail(kState, "no model: call ObitCUDASkyModelDFTSetMod first");
  int rc = desc_layout(gi, vis, nrow, cpt, d);
  if (rc) return rc;
  // per-IF frequency lines, then every selected channel's own value
  std::vector<double> fb, fdl, fch;
  rc = frequency_lines(st, vi, fb, fdl, &fch);
  if (rc) return rc;
  const int nfq = 2 * d.nIFsel + d.nIFsel * d.nchsel;
  if (nfq > st->freqCap) {
    destroy_graph(st);  // a captured launch reads the old table
    if (st->dfreq) OBIT_HIP_TRY(hipFree(st->dfreq));
    if (st->hfreq) OBIT_HIP_TRY(hipHostFree(st->hfreq));
    st->dfreq = nullptr;
    st->hfreq = nullptr;
    OBIT_HIP_TRY(hipMalloc(&st->dfreq, sizeof(double) * nfq));
    OBIT_HIP_TRY(hipHostMalloc(&st->hfreq, sizeof(double) * nfq, hipHostMallocDefault));
    st->freqCap = nfq;
  }
  OBIT_HIP_TRY(hipStreamSynchronize(st->stream));  // staging buffer reuse
  for (int i = 0; i < d.nIFsel; ++i) { st->hfreq[i] = fb[i]; st->hfreq[d.nIFsel + i] = fdl[i]; }
  std::memcpy(st->hfreq + 2 * d.nIFsel, fch.data(), sizeof(double) * fch.size());
  OBIT_HIP_TRY(hipMemcpyAsync(st->dfreq, st->hfreq, sizeof(double) * nfq, hipMemcpyHostToDevice, st->stream));
  d.fbase = st->dfreq;
  d.fdel = st->dfreq + d.nIFsel;
  d.fchan = st->dfreq + 2 * d.nIFsel;
  d.ncomp = st->ncomp;
  d.cx = st->cx; d.cy = st->cy; d.cz = st->cz;
  d.camp = st->camp; d.cg = st->cg;
  if (model_is_spec(st->type)) {
    rc = ensure_spec_table(st, gi);
    if (rc) return rc;
    d.tab = st->tab;
    d.tab_nchpad = (d.nchsel + 63) / 64 * 64;
    d.tab_ldk = d.nIFsel * d.tab_nchpad;
  }
  return kOK;
}

// Whether k_dft_mfma can run this descriptor: point components and one
// channel step for every IF (within 1e-12; the kernel uses the first
// IF's).  A selection that is not whole 64-channel groups runs its last
// group in full and writes only the selected channels.  make_desc has left
// the per-IF lines in st->hfreq.
bool channels_ok(const State *st, const Desc &d) {
  if (d.nchsel < 1 || !st->hfreq) return false;
  const double d0 = st->hfreq[d.nIFsel];
  for (int i = 1; i < d.nIFsel; ++i)
    if (std::fabs(st->hfreq[d.nIFsel + i] - d0) > 1e-12 * std::fabs(d0)) return false;
  return true;
}

bool mfma_ok(const State *st, const Desc &d) {
  return model_gauss_off(st->type) < 0 && !model_is_spec(st->type) && channels_ok(st, d);
}

// A Gaussian table (no spectra) with npoint components without Gaussian
// terms (SetMod keeps a copy of the table ordered points first),
// subtracted over whole rows: the points run on k_dft_mfma and the
// Gaussians on the VALU kernel, each launch subtracting its own part of the
// model.
bool mixed_ok(const State *st, const Desc &d) {
  return st->type == GPUModTypeGauss && st->npoint > 0 && d.op == kSub && !d.part && channels_ok(st, d);
}

typedef void (*MfmaFn)(Desc, int, int, int);
MfmaFn mfma_fn(const State *st) {
  return st->prec == 0 ? k_dft_mfma<0> : st->prec == 2 ? k_dft_mfma<2> : k_dft_mfma<1>;
}

// Resident workgroups per CU of a kernel (cached).
int occupancy_of(State *st, const void *fn) {
  auto it = st->occ.find(fn);
  if (it != st->occ.end()) return it->second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kRows, 0) != hipSuccess || nb < 1) nb = 1;
  st->occ[fn] = nb;
  return nb;
}

// wgPerCU > 0: a persistent grid of wgPerCU workgroups per CU striding over
// the items; wgPerCU = 0: one workgroup per item, placed by the dispatcher
// as CUs free up.  cpt = 64 selects k_dft_mfma where mfma_ok allows it (a
// persistent grid of the resident workgroups over row quads), else 32.
// k_dft_mfma over d's rows and components (its own items: component
// parts x row quads), a persistent grid of its resident workgroups.  The
// caller has checked the channels (channels_ok) and that d's components are
// points.
int launch_mfma(State *st, const Desc &d, hipStream_t stream) {
  const long long nquad = (d.nrow + 3) / 4;
  MfmaFn fn = mfma_fn(st);
  const long long grid =
      std::min<long long>(nquad * std::max(d.nsplit, 1), (long long)st->nCU * occupancy_of(st, (const void *)fn));
  const int nchunk64 = (d.nchsel + 63) / 64;
  hipLaunchKernelGGL(fn, dim3((unsigned)std::max<long long>(grid, 1)), dim3(kRows), 0,
                     stream ? stream : st->stream, d, d.nIFsel * nchunk64, nchunk64, st->ampShift);
  OBIT_HIP_TRY(hipGetLastError());
  return kOK;
}

// The points-first copy of a Gaussian table (SetMod): its points and its
// Gaussians as two descriptors over d's rows.
void mixed_descs(const State *st, const Desc &d, Desc &dp, Desc &dg) {
  dp = d;
  dp.cx = st->mx;
  dp.cy = st->my;
  dp.cz = st->mz;
  dp.camp = st->mamp;
  dp.ncomp = st->npoint;
  dg = d;
  dg.cx = st->mx + st->npoint;
  dg.cy = st->my + st->npoint;
  dg.cz = st->mz + st->npoint;
  dg.camp = st->mamp + st->npoint;
  dg.cg = st->cgG;
  dg.ncomp = d.ncomp - st->npoint;
  dg.nchunk = (d.nchsel + 31) / 32;
}

int launch_split(State *st, const Desc &d, int cpt, int tile, int wgPerCU, hipStream_t stream = nullptr) {
  if (d.nitems <= 0) return kOK;
  if (cpt == 64) {
    if (mfma_ok(st, d)) return launch_mfma(st, d, stream);
    if (mixed_ok(st, d)) {
      Desc dp, dg;
      mixed_descs(st, d, dp, dg);
      int rc = launch_mfma(st, dp, stream);
      if (rc || st->npoint == d.ncomp) return rc;
      dg.nitems = (long long)d.nrowblk * d.nIFsel * dg.nchunk;
      return launch_split(st, dg, 32, tile, wgPerCU, stream);
    }
    // the VALU kernel at 32 channels a thread, its items counted for that
    Desc d32 = d;
    d32.nchunk = (d.nchsel + 31) / 32;
    d32.nitems = (long long)d.nrowblk * d.nIFsel * d32.nchunk * std::max(d.nsplit, 1);
    return launch_split(st, d32, 32, tile, wgPerCU, stream);
  }
  const bool g = model_gauss_off(st->type) >= 0, s = model_is_spec(st->type);
  KernelFn fn = pick_split(cpt, tile, g, s, st->prec);
  long long grid = wgPerCU > 0 ? (long long)st->nCU * wgPerCU : d.nitems;
  if (grid > d.nitems) grid = d.nitems;
  if (grid > INT_MAX) return fail(kInvalid, "%lld work items exceed one launch", d.nitems);
  Desc dd = d;
  void *args[] = {&dd};
  OBIT_HIP_TRY(hipLaunchKernel((const void *)fn, dim3((unsigned)grid), dim3(kRows), args, 0,
                               stream ? stream : st->stream));
  return kOK;
}

int launch_fp64(State *st, const Desc &d) {
  const long long nitems = (long long)d.nrowblk * d.nIFsel * d.nchsel;
  if (nitems <= 0) return kOK;
  const bool g = model_gauss_off(st->type) >= 0, s = model_is_spec(st->type);
  long long grid = (long long)st->nCU * 4;
  if (grid > nitems) grid = nitems;
  Desc dd = d;
  void *args[] = {&dd};
  OBIT_HIP_TRY(hipLaunchKernel((const void *)pick_fp64(g, s), dim3((unsigned)grid), dim3(kRows), args, 0,
                               st->stream));
  return kOK;
}

// Run the operation over nrow rows at a device pointer.
int run_device(State *st, const GPUInfo *gi, float *vis, long long nrow, bool fp64) {
  Desc d;
  int rc = make_desc(st, gi, vis, nrow, st->cpt, d);
  if (rc) return rc;
  if (!fp64 && !st->uniform && !st->fallbackNoted) {
    // the split kernel walks each IF's channels as a line; a table off a
    // line goes to the FP64 kernel, which is many times slower
    std::fprintf(stderr,
                 "obit_hip: channel frequencies are not uniform within an IF; the DFT runs on the "
                 "FP64 kernel (ObitHIPSkyModelDFTGetPath reports 1)\n");
    st->fallbackNoted = true;
  }
  st->lastPath = (fp64 || !st->uniform) ? 1 : 0;
  if (fp64 || !st->uniform) return launch_fp64(st, d);
  return launch_split(st, d, st->cpt, st->tile, st->wgPerCU);
}

// ---- the per-buffer launch plan (ObitCUDASkyModelDFTCalc) ----

int occupancy(State *st, KernelFn fn) { return occupancy_of(st, (const void *)fn); }

struct BufPlan {
  int cpt = 32, tile = 64, nsplit = 1, kpart = 1;
  // a Gaussian table's points on k_dft_mfma (nsplit parts of kpart) and
  // its Gaussians on the VALU kernel (nsplitG parts of kpartG), all parts'
  // sums added by k_dft_combine
  bool mixed = false;
  int nsplitG = 0, kpartG = 1;
  bool serial = false;      // each segment's kernel waits for the one before
  std::vector<int> segBlk;  // row blocks of each segment
};

// The plan of one per-buffer call.  The kernel is bound by the VALU of
// each CU, and a buffer of nVisPIO rows makes few (256-row block, IF,
// channel chunk) items: 320 for a 10,000-row buffer of 4 IFs x 64 channels
// at 32 channels per thread, against 512 resident workgroups.  So:
//   * channels per thread: 32, or 16 where that wastes fewer padded
//     channels (a 16-channel chunk runs at about 0.9 of the rate);
//   * the component list is cut into nsplit parts of whole 64-component
//     tiles (4 tiles at least), so the buffer makes about 10 rounds of
//     resident workgroups, each placed by the dispatcher as a CU frees up;
//     their FP64 sums go to a scratch (at most 1 GiB) that k_dft_combine
//     adds in part order and applies;
//   * segments of whole rounds: one round first (its upload is the only
//     one nothing overlaps), two rounds each in the middle, and a short
//     last one (its download is the only one nothing overlaps).
// Measured on workload 2's 10,000-row buffers (4 IFs x 64 channels, 10,000
// components; 2 workgroups per CU resident): 16 parts and segments of
// 4, 8, 8, 8, 8, 4 row blocks ran at 5.7e12 evaluations/s, against 5.5e12
// with five equal segments, 4.9e12 with three, and 5.3-5.45e12 with 12,
// 23 or 32 parts.
//
// Where k_dft_mfma can run the buffer (mfma_ok) and fills at least 0.6 of
// its channel slots (whole groups of 64 channels, 4 a wave) it does, with
// its own items: (component part, row quad).  The component list is cut into parts
// of whole 128-component groups (the kernel's fold interval) so that the
// buffer makes about 10 rounds of its resident workgroups, and segments
// are sized as above with the middle ones a sixth of the buffer.
// Measured on the same buffers (100,000 rows): 9.9e12 evaluations/s with
// 3 or 4 parts (7 or 10 rounds), 8.5-8.7e12 with 7 or 12 parts, 8.8e12
// with 2; middle segments of a quarter, a fifth, an eighth or a tenth of
// the buffer 7.6-9.6e12.
BufPlan plan_buffer(State *st, const Desc &d, size_t rowBytes) {
  (void)rowBytes;
  BufPlan p;
  p.kpart = std::max(d.ncomp, 1);
  // k_dft_mfma forms 4 groups of 64 channels a wave: taken here where at
  // least 0.6 of those slots are selected channels
  const int ngrp = d.nIFsel * ((d.nchsel + 63) / 64);
  const double fill = (double)d.nIFsel * d.nchsel / (256.0 * ((ngrp + 3) / 4));
  const double chans = (double)d.nrow * d.nIFsel * d.nchsel;
  const double capParts = double(size_t(1) << 30) / (chans * sizeof(double2));  // parts the scratch holds
  const bool mixed = st->type == GPUModTypeGauss && st->npoint > 0 && st->npoint < d.ncomp &&
                     channels_ok(st, d) && fill >= 0.6 && capParts >= 2.0;
  if ((mfma_ok(st, d) && fill >= 0.6) || mixed) {
    p.cpt = 64;
    const long long slots = (long long)std::max(st->nCU, 1) * occupancy_of(st, (const void *)mfma_fn(st));
    const long long nquad = (d.nrow + 3) / 4;
    const int npt = mixed ? st->npoint : d.ncomp;
    if (mixed) {
      // the Gaussians: VALU items of whole 64-component tiles, about 10
      // rounds of that kernel's resident workgroups
      const int ng = d.ncomp - st->npoint;
      const long long slotsV =
          (long long)std::max(st->nCU, 1) * occupancy(st, pick_split(32, p.tile, true, false, st->prec));
      const long long perPart = (long long)d.nrowblk * d.nIFsel * ((d.nchsel + 31) / 32);
      int q = (int)std::min<long long>(16, std::max<long long>(1, (10 * slotsV + perPart - 1) / perPart));
      q = std::min(q, std::max(1, (int)capParts - 1));
      p.kpartG = ((ng + q - 1) / q + p.tile - 1) / p.tile * p.tile;
      p.nsplitG = (ng + p.kpartG - 1) / p.kpartG;
      p.mixed = true;
      p.kpart = std::max(npt, 1);
    }
    if (npt > 0) {
      const int want = (int)std::min<long long>(16, std::max<long long>(1, (10 * slots + nquad - 1) / nquad));
      const int kpt = ((npt + want - 1) / want + 127) / 128 * 128;
      const int ns = (npt + kpt - 1) / kpt;
      if (ns > 1 && ns + p.nsplitG <= capParts) {
        p.nsplit = ns;
        p.kpart = kpt;
      }
    }
    const long long perBlock = (long long)(kRows / 4) * p.nsplit;  // items of one row block
    const int round = (int)std::max<long long>(1, slots / perBlock);
    const int mid = std::max(2 * round, (d.nrowblk + 5) / 6);
    int b = 0;
    for (int i = 0; b < d.nrowblk; ++i) {
      int want = (i == 0) ? round : mid;
      const int left = d.nrowblk - b;
      if (left <= want) want = left;
      else if (left < want + round) want = left - round;
      p.segBlk.push_back(want);
      b += want;
    }
    return p;
  }
  const int n16 = (d.nchsel + 15) / 16, n32 = (d.nchsel + 31) / 32;
  p.cpt = (n16 * 16 / 0.9 < n32 * 32.0) ? 16 : 32;
  const int nchunk = p.cpt == 16 ? n16 : n32;
  const bool g = model_gauss_off(st->type) >= 0, s = model_is_spec(st->type);
  const long long slots = (long long)std::max(st->nCU, 1) * occupancy(st, pick_split(p.cpt, p.tile, g, s, st->prec));
  const long long perBlock = (long long)d.nIFsel * nchunk;  // items of one row block, one part
  const long long items = perBlock * d.nrowblk;
  if (d.ncomp > 0) {
    const double chans = (double)d.nrow * d.nIFsel * d.nchsel;
    int want = (int)std::min<long long>(16, std::max<long long>(1, (10 * slots + items - 1) / items));
    for (; want > 1; --want) {
      const int kpt = ((d.ncomp + want - 1) / want + p.tile - 1) / p.tile * p.tile;
      const int ns = (d.ncomp + kpt - 1) / kpt;
      if (kpt < 4 * p.tile || (double)ns * chans * sizeof(double2) > double(size_t(1) << 30)) continue;
      p.nsplit = ns;
      p.kpart = kpt;
      break;
    }
  }
  const int round = (int)std::max<long long>(1, slots / (perBlock * p.nsplit));  // row blocks per round
  const int mid = std::max(2 * round, (d.nrowblk + 15) / 16);                     // at most ~17 segments
  int b = 0;
  for (int i = 0; b < d.nrowblk; ++i) {
    int want = (i == 0) ? round : mid;
    const int left = d.nrowblk - b;
    // the last segment: what is left, at most one round; the one before
    // leaves it
    if (left <= want) want = left;
    else if (left < want + round) want = left - round;
    p.segBlk.push_back(want);
    b += want;
  }
  return p;
}

// ---- the call that allocates nothing (ObitHIPSkyModelDFTCalcNoAlloc) ----

// Device scratch of the no-allocation call: the per-IF frequency lines and
// the prepared component table live here instead of in library-allocated
// buffers.  One copy per GPU, held until the process ends; calls on one GPU
// share it.
constexpr size_t kNoAllocBytes = size_t(32) << 20;
constexpr int kNoAllocMaxIF = 64;
constexpr size_t kNoAllocLineBytes = 2 * kNoAllocMaxIF * sizeof(double);
constexpr size_t kNoAllocCompBytes = 3 * sizeof(double) + 4 * sizeof(float);  // x, y, z; amplitude, 3 Gaussian
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scr_mu;
void *g_scr_ptr[kMaxDevices] = {};

// The device address of this GPU's scratch (in host code the array's name
// is only a stand-in): looked up once per device, with that device current.
int scratch_ptr(int device, void **out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(kInvalid, "device %d outside 0..%d", device, kMaxDevices - 1);
  std::lock_guard<std::mutex> lk(g_scr_mu);
  if (!g_scr_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(kHipError, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scr_ptr[device] = p;
  }
  *out = g_scr_ptr[device];
  return kOK;
}

struct Lines {
  double fb[kNoAllocMaxIF], fdl[kNoAllocMaxIF];
};

__global__ void k_store_lines(Lines v, int n, double *fb, double *fdl) {
  const int i = threadIdx.x;
  if (i < n) { fb[i] = v.fb[i]; fdl[i] = v.fdl[i]; }
}

// frequency_lines for the no-allocation call: the same arithmetic from the
// host copies taken at SetFreq / SetMod, into fixed storage; refuses tables
// that are not a line within each IF (the FP64 kernel needs a per-channel
// table).
int lines_fixed(const State *st, const GPUVisInfo *vi, Lines &L) {
  const int nchsel = vi->chane - vi->chanb + 1, nIFsel = vi->IFe - vi->IFb + 1;
  if (nIFsel > kNoAllocMaxIF) return fail(kInvalid, "%d IFs selected; the no-allocation call takes at most %d",
                                          nIFsel, kNoAllocMaxIF);
  const int nfreq = vi->nchan * vi->nIF;
  const bool f64 = !st->freq64.empty() && st->uvRef > 0.0;
  if (f64 && (int)st->freq64.size() < nfreq)
    return fail(kInvalid, "frequency table has %zu values, need %d", st->freq64.size(), nfreq);
  if (!f64 && (st->fscaleHost.size() < (size_t)nfreq || st->fscalePtr != vi->freqScale))
    return fail(kState, "no host copy of this freqScale table: call ObitCUDASkyModelDFTSetMod first");
  auto val = [&](int IF, int c) {
    const int i = IF * vi->kincif + (vi->chanb + c) * vi->kincf;
    return f64 ? st->freq64[i] / st->uvRef : (double)st->fscaleHost[i];
  };
  for (int is = 0; is < nIFsel; ++is) {
    const int IF = vi->IFb + is;
    double a = val(IF, 0), b = 0.0;
    if (nchsel > 1) {
      if (f64) {
        b = (val(IF, nchsel - 1) - val(IF, 0)) / (nchsel - 1);
      } else {  // least squares through FP32 values, as frequency_lines
        double sx = 0, sy = 0, sxx = 0, sxy = 0;
        for (int c = 0; c < nchsel; ++c) {
          const double y = val(IF, c);
          sx += c; sy += y; sxx += (double)c * c; sxy += c * y;
        }
        const double n = nchsel;
        b = (n * sxy - sx * sy) / (n * sxx - sx * sx);
        a = (sy - b * sx) / n;
      }
      for (int c = 0; c < nchsel; ++c) {
        const double y = val(IF, c);
        const double dev = std::fabs(y - (a + b * c));
        const double tol = f64 ? 1e-12 * std::fabs(y) : 4e-7 * std::fabs(y);
        if (dev > tol)
          return fail(kInvalid, "channel frequencies of IF %d are not uniform: the no-allocation call runs the "
                                "split kernel only", IF);
      }
    }
    L.fb[is] = a;
    L.fdl[is] = b;
  }
  return kOK;
}

}  // namespace dft
}  // namespace obit

using obit::dft::State;
using obit::dft::find_state;

extern "C" {

void ObitCUDASkyModelDFTInit(GPUInfo *gpuInfo) {
  if (!gpuInfo) return;
  State *st = find_state(gpuInfo);
  if (!st) {
    st = new State();
    std::lock_guard<std::mutex> lk(obit::dft::g_mu);
    obit::dft::g_states[gpuInfo] = st;
  }
  st->device = gpuInfo->cuda_device < 0 ? 0 : gpuInfo->cuda_device;
  OBIT_HIP_CHECK(hipSetDevice(st->device));
  if (!st->stream) OBIT_HIP_CHECK(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  OBIT_HIP_CHECK(hipGetDeviceProperties(&prop, st->device));
  st->nCU = prop.multiProcessorCount;
  {
    // this GPU's scratch address for the no-allocation call
    void *p = nullptr;
    if (obit::dft::scratch_ptr(st->device, &p)) {
      std::fprintf(stderr, "obit_hip: DFTInit: %s\n", ObitHIPLastError());
      std::abort();
    }
  }
  st->fscaleHost.clear();
}

void ObitCUDASkyModelDFTSetMod(GPUInfo *gpuInfo) {
  State *st = find_state(gpuInfo);
  if (!st) { ObitCUDASkyModelDFTInit(gpuInfo); st = find_state(gpuInfo); }
  const GPUModelInfo *mi = gpuInfo->h_modelInfo;
  if (!mi) { std::fprintf(stderr, "obit_hip: DFTSetMod without h_modelInfo\n"); std::abort(); }
  OBIT_HIP_CHECK(hipSetDevice(st->device));
  const int n = mi->nmodel[0], size = mi->size;
  if (n < 0 || size < 4) { std::fprintf(stderr, "obit_hip: bad model shape %d x %d\n", n, size); std::abort(); }
  st->type = (int)mi->type;
  st->nterm = mi->nterm;
  st->size = size;
  st->ncomp = n;
  st->modelDev = mi->model[0];
  const int gOff = obit::dft::model_gauss_off(st->type);
  if (gOff >= 0 && size < 7) { std::fprintf(stderr, "obit_hip: Gaussian model needs 7 floats per row\n"); std::abort(); }
  if (n > st->capComp) {
    (void)hipFree(st->cx); (void)hipFree(st->cy); (void)hipFree(st->cz); (void)hipFree(st->camp); (void)hipFree(st->cg);
    const size_t m = (size_t)std::max(n, 1);
    OBIT_HIP_CHECK(hipMalloc(&st->cx, m * sizeof(double)));
    OBIT_HIP_CHECK(hipMalloc(&st->cy, m * sizeof(double)));
    OBIT_HIP_CHECK(hipMalloc(&st->cz, m * sizeof(double)));
    OBIT_HIP_CHECK(hipMalloc(&st->camp, m * sizeof(float)));
    OBIT_HIP_CHECK(hipMalloc(&st->cg, 3 * m * sizeof(float)));
    st->capComp = n;
  }
  if (n > 0) {
    hipLaunchKernelGGL(obit::dft::k_prep_comps, dim3((n + 255) / 256), dim3(256), 0, st->stream,
                       mi->model[0], size, n, gOff, st->cx, st->cy, st->cz, st->camp, st->cg);
    OBIT_HIP_CHECK(hipGetLastError());
  }
  // a Gaussian table: a copy with its components without Gaussian terms
  // first, and the Gaussian terms of the rest (the table itself keeps its
  // order for every other launch)
  st->npoint = 0;
  if (n > 0 && st->type == GPUModTypeGauss) {
    OBIT_HIP_CHECK(hipStreamSynchronize(st->stream));
    std::vector<double> x(n), y(n), z(n);
    std::vector<float> a(n), g(3 * (size_t)n);
    OBIT_HIP_CHECK(hipMemcpyAsync(x.data(), st->cx, n * sizeof(double), hipMemcpyDeviceToHost, st->stream));
    OBIT_HIP_CHECK(hipMemcpyAsync(y.data(), st->cy, n * sizeof(double), hipMemcpyDeviceToHost, st->stream));
    OBIT_HIP_CHECK(hipMemcpyAsync(z.data(), st->cz, n * sizeof(double), hipMemcpyDeviceToHost, st->stream));
    OBIT_HIP_CHECK(hipMemcpyAsync(a.data(), st->camp, n * sizeof(float), hipMemcpyDeviceToHost, st->stream));
    OBIT_HIP_CHECK(hipMemcpyAsync(g.data(), st->cg, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, st->stream));
    OBIT_HIP_CHECK(hipStreamSynchronize(st->stream));
    auto isPoint = [&](int k) { return g[k] == 0.0f && g[n + k] == 0.0f && g[2 * (size_t)n + k] == 0.0f; };
    std::vector<int> order;
    order.reserve(n);
    for (int k = 0; k < n; ++k)
      if (isPoint(k)) order.push_back(k);
    const int np = (int)order.size();
    for (int k = 0; k < n; ++k)
      if (!isPoint(k)) order.push_back(k);
    if (np > 0) {
      const int ng = n - np;
      // one buffer: mx, my, mz (doubles), mamp, cgG (floats)
      const size_t need = 3 * (size_t)n * sizeof(double) + (size_t)n * sizeof(float) + 3 * (size_t)ng * sizeof(float);
      if (need > st->mbufCap) {
        (void)hipFree(st->mbuf);
        st->mbuf = nullptr;
        st->mbufCap = 0;
        OBIT_HIP_CHECK(hipMalloc(&st->mbuf, need));
        st->mbufCap = need;
      }
      std::vector<unsigned char> h(need);
      double *hx = reinterpret_cast<double *>(h.data()), *hy = hx + n, *hz = hy + n;
      float *ha = reinterpret_cast<float *>(hz + n), *hg = ha + n;
      for (int i = 0; i < n; ++i) {
        const int k = order[i];
        hx[i] = x[k]; hy[i] = y[k]; hz[i] = z[k]; ha[i] = a[k];
        if (i >= np)
          for (int t = 0; t < 3; ++t) hg[t * (size_t)ng + (i - np)] = g[t * (size_t)n + k];
      }
      OBIT_HIP_CHECK(hipMemcpyAsync(st->mbuf, h.data(), need, hipMemcpyHostToDevice, st->stream));
      // the copy reads the host vector: done before it goes
      OBIT_HIP_CHECK(hipStreamSynchronize(st->stream));
      st->mx = reinterpret_cast<double *>(st->mbuf);
      st->my = st->mx + n;
      st->mz = st->my + n;
      st->mamp = reinterpret_cast<float *>(st->mz + n);
      st->cgG = st->mamp + n;
      st->npoint = np;
    }
  }
  // k_dft_mfma's scale: the largest amplitude to [2^13, 2^14)
  st->ampShift = 0;
  if (n > 0) {
    if (!st->dAmax) OBIT_HIP_CHECK(hipMalloc(&st->dAmax, sizeof(float)));
    hipLaunchKernelGGL(obit::dft::k_amp_max, dim3(1), dim3(256), 0, st->stream, st->camp, n, st->dAmax);
    OBIT_HIP_CHECK(hipGetLastError());
    float amax = 0.0f;
    OBIT_HIP_CHECK(hipMemcpyAsync(&amax, st->dAmax, sizeof(float), hipMemcpyDeviceToHost, st->stream));
    OBIT_HIP_CHECK(hipStreamSynchronize(st->stream));
    int e = 0;
    if (amax > 0.0f && std::isfinite(amax)) {
      (void)std::frexp(amax, &e);
      st->ampShift = 14 - e;
    }
  }
  OBIT_HIP_CHECK(hipStreamSynchronize(st->stream));
  st->modelReady = true;
  st->tabReady = false;
  // a new model may come with a new frequency table: keep the host copy of
  // the one the descriptor names now
  if (gpuInfo->h_visInfo && gpuInfo->h_visInfo->freqScale && st->freq64.empty() &&
      obit::dft::refresh_fscale(st, gpuInfo->h_visInfo, true)) {
    std::fprintf(stderr, "obit_hip: DFTSetMod: %s\n", ObitHIPLastError());
    std::abort();
  }
  obit::dft::destroy_graph(st);
}

void ObitCUDASkyModelDFTCalc(GPUInfo *gpuInfo) {
  using namespace obit::dft;
  State *st = find_state(gpuInfo);
  if (!st || !st->modelReady) {
    std::fprintf(stderr, "obit_hip: DFTCalc before DFTInit/DFTSetMod\n");
    std::abort();
  }
  OBIT_HIP_CHECK(hipSetDevice(st->device));
  const long long nvis = gpuInfo->nvis;
  const size_t rowBytes = (size_t)gpuInfo->lenvis * sizeof(float);
  const size_t bytes = (size_t)nvis * rowBytes;
  if (nvis <= 0) return;
  if (bytes > st->stageCap) {
    if (st->stage) OBIT_HIP_CHECK(hipFree(st->stage));
    OBIT_HIP_CHECK(hipMalloc(&st->stage, bytes));
    st->stageCap = bytes;
  }
  Desc d;
  int rc = make_desc(st, gpuInfo, st->stage, nvis, 32, d);
  if (rc) { std::fprintf(stderr, "obit_hip: DFTCalc: %s\n", ObitHIPLastError()); std::abort(); }
  if (!st->uniform) {
    // the FP64 kernel, in one piece on the library's stream
    OBIT_HIP_CHECK(hipMemcpyAsync(st->stage, gpuInfo->h_data, bytes, hipMemcpyHostToDevice, st->stream));
    rc = run_device(st, gpuInfo, st->stage, nvis, false);
    if (rc) { std::fprintf(stderr, "obit_hip: DFTCalc: %s\n", ObitHIPLastError()); std::abort(); }
    OBIT_HIP_CHECK(hipMemcpyAsync(gpuInfo->h_data, st->stage, bytes, hipMemcpyDeviceToHost, st->stream));
    OBIT_HIP_CHECK(hipStreamSynchronize(st->stream));
    return;
  }
  const BufPlan plan = plan_buffer(st, d, rowBytes);
  const int cpt = plan.cpt, nsplit = plan.nsplit, nseg = (int)plan.segBlk.size();
  const int ntot = nsplit + plan.nsplitG;  // parts in the scratch
  d.nchunk = (d.nchsel + cpt - 1) / cpt;
  if (ntot > 1) {
    const size_t need = (size_t)ntot * (size_t)nvis * d.nIFsel * d.nchsel * sizeof(double2);
    if (need > st->partCap) {
      if (st->part) OBIT_HIP_CHECK(hipFree(st->part));
      st->part = nullptr;
      st->partCap = 0;
      OBIT_HIP_CHECK(hipMalloc(&st->part, need));
      st->partCap = need;
    }
  }
  st->lastPath = 0;
  st->lastCpt = cpt;
  st->lastSplit = ntot;
  // The buffer in row segments on the caller's streams (ObitGPUSkyModel-
  // DFTInit creates them), round robin: each segment's upload, launch and
  // download queue on its stream, so kernels run while later segments are
  // uploaded and earlier ones downloaded (as the reference's
  // ObitCUDASkyModel.cu:804-963 pipelines its streams).  The uploads are
  // chained (each waits for the one before) so every segment's rows arrive
  // at the full link rate; with plan.serial each kernel also waits for the
  // one before, so segments finish, and download, in order.  At most 3
  // streams: a fourth would share a hardware queue with another (4 queues
  // per process, one taken by the library's own stream) and run behind it.
  int nstr = 0;
  hipStream_t segs[3];
  for (int i = 0; i < gpuInfo->nstream && i < STREAM_COUNT && nstr < 3; ++i)
    if (gpuInfo->stream && gpuInfo->stream[i]) segs[nstr++] = reinterpret_cast<hipStream_t>(gpuInfo->stream[i]);
  if (nstr == 0) segs[nstr++] = st->stream;
  while ((int)st->events.size() < 2 * nseg + 1) {
    hipEvent_t e;
    OBIT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    st->events.push_back(e);
  }
  hipEvent_t ready = st->events[0];
  hipEvent_t *up = st->events.data() + 1, *kd = st->events.data() + 1 + nseg;
  // the frequency lines (and any spectral table) were queued on the
  // library's stream: the segments start after them
  OBIT_HIP_CHECK(hipEventRecord(ready, st->stream));
  long long r0 = 0;
  for (int i = 0; i < nseg && r0 < nvis; ++i) {
    const long long n = std::min((long long)plan.segBlk[i] * kRows, nvis - r0);
    hipStream_t s = segs[i % nstr];
    OBIT_HIP_CHECK(hipStreamWaitEvent(s, i ? up[i - 1] : ready, 0));
    float *dv = st->stage + r0 * gpuInfo->lenvis;
    OBIT_HIP_CHECK(hipMemcpyAsync(dv, gpuInfo->h_data + r0 * gpuInfo->lenvis, (size_t)n * rowBytes,
                                  hipMemcpyHostToDevice, s));
    OBIT_HIP_CHECK(hipEventRecord(up[i], s));
    if (plan.serial && i) OBIT_HIP_CHECK(hipStreamWaitEvent(s, kd[i - 1], 0));
    Desc ds = d;
    ds.vis = dv;
    ds.nrow = n;
    ds.nrowblk = (int)((n + kRows - 1) / kRows);
    ds.nsplit = nsplit;
    ds.kpart = plan.kpart;
    ds.part = ntot > 1 ? st->part + (size_t)ntot * (size_t)r0 * d.nIFsel * d.nchsel : nullptr;
    ds.nitems = (long long)ds.nrowblk * ds.nIFsel * ds.nchunk * nsplit;
    if (plan.mixed) {
      // points into parts 0 .. nsplit-1, Gaussians into the next nsplitG
      Desc dp, dg;
      mixed_descs(st, ds, dp, dg);
      rc = launch_mfma(st, dp, s);
      dg.nsplit = plan.nsplitG;
      dg.kpart = plan.kpartG;
      dg.part = ds.part + (size_t)nsplit * (size_t)n * d.nIFsel * d.nchsel;
      dg.nitems = (long long)dg.nrowblk * dg.nIFsel * dg.nchunk * plan.nsplitG;
      if (!rc) rc = launch_split(st, dg, 32, plan.tile, 0, s);
      ds.nsplit = ntot;
    } else {
      rc = launch_split(st, ds, cpt, plan.tile, 0, s);
    }
    if (!rc && ntot > 1) {
      const long long m = n * ds.nIFsel * ds.nchsel;
      const int nb = (int)std::min<long long>((m + 255) / 256, (long long)st->nCU * 8);
      hipLaunchKernelGGL(k_dft_combine, dim3(nb), dim3(256), 0, s, ds);
      if (hipGetLastError() != hipSuccess) rc = obit::fail(obit::kHipError, "k_dft_combine launch");
    }
    if (rc) { std::fprintf(stderr, "obit_hip: DFTCalc: %s\n", ObitHIPLastError()); std::abort(); }
    OBIT_HIP_CHECK(hipEventRecord(kd[i], s));
    OBIT_HIP_CHECK(hipMemcpyAsync(gpuInfo->h_data + r0 * gpuInfo->lenvis, dv, (size_t)n * rowBytes,
                                  hipMemcpyDeviceToHost, s));
    r0 += n;
  }
  for (int i = 0; i < nstr; ++i) OBIT_HIP_CHECK(hipStreamSynchronize(segs[i]));
}

void ObitCUDASkyModelDFTShutdown(GPUInfo *gpuInfo) {
  State *st = nullptr;
  {
    std::lock_guard<std::mutex> lk(obit::dft::g_mu);
    auto it = obit::dft::g_states.find(gpuInfo);
    if (it != obit::dft::g_states.end()) { st = it->second; obit::dft::g_states.erase(it); }
  }
  if (!st) return;
  (void)hipSetDevice(st->device);
  (void)hipStreamSynchronize(st->stream);
  obit::dft::destroy_graph(st);
  (void)hipFree(st->cx); (void)hipFree(st->cy); (void)hipFree(st->cz); (void)hipFree(st->camp); (void)hipFree(st->cg);
  (void)hipFree(st->tab); (void)hipFree(st->dfreq); (void)hipFree(st->stage); (void)hipFree(st->part);
  (void)hipFree(st->dAmax);
  (void)hipFree(st->mbuf);
  for (hipEvent_t e : st->events) (void)hipEventDestroy(e);
  if (st->hfreq) (void)hipHostFree(st->hfreq);
  if (st->ownVis) (void)hipFree(st->dvis);
  (void)hipStreamDestroy(st->stream);
  delete st;
}

// ---- extensions ----

int ObitHIPSkyModelDFTSetFreq(GPUInfo *gpuInfo, const double *freqArr, int nfreq, double uvRefFreq) {
  State *st = find_state(gpuInfo);
  if (!st) return obit::fail(obit::kState, "DFTSetFreq before DFTInit");
  if (!freqArr || nfreq <= 0 || !(uvRefFreq > 0.0)) return obit::fail(obit::kInvalid, "bad frequency table");
  st->freq64.assign(freqArr, freqArr + nfreq);
  st->uvRef = uvRefFreq;
  st->tabReady = false;
  obit::dft::destroy_graph(st);
  return obit::kOK;
}

int ObitHIPSkyModelDFTGetBufferShape(GPUInfo *gpuInfo, int *cpt, int *nsplit) {
  State *st = find_state(gpuInfo);
  if (!st) return obit::fail(obit::kState, "DFTGetBufferShape before DFTInit");
  if (!cpt || !nsplit) return obit::fail(obit::kInvalid, "NULL output");
  *cpt = st->lastCpt;
  *nsplit = st->lastSplit;
  return obit::kOK;
}

int ObitHIPSkyModelDFTGetPath(GPUInfo *gpuInfo, int *path) {
  State *st = find_state(gpuInfo);
  if (!st) return obit::fail(obit::kState, "DFTGetPath before DFTInit");
  if (!path) return obit::fail(obit::kInvalid, "path is NULL");
  *path = st->lastPath;
  return obit::kOK;
}

int ObitHIPSkyModelDFTCalcNoAlloc(GPUInfo *gpuInfo, const float *d_model, int ncomp, float *d_vis,
                                  long long nvis) {
  using namespace obit;
  using namespace obit::dft;
  State *st = find_state(gpuInfo);
  if (!st) return fail(kState, "DFTCalcNoAlloc before DFTInit");
  if (!st->modelReady) return fail(kState, "DFTCalcNoAlloc: call ObitCUDASkyModelDFTSetMod first (model type)");
  if (ncomp < 0 || nvis < 0) return fail(kInvalid, "negative component or row count");
  if ((ncomp > 0 && !d_model) || (nvis > 0 && !d_vis)) return fail(kInvalid, "null device pointer");
  if (model_is_spec(st->type))
    return fail(kInvalid, "spectral model types need the library's spectral table: use CalcResident");
  const size_t need = kNoAllocLineBytes + (size_t)ncomp * kNoAllocCompBytes + 4 * 256;
  if (need > kNoAllocBytes)
    return fail(kInvalid, "%d components need %zu bytes of scratch; the no-allocation call holds %zu (%zu components)",
                ncomp, need, kNoAllocBytes, (kNoAllocBytes - kNoAllocLineBytes - 4 * 256) / kNoAllocCompBytes);
  const int gOff = model_gauss_off(st->type);
  if (gOff >= 0 && st->size < 7) return fail(kInvalid, "Gaussian model needs 7 floats per row");
  if (nvis == 0) return kOK;
  Desc d;
  int rc = desc_layout(gpuInfo, d_vis, nvis, st->cpt, d);
  if (rc) return rc;
  Lines L;
  rc = lines_fixed(st, gpuInfo->h_visInfo, L);
  if (rc) return rc;
  OBIT_HIP_TRY(hipSetDevice(st->device));
  void *base = nullptr;
  rc = scratch_ptr(st->device, &base);
  if (rc) return rc;
  // scratch layout: lines, then x, y, z (FP64), amplitudes, Gaussian terms
  auto align = [](size_t o) { return (o + 255) & ~size_t(255); };
  unsigned char *p = static_cast<unsigned char *>(base);
  size_t o = 0;
  double *fb = reinterpret_cast<double *>(p + o);
  double *fdl = fb + kNoAllocMaxIF;
  o = align(o + kNoAllocLineBytes);
  double *cx = reinterpret_cast<double *>(p + o); o = align(o + sizeof(double) * ncomp);
  double *cy = reinterpret_cast<double *>(p + o); o = align(o + sizeof(double) * ncomp);
  double *cz = reinterpret_cast<double *>(p + o); o = align(o + sizeof(double) * ncomp);
  float *camp = reinterpret_cast<float *>(p + o); o = align(o + sizeof(float) * ncomp);
  float *cg = reinterpret_cast<float *>(p + o); o += 3 * sizeof(float) * ncomp;
  if (o > kNoAllocBytes) return fail(kInvalid, "scratch layout overflow (%zu bytes)", o);
  hipLaunchKernelGGL(k_store_lines, dim3(1), dim3(kNoAllocMaxIF), 0, st->stream, L, d.nIFsel, fb, fdl);
  OBIT_HIP_TRY(hipGetLastError());
  if (ncomp > 0) {
    hipLaunchKernelGGL(k_prep_comps, dim3((ncomp + 255) / 256), dim3(256), 0, st->stream, d_model, st->size,
                       ncomp, gOff, cx, cy, cz, camp, cg);
    OBIT_HIP_TRY(hipGetLastError());
  }
  d.fbase = fb;
  d.fdel = fdl;
  d.ncomp = ncomp;
  d.cx = cx; d.cy = cy; d.cz = cz;
  d.camp = camp; d.cg = cg;
  st->lastPath = 0;
  // the VALU kernel (the matrix-core kernel's eligibility reads the host
  // copy of the lines, which this call does not make)
  return launch_split(st, d, st->cpt == 64 ? 32 : st->cpt, st->tile, st->wgPerCU);
}

int ObitHIPSkyModelDFTSetSpecRef(GPUInfo *gpuInfo, const double *refFreq, int n) {
  State *st = find_state(gpuInfo);
  if (!st) return obit::fail(obit::kState, "DFTSetSpecRef before DFTInit");
  if (!refFreq || n <= 0) return obit::fail(obit::kInvalid, "bad reference frequencies");
  st->specRef.assign(refFreq, refFreq + n);
  st->tabReady = false;
  return obit::kOK;
}

int ObitHIPSkyModelDFTSetVis(GPUInfo *gpuInfo, const float *h_vis, long long nvis) {
  State *st = find_state(gpuInfo);
  if (!st) return obit::fail(obit::kState, "DFTSetVis before DFTInit");
  if (nvis < 0 || (nvis > 0 && !h_vis)) return obit::fail(obit::kInvalid, "bad visibility buffer");
  OBIT_HIP_TRY(hipSetDevice(st->device));
  const size_t bytes = (size_t)nvis * gpuInfo->lenvis * sizeof(float);
  if (st->ownVis && st->dvis) OBIT_HIP_TRY(hipFree(st->dvis));
  st->dvis = nullptr;
  st->ownVis = true;
  hipError_t e = hipMalloc(&st->dvis, bytes ? bytes : 4);
  if (e != hipSuccess) return obit::fail(obit::kNoMemory, "resident vis (%zu B): %s", bytes, hipGetErrorString(e));
  OBIT_HIP_TRY(hipMemcpy(st->dvis, h_vis, bytes, hipMemcpyHostToDevice));
  st->nres = nvis;
  obit::dft::destroy_graph(st);
  return obit::kOK;
}

int ObitHIPSkyModelDFTAttachVis(GPUInfo *gpuInfo, float *d_vis, long long nvis) {
  State *st = find_state(gpuInfo);
  if (!st) return obit::fail(obit::kState, "DFTAttachVis before DFTInit");
  if (nvis < 0 || (nvis > 0 && !d_vis)) return obit::fail(obit::kInvalid, "bad device buffer");
  if (st->ownVis && st->dvis) OBIT_HIP_TRY(hipFree(st->dvis));
  st->dvis = d_vis;
  st->ownVis = false;
  st->nres = nvis;
  obit::dft::destroy_graph(st);
  return obit::kOK;
}

int ObitHIPSkyModelDFTGetVis(GPUInfo *gpuInfo, float *h_vis, long long nvis) {
  State *st = find_state(gpuInfo);
  if (!st || !st->dvis) return obit::fail(obit::kState, "no resident visibilities");
  if (nvis > st->nres) return obit::fail(obit::kInvalid, "asked for %lld rows, %lld resident", nvis, st->nres);
  OBIT_HIP_TRY(hipSetDevice(st->device));
  OBIT_HIP_TRY(hipStreamSynchronize(st->stream));
  OBIT_HIP_TRY(hipMemcpy(h_vis, st->dvis, (size_t)nvis * gpuInfo->lenvis * sizeof(float), hipMemcpyDeviceToHost));
  return obit::kOK;
}

int ObitHIPSkyModelDFTCalcResident(GPUInfo *gpuInfo) {
  State *st = find_state(gpuInfo);
  if (!st || !st->dvis) return obit::fail(obit::kState, "no resident visibilities");
  OBIT_HIP_TRY(hipSetDevice(st->device));
  return obit::dft::run_device(st, gpuInfo, st->dvis, st->nres, false);
}

int ObitHIPSkyModelDFTCalcFP64(GPUInfo *gpuInfo) {
  State *st = find_state(gpuInfo);
  if (!st || !st->dvis) return obit::fail(obit::kState, "no resident visibilities");
  OBIT_HIP_TRY(hipSetDevice(st->device));
  return obit::dft::run_device(st, gpuInfo, st->dvis, st->nres, true);
}

int ObitHIPSkyModelDFTGraphLaunch(GPUInfo *gpuInfo) {
  State *st = find_state(gpuInfo);
  if (!st || !st->dvis) return obit::fail(obit::kState, "no resident visibilities");
  OBIT_HIP_TRY(hipSetDevice(st->device));
  const long long key = (long long)(size_t)st->dvis ^ (st->nres << 20) ^ ((long long)st->cpt << 8) ^
                        ((long long)st->tile << 12) ^ ((long long)st->wgPerCU << 4) ^ st->prec;
  if (!st->gexec || st->graphKey != key) {
    obit::dft::destroy_graph(st);
    obit::dft::Desc d;
    int rc = obit::dft::make_desc(st, gpuInfo, st->dvis, st->nres, st->cpt, d);
    if (rc) return rc;
    OBIT_HIP_TRY(hipStreamSynchronize(st->stream));
    OBIT_HIP_TRY(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal));
    rc = st->uniform ? obit::dft::launch_split(st, d, st->cpt, st->tile, st->wgPerCU)
                     : obit::dft::launch_fp64(st, d);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return obit::fail(obit::kHipError, "graph capture: %s", hipGetErrorString(e));
    st->graph = g;
    OBIT_HIP_TRY(hipGraphInstantiate(&st->gexec, st->graph, nullptr, nullptr, 0));
    st->graphKey = key;
  }
  OBIT_HIP_TRY(hipGraphLaunch(st->gexec, st->stream));
  return obit::kOK;
}

int ObitHIPSkyModelDFTSetVariant(GPUInfo *gpuInfo, int cpt, int tile, int wgPerCU) {
  State *st = find_state(gpuInfo);
  if (!st) return obit::fail(obit::kState, "DFTSetVariant before DFTInit");
  if (cpt != 8 && cpt != 16 && cpt != 32 && cpt != 64)
    return obit::fail(obit::kInvalid, "cpt must be 8, 16, 32 or 64 (the matrix-core kernel)");
  if (tile != 32 && tile != 64 && tile != 128) return obit::fail(obit::kInvalid, "tile must be 32, 64 or 128");
  if (wgPerCU < 1 || wgPerCU > 8) return obit::fail(obit::kInvalid, "wgPerCU must be 1..8");
  st->cpt = cpt; st->tile = tile; st->wgPerCU = wgPerCU;
  return obit::kOK;
}

int ObitHIPSkyModelDFTGetVariant(GPUInfo *gpuInfo, int *cpt, int *tile, int *wgPerCU) {
  State *st = find_state(gpuInfo);
  if (!st) return obit::fail(obit::kState, "DFTGetVariant before DFTInit");
  if (cpt) *cpt = st->cpt;
  if (tile) *tile = st->tile;
  if (wgPerCU) *wgPerCU = st->wgPerCU;
  return obit::kOK;
}

// Precision policy of the per-channel step: 0 hardware sin/cos for both
// phasors, 1 hardware for the centre phasor and the polynomial for the
// step (default), 2 polynomial for both.
int ObitHIPSkyModelDFTSetPrecision(GPUInfo *gpuInfo, int prec) {
  State *st = find_state(gpuInfo);
  if (!st) return obit::fail(obit::kState, "DFTSetPrecision before DFTInit");
  if (prec < 0 || prec > 2) return obit::fail(obit::kInvalid, "precision policy must be 0..2");
  st->prec = prec;
  return obit::kOK;
}

int ObitHIPSkyModelDFTAutotune(GPUInfo *gpuInfo, long long sampleRows, int reps) {
  State *st = find_state(gpuInfo);
  if (!st || !st->dvis) return obit::fail(obit::kState, "autotune needs resident visibilities");
  OBIT_HIP_TRY(hipSetDevice(st->device));
  if (reps < 1) reps = 1;
  // The sample must give the widest grid (4 workgroups per CU) at least 8
  // rounds of items with the coarsest chunking (32 channels per thread);
  // with fewer items than workgroups every grid width runs the same launch
  // and the choice of width is noise.
  long long n = std::max(sampleRows, 1LL);
  {
    obit::dft::Desc d1;
    int rc1 = obit::dft::make_desc(st, gpuInfo, st->dvis, obit::dft::kRows, 32, d1);
    if (rc1) return rc1;
    const long long perBlk = std::max(d1.nitems, 1LL);
    const long long blocks = (8LL * st->nCU * 4 + perBlk - 1) / perBlk;
    n = std::max(n, blocks * obit::dft::kRows);
  }
  n = std::min(n, st->nres);
  const size_t bytes = (size_t)n * gpuInfo->lenvis * sizeof(float);
  float *tmp = nullptr;
  OBIT_HIP_TRY(hipMalloc(&tmp, bytes));
  OBIT_HIP_TRY(hipMemcpy(tmp, st->dvis, bytes, hipMemcpyDeviceToDevice));
  hipEvent_t e0, e1;
  OBIT_HIP_TRY(hipEventCreate(&e0));
  OBIT_HIP_TRY(hipEventCreate(&e1));
  const int cpts[2] = {16, 32}, tiles[3] = {32, 64, 128}, wgs[3] = {1, 2, 4};
  float best = 1e30f;
  int bc = st->cpt, bt = st->tile, bw = st->wgPerCU;
  int rc = obit::kOK;
  for (int ci = 0; ci < 2 && !rc; ++ci) {
    obit::dft::Desc d, dw;
    rc = obit::dft::make_desc(st, gpuInfo, tmp, n, cpts[ci], d);
    if (!rc) rc = obit::dft::make_desc(st, gpuInfo, tmp, std::min(n, (long long)obit::dft::kRows), cpts[ci], dw);
    if (rc) break;
    for (int ti = 0; ti < 3 && !rc; ++ti) {
      for (int wi = 0; wi < 3 && !rc; ++wi) {
        rc = obit::dft::launch_split(st, dw, cpts[ci], tiles[ti], wgs[wi]);  // load the code object
        if (rc) break;
        (void)hipEventRecord(e0, st->stream);
        for (int r = 0; r < reps && !rc; ++r) rc = obit::dft::launch_split(st, d, cpts[ci], tiles[ti], wgs[wi]);
        (void)hipEventRecord(e1, st->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) { best = ms; bc = cpts[ci]; bt = tiles[ti]; bw = wgs[wi]; }
      }
    }
  }
  if (!rc) {
    // the matrix-core kernel, where the model and channels allow it
    obit::dft::Desc d, dw;
    rc = obit::dft::make_desc(st, gpuInfo, tmp, n, 64, d);
    if (!rc) rc = obit::dft::make_desc(st, gpuInfo, tmp, std::min(n, (long long)obit::dft::kRows), 64, dw);
    if (!rc && (obit::dft::mfma_ok(st, d) || obit::dft::mixed_ok(st, d))) {
      rc = obit::dft::launch_split(st, dw, 64, bt, bw);
      if (!rc) {
        (void)hipEventRecord(e0, st->stream);
        for (int r = 0; r < reps && !rc; ++r) rc = obit::dft::launch_split(st, d, 64, bt, bw);
        (void)hipEventRecord(e1, st->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) { best = ms; bc = 64; }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(tmp);
  if (rc) return rc;
  st->cpt = bc; st->tile = bt; st->wgPerCU = bw;
  obit::dft::destroy_graph(st);
  return obit::kOK;
}

void *ObitHIPSkyModelDFTStream(GPUInfo *gpuInfo) {
  State *st = find_state(gpuInfo);
  return st ? (void *)st->stream : nullptr;
}

}  // extern "C"
