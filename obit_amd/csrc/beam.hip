// beam.hip -- Jones-corrected (antenna-beam-corrected) DFT sky model for
// linear feeds (ObitHIPBeam.h).
//
// Replaces the evaluation of reference src/ObitCUDABeamModel.cu
// (beamPointKernel :139-310, beamGaussKernel :687-858), whose inner loop
// forms, per (visibility, channel, component), the four products of the
// baseline's two Jones matrices with the Stokes coherency
// (include/CUDALinCorSum.h).  Those products depend on the component, the
// two antenna types, the beam channel and the parallactic angle, not on
// the visibility, so they are formed once per call:
//   k_beam_coh  C[pair][comp][jchan] = J1 P J2^H   (pair = type1 * ntype + type2)
// and the DFT is then a sum of complex multiply-adds of the component's
// phasor into each correlation:
//   k_beam_dft  S_e(row, chan) = sum_k a_k exp(2 pi i psi_k nu/nu0) C_e
// with the phasor walk of the split DFT (FP64 phase and range reduction,
// FP32 sincos at the centre of the thread's channel chunk, the step as three
// shears) and FP32 tile sums folded into FP64.
//
// Work decomposition: a workgroup is 256 consecutive rows x one (IF,
// chunk of 8 channels with 2 correlations, 4 with 4); each thread owns one
// row.  When the chunk's channels
// share one beam channel (the usual case: beam channels are coarse), the
// tile's products for every antenna-type pair are staged in LDS and each
// thread reads its pair's once per component; otherwise each channel
// reads its product from the table in global memory.
#include "common.h"
#include "ObitHIPBeam.h"
#include "ObitHIPBeamModel.h"
#include "ObitHIPVis.h"
#include "phasor.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace obit {
namespace beam {

constexpr int kRows = 256;     // rows (threads) per workgroup
// Channels per thread.  16 with 2 correlations needs 310 VGPRs (one wave
// per SIMD) and ran 1.54x slower than 8 (two waves per SIMD); with 4
// correlations 8 needs 314 VGPRs and ran 1.44x slower than 4 (173 VGPRs).
constexpr int kCPT2 = 8;       // 2 correlations
constexpr int kCPT4 = 4;       // 4 correlations
constexpr int kTile = 64;      // components per LDS tile
constexpr int kMaxPairs = 16;  // type pairs staged in LDS (4 antenna types)
constexpr int kMaxTypes = 16;

struct Desc {
  float *vis;
  long long nrow;
  int lrec, nrparm, ilocu, ilocv, ilocw, ilocb, iloca1, iloca2;
  int nif, nchan, incf, incif;
  int nchunk, nrowblk;
  long long nitems;
  int ncomp, ntype, nant, njchan, npair;
  const double *fIF, *dfIF;
  const double *cx, *cy, *cz;  // turns per wavelength at nu0
  const float *camp;
  const float *cg;             // Gaussian coefficients [3][ncomp]
  const int *antType, *jmap;
  const float2 *coh;           // [pair][comp][jchan][4]: XX, XY, YX, YY
  const float *tab;            // per-channel amplitudes [comp][nif * nchan], or null
  int op;
  float sf[4];                 // XX, XY, YX, YY
  int chanb, chane, IFb, IFe;  // the selection the model is applied to
};

// One product per (pair, component, beam channel): beam1 * P * beam2^H,
// in the CPU's order (model * beam2^H, then beam1 * that;
// src/ObitSkyModelVMBeamMF.c:2752-2755).
__global__ __launch_bounds__(256) void k_beam_coh(const float *__restrict__ jones,
                                                  float2 *__restrict__ coh, int ntype, int ncomp,
                                                  int njchan, int stokType, float c, float s) {
  const long long n = (long long)ntype * ntype * ncomp * njchan;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int jc = (int)(i % njchan);
  const int k = (int)((i / njchan) % ncomp);
  const int p = (int)(i / ((long long)njchan * ncomp));
  const int t1 = p / ntype, t2 = p - t1 * ntype;
  const float *A = jones + (((long long)t1 * ncomp + k) * njchan + jc) * 8;
  const float *B = jones + (((long long)t2 * ncomp + k) * njchan + jc) * 8;
  // coherency P of the model's Stokes parameter, linear feeds
  float Pr[4] = {1.0f, 0.0f, 0.0f, 1.0f}, Pi[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (stokType == 2) {
    Pr[0] = c; Pr[1] = -s; Pr[2] = -s; Pr[3] = -c;
  } else if (stokType == 3) {
    Pr[0] = s; Pr[1] = c; Pr[2] = c; Pr[3] = -s;
  } else if (stokType == 4) {
    Pr[0] = 0.0f; Pr[1] = 0.0f; Pr[2] = 0.0f; Pr[3] = 0.0f;
    Pi[1] = 1.0f; Pi[2] = -1.0f;
  }
  // T = P * B^H: T[r][q] = sum_m P[r][m] conj(B[q][m])
  float Tr[4], Ti[4];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      float re = 0.0f, im = 0.0f;
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const float br = B[2 * (2 * q + m)], bi = -B[2 * (2 * q + m) + 1];
        const float pr = Pr[2 * r + m], pi = Pi[2 * r + m];
        re += pr * br - pi * bi;
        im += pr * bi + pi * br;
      }
      Tr[2 * r + q] = re;
      Ti[2 * r + q] = im;
    }
  }
  // M = A * T
  float2 *out = coh + i * 4;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      float re = 0.0f, im = 0.0f;
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const float ar = A[2 * (2 * r + m)], ai = A[2 * (2 * r + m) + 1];
        re += ar * Tr[2 * m + q] - ai * Ti[2 * m + q];
        im += ar * Ti[2 * m + q] + ai * Tr[2 * m + q];
      }
      out[2 * r + q] = make_float2(re, im);
    }
  }
}

// Antenna numbers (0-rel) from BASELINE = 256 a1 + a2 (+ subarray
// fraction) or from ANTENNA1/ANTENNA2 (the reference's GetAnts,
// ObitCUDABeamModel.cu:82-98).
__device__ __forceinline__ void row_ants(const Desc &d, const float *rp, int *a1, int *a2) {
  if (d.ilocb >= 0) {
    const int cb = (int)rp[d.ilocb];
    const int x = (int)((double)cb / 256.0 + 0.001);
    *a1 = x - 1;
    *a2 = (int)((double)(cb - x * 256) + 0.001) - 1;
  } else {
    *a1 = (int)(rp[d.iloca1] + 0.5f) - 1;
    *a2 = (int)(rp[d.iloca2] + 0.5f) - 1;
  }
}

// Accumulate v = (vr, vi) times product m into (ar, ai).
__device__ __forceinline__ void cmac(float vr, float vi, float2 m, float &ar, float &ai) {
  ar = fmaf(vr, m.x, fmaf(-vi, m.y, ar));
  ai = fmaf(vr, m.y, fmaf(vi, m.x, ai));
}

// NC correlations: 2 (XX, YY: products 0 and 3) or 4 (XX, XY, YX, YY).
template <int NC>
__device__ __forceinline__ int prod_index(int e) {
  return NC == 2 ? 3 * e : e;
}

// Three workgroups per CU (168 VGPRs, three waves per SIMD): a few
// registers spill, and it still runs 1.16x (2 correlations) and 1.11x (4)
// faster than two workgroups with no spills.
constexpr int kBlocksPerCU = 3;

template <int NC, int kCPT, bool GAUSS, bool SPEC>
__global__ __launch_bounds__(kRows, kBlocksPerCU) void k_beam_dft(Desc d) {
  __shared__ double s_x[kTile], s_y[kTile], s_z[kTile];
  __shared__ float s_a[kTile];
  __shared__ float s_g[GAUSS ? 3 * kTile : 1];
  __shared__ float s_t[SPEC ? kTile * kCPT : 1];
  __shared__ float2 s_m[kTile * kMaxPairs * NC];
  const int tid = threadIdx.x;

  for (long long item = blockIdx.x; item < d.nitems; item += gridDim.x) {
    const int rb = (int)(item % d.nrowblk);
    const int ic = (int)(item / d.nrowblk);
    const int iIF = ic / d.nchunk;
    const int c0 = (ic - iIF * d.nchunk) * kCPT;
    // items wholly outside the selection (the same for the whole block)
    if (iIF < d.IFb || iIF > d.IFe || c0 + kCPT - 1 < d.chanb || c0 > d.chane) continue;
    const long long row = (long long)rb * kRows + tid;
    const bool valid = row < d.nrow;
    const float *rp = d.vis + (valid ? row : 0LL) * (long long)d.lrec;
    int a1 = 0, a2 = 0;
    if (valid) row_ants(d, rp, &a1, &a2);
    const bool live = valid && a1 >= 0 && a1 < d.nant && a2 >= 0 && a2 < d.nant;
    const int pair = live ? d.antType[a1] * d.ntype + d.antType[a2] : 0;
    const float uf = live ? rp[d.ilocu] : 0.0f;
    const float vf = live ? rp[d.ilocv] : 0.0f;
    const float wf = live ? rp[d.ilocw] : 0.0f;
    const double u = uf, v = vf, w = wf;
    // beam channel of each channel of the chunk (the same for every lane)
    const int *jm = d.jmap + (long long)iIF * d.nchan;
    int jcs[kCPT];
    bool uniform = true;
#pragma unroll
    for (int j = 0; j < kCPT; ++j) {
      jcs[j] = __builtin_amdgcn_readfirstlane(jm[min(c0 + j, d.nchan - 1)]);
      uniform = uniform && (jcs[j] == jcs[0]);
    }
    const bool useLds = uniform && d.npair <= kMaxPairs;
    const double fd = d.dfIF[iIF];
    const double fm = d.fIF[iIF] + (double)(c0 + kCPT / 2) * fd;
    const float fmf = (float)fm, fdf = (float)fd;
    const float uu = uf * uf, vv = vf * vf, uv = uf * vf;

    float accr[NC][kCPT], acci[NC][kCPT];
    double outr[NC][kCPT], outi[NC][kCPT];
#pragma unroll
    for (int e = 0; e < NC; ++e)
#pragma unroll
      for (int j = 0; j < kCPT; ++j) {
        accr[e][j] = acci[e][j] = 0.0f;
        outr[e][j] = outi[e][j] = 0.0;
      }

    for (int k0 = 0; k0 < d.ncomp; k0 += kTile) {
      __syncthreads();
      for (int t = tid; t < kTile; t += kRows) {
        const int k = k0 + t;
        const bool ok = k < d.ncomp;
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
        const long long ld = (long long)d.nif * d.nchan;
        for (int t = tid; t < kTile * kCPT; t += kRows) {
          const int kk = t / kCPT, cc = t - kk * kCPT;
          const int k = k0 + kk, ch = c0 + cc;
          s_t[t] = (k < d.ncomp && ch < d.nchan) ? d.tab[(long long)k * ld + (long long)iIF * d.nchan + ch] : 0.0f;
        }
      }
      if (useLds) {
        const int per = d.npair * NC;
        for (int t = tid; t < kTile * per; t += kRows) {
          const int kk = t / per, r = t - kk * per;
          const int p = r / NC, e = r - p * NC;
          const int k = k0 + kk;
          s_m[t] = (k < d.ncomp)
                       ? d.coh[(((long long)p * d.ncomp + k) * d.njchan + jcs[0]) * 4 + prod_index<NC>(e)]
                       : make_float2(0.0f, 0.0f);
        }
      }
      __syncthreads();
      const int kn = min(kTile, d.ncomp - k0);
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
        sincos_turns_poly((float)pm, &sm, &cm);
        sincos_halfturns_small((float)pd, &sh, &ch);
        const float tau = sh * __builtin_amdgcn_rcpf(ch);
        const float sig = 2.0f * sh * ch;
        const float a0 = s_a[kk];
        const float sgn = flip ? -1.0f : 1.0f;
        const float ao = flip ? -a0 : a0;
        float gf = a0, rf = 0.0f, rbk = 0.0f, qg = 0.0f;
        if (GAUSS) {
          const float G = fmaf(s_g[3 * kk + 2], uv, fmaf(s_g[3 * kk + 1], vv, s_g[3 * kk] * uu));
          const float e0 = G * fmf * fmf;
          const float Gd = G * fdf;
          const bool on = e0 > -60.0f;
          gf = on ? (SPEC ? 1.0f : a0) * __expf(e0) : 0.0f;
          rf = on ? sgn * __expf(Gd * (2.0f * fmf + fdf)) : 0.0f;
          rbk = on ? sgn * __expf(Gd * (fdf - 2.0f * fmf)) : 0.0f;
          qg = on ? __expf(2.0f * Gd * fdf) : 0.0f;
        }
        // this component's products: from LDS once, or per channel
        float mur[NC], mui[NC];
        const float2 *mg = d.coh + ((long long)pair * d.ncomp + k0 + kk) * d.njchan * 4;
#pragma unroll
        for (int e = 0; e < NC; ++e) {
          const float2 m = useLds ? s_m[(kk * d.npair + pair) * NC + e] : make_float2(0.0f, 0.0f);
          mur[e] = m.x;
          mui[e] = m.y;
        }
        // forward from the centre channel
        float x = cm, y = sm, g = gf, r = rf;
#pragma unroll
        for (int j = kCPT / 2; j < kCPT; ++j) {
          const bool odd = (j - kCPT / 2) & 1;
          {
            const float amp = GAUSS ? (SPEC ? s_t[kk * kCPT + j] * g : g)
                                    : (SPEC ? (odd ? s_t[kk * kCPT + j] * sgn : s_t[kk * kCPT + j])
                                            : (odd ? ao : a0));
            const float vr = amp * x, vi = amp * y;
#pragma unroll
            for (int e = 0; e < NC; ++e) {
              const float2 m = useLds ? make_float2(mur[e], mui[e]) : mg[jcs[j] * 4 + prod_index<NC>(e)];
              cmac(vr, vi, m, accr[e][j], acci[e][j]);
            }
          }
          if (j + 1 < kCPT) {
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
        for (int j = kCPT / 2 - 1; j >= 0; --j) {
          const bool odd = (kCPT / 2 - j) & 1;
          const float x1 = fmaf(tau, y, x);
          y = fmaf(-sig, x1, y);
          x = fmaf(tau, y, x1);
          {
            const float amp = GAUSS ? (SPEC ? s_t[kk * kCPT + j] * gb : gb)
                                    : (SPEC ? (odd ? s_t[kk * kCPT + j] * sgn : s_t[kk * kCPT + j])
                                            : (odd ? ao : a0));
            const float vr = amp * x, vi = amp * y;
#pragma unroll
            for (int e = 0; e < NC; ++e) {
              const float2 m = useLds ? make_float2(mur[e], mui[e]) : mg[jcs[j] * 4 + prod_index<NC>(e)];
              cmac(vr, vi, m, accr[e][j], acci[e][j]);
            }
          }
          if (GAUSS && j > 0) { gb *= rb2; rb2 *= qg; }
        }
      }
#pragma unroll
      for (int e = 0; e < NC; ++e)
#pragma unroll
        for (int j = 0; j < kCPT; ++j) {
          outr[e][j] += (double)accr[e][j];
          outi[e][j] += (double)acci[e][j];
          accr[e][j] = acci[e][j] = 0.0f;
        }
    }

    if (!live) continue;
    // apply, following src/ObitSkyModelVMBeamMF.c:3711-3800
    float *base = d.vis + row * (long long)d.lrec + d.nrparm + (long long)iIF * d.incif;
#pragma unroll
    for (int j = 0; j < kCPT; ++j) {
      if (c0 + j >= d.nchan) break;
      if (c0 + j < d.chanb || c0 + j > d.chane) continue;
      float *p = base + (long long)(c0 + j) * d.incf;
      const bool repl = d.op == ObitHIPBeamOpRepl;
      bool allBad = (p[2] <= 0.0f) && (p[5] <= 0.0f);
      if (NC == 4) allBad = allBad && (p[8] <= 0.0f) && (p[11] <= 0.0f);
      if (allBad && !repl) continue;
      // XX and YY (products 0 and 3)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int e = (NC == 2) ? h : 3 * h;
        const int prod = 3 * h;
        float mr = (float)outr[e][j] * d.sf[prod], mi = (float)outi[e][j] * d.sf[prod];
        float *q = p + 3 * h;
        if (!(q[2] > 0.0f || repl)) continue;
        if (d.op == ObitHIPBeamOpDiv) {
          float wm = mr * mr + mi * mi;
          mr = mr / wm;
          mi = mi / wm;
          wm = sqrtf(wm);
          const float re = q[0], im = q[1];
          q[0] = mr * re + mi * im;
          q[1] = mr * im - mi * re;
          q[2] = q[2] * wm;
        } else if (repl) {
          q[0] = mr;
          q[1] = mi;
        } else {
          q[0] -= mr;
          q[1] -= mi;
        }
      }
      // XY (product 1) at +6 and YX (product 2) at +9: not when dividing
      if (NC == 4 && d.op != ObitHIPBeamOpDiv) {
#pragma unroll
        for (int h = 1; h <= 2; ++h) {
          const float mr = (float)outr[h][j] * d.sf[h], mi = (float)outi[h][j] * d.sf[h];
          float *q = p + 3 + 3 * h;
          if (!(q[2] > 0.0f || repl)) continue;
          if (repl) {
            q[0] = mr;
            q[1] = mi;
          } else {
            q[0] -= mr;
            q[1] -= mi;
          }
        }
      }
    }
  }
}

// ---- Jones matrices from beam images (the reference's GetJonesKernel,
// ObitCUDABeamModel.cu:1227-1308, with CUDABeamInterpPixel and
// SetConvKernal, include/CUDABeamInterp.h:81-145, 233-258, and
// CUDAPositionXYpix(LM), include/CUDASkyGeom.h:90-312)

struct BeamImg {
  int nx, ny, proj, hwidth;
  double crval[2];
  float crpix[2], cdelt[2];
  float denom[10];      // reciprocals of the Lagrange denominators
  long long off;        // first float of this type's 8 images
};

// Pixel (1-rel) of a position (deg) in a beam image rotated by rot
// (radians); nonzero when the projection cannot place it.
__device__ int beam_pixel(const BeamImg &b, double x, double y, float rot, float *px, float *py) {
  const double dt = x - b.crval[0];
  if (dt > 180.0) x -= 360.0;
  if (dt < -180.0) x += 360.0;
  double d0 = x - b.crval[0], d1 = y - b.crval[1];
  if (b.cdelt[0] == 0.0f || b.cdelt[1] == 0.0f) return 2;
  if (b.proj != 0) {
    const double d2r = 1.745329252e-2;   // the reference's factor
    const double ra0 = b.crval[0] * d2r, dec0 = b.crval[1] * d2r, ra = x * d2r, dec = y * d2r;
    const double coss = cos(dec), sins = sin(dec), cda = cos(ra - ra0);
    double l = sin(ra - ra0) * coss;
    const double sint = sins * sin(dec0) + coss * cos(dec0) * cda;
    const double mm = sins * cos(dec0) - coss * sin(dec0) * cda;
    double m;
    if (b.proj == 1) {            // -SIN
      if (sint < 0.0) return 1;
      m = mm;
    } else if (b.proj == 2) {     // -TAN
      if (sint <= 0.0) return 1;
      l = l / sint;
      m = mm / sint;
    } else {                      // -ARC
      double c = fmin(1.0, fmax(-1.0, sint));
      c = acos(c);
      c = (c != 0.0) ? c / sin(c) : 1.0;
      l = l * c;
      m = mm * c;
    }
    d0 = l / d2r;
    d1 = m / d2r;
  }
  float sr, cr;
  sincosf(rot, &sr, &cr);
  const double dz = d0 * cr + d1 * sr;
  d1 = d1 * cr - d0 * sr;
  d0 = dz;
  *px = (float)(d0 / b.cdelt[0] + b.crpix[0]);
  *py = (float)(d1 / b.cdelt[1] + b.crpix[1]);
  return 0;
}

// Lagrange weights for 1-rel pixel p on an axis of n pixels.
__device__ __forceinline__ void lagrange(float p, int n, int hw, const float *denom, int *start, float *k) {
  const int iwid = 2 * hw + 1;
  int cen = (int)(p + 0.5f) - hw;
  cen = max(1, min(cen, n - iwid + 1)) - 1;
  *start = cen;
  const float xx = p - (float)cen;
  for (int j = 0; j < iwid; ++j) {
    float prod = denom[j];
    for (int i = 0; i < iwid; ++i)
      if (i != j) prod *= (xx - (float)(i + 1));
    k[j] = prod;
  }
}

// Interpolated value of one image plane at (px, py); blank outside the
// image or where blanks carry more than a tenth of the weight.
__device__ float beam_interp(const float *plane, const BeamImg &b, float px, float py, float fblank) {
  if (px < 1.0f || px > (float)b.nx || py < 1.0f || py > (float)b.ny) return fblank;
  float xk[9], yk[9];
  int xs, ys;
  lagrange(px, b.nx, b.hwidth, b.denom, &xs, xk);
  lagrange(py, b.ny, b.hwidth, b.denom, &ys, yk);
  const int iwid = 2 * b.hwidth + 1;
  float sum = 0.0f, sumwt = 0.0f;
  for (int j = 0; j < iwid; ++j) {
    const float *row = plane + (long long)(ys + j) * b.nx + xs;
    for (int i = 0; i < iwid; ++i) {
      const float v = row[i];
      if (v != fblank) {
        const float wt = xk[i] * yk[j];
        sumwt += wt;
        sum += v * wt;
      }
    }
  }
  return (sumwt > 0.90f) ? sum / sumwt : fblank;
}

// 1/sqrt of a cos^2-shaped primary-beam gain (the reference's InvBeam,
// ObitCUDABeamModel.cu:66-71); 100 beyond the first null.
__device__ __forceinline__ float inv_beam(float fact, float freq, float ang) {
  const float arg = fact * freq * ang;
  if (arg > 1.02f) return 100.0f;
  const float div = (float)(1.0 - 4.0 * ((double)arg * (double)arg));
  return div / cosf((float)(3.141592653589793 * (double)arg));
}

// One thread per (component, antenna type), looping over beam channels.
__global__ __launch_bounds__(256) void k_beam_jones(const float *__restrict__ images,
                                                    const BeamImg *__restrict__ imgs,
                                                    const float *__restrict__ pos,
                                                    const float *__restrict__ bfreq, float refSizeFact,
                                                    int ncomp, int nchan, float parAng, float fblank,
                                                    float *__restrict__ jones) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (k >= ncomp) return;
  const BeamImg b = imgs[t];
  float *J = jones + ((long long)t * ncomp + k) * nchan * 8;
  const double p0 = -(double)pos[2 * k], p1 = (double)pos[2 * k + 1];
  float px, py;
  if (beam_pixel(b, p0, p1, parAng, &px, &py) != 0) {
    for (int c = 0; c < nchan; ++c) {
      float *q = J + 8 * c;
      q[0] = 1.0f; q[1] = 0.0f; q[2] = 0.0f; q[3] = 0.0f;
      q[4] = 0.0f; q[5] = 0.0f; q[6] = 1.0f; q[7] = 0.0f;
    }
    return;
  }
  const float ang = sqrtf((float)(p0 * p0 + p1 * p1));
  const long long planeSz = (long long)b.nx * b.ny, imgSz = planeSz * nchan;
  for (int c = 0; c < nchan; ++c) {
    const float ipb = inv_beam(refSizeFact, bfreq[c], ang);
    const float *base = images + b.off + (long long)c * planeSz;
    // images in the reference's order: P1 re, P1 im, P2 re, P2 im,
    // X1 re, X1 im, X2 re, X2 im
    float P1r = ipb * beam_interp(base, b, px, py, fblank);
    float P2r = ipb * beam_interp(base + 2 * imgSz, b, px, py, fblank);
    float P1i, P2i, X1r, X1i, X2r, X2i;
    // the reference's sanity test (it tests P1r twice)
    if (ipb < 10.0f && fabsf(P1r) <= 1.3f && fabsf(P1r) <= 1.3f) {
      P1i = ipb * beam_interp(base + 1 * imgSz, b, px, py, fblank);
      P2i = ipb * beam_interp(base + 3 * imgSz, b, px, py, fblank);
      X1r = ipb * beam_interp(base + 4 * imgSz, b, px, py, fblank);
      X1i = ipb * beam_interp(base + 5 * imgSz, b, px, py, fblank);
      X2r = ipb * beam_interp(base + 6 * imgSz, b, px, py, fblank);
      X2i = ipb * beam_interp(base + 7 * imgSz, b, px, py, fblank);
    } else {
      P1r = P2r = 1.0f;
      P1i = P2i = 0.0f;
      X1r = X1i = X2r = X2i = 0.0f;
    }
    float *q = J + 8 * c;
    q[0] = P1r; q[1] = P1i; q[2] = X1r; q[3] = X1i;
    q[4] = X2r; q[5] = X2i; q[6] = P2r; q[7] = P2i;
  }
}

}  // namespace beam
}  // namespace obit

struct ObitHIPBeam {
  ObitHIPBeamDesc desc;
  hipStream_t stream = nullptr;
  int ncu = 0;
  bool gauss = false, anyFlux = false, anyTab = false, haveMod = false;
  double *d_freq = nullptr;   // fIF[nif], dfIF[nif]
  int *d_antType = nullptr, *d_jmap = nullptr;
  double *d_pos = nullptr;    // cx, cy, cz [3][ncomp]
  float *d_amp = nullptr, *d_g = nullptr;
  float *d_jones = nullptr;
  float2 *d_coh = nullptr;
  // beam images (ObitHIPBeamDFTSetBeams)
  float *d_images = nullptr, *d_bfreq = nullptr, *d_cpos = nullptr;
  float *d_tab = nullptr;     // per-channel amplitudes (ObitHIPBeamDFTSetAmps)
  obit::beam::BeamImg *d_imgs = nullptr;
  float refSizeFact = 0.0f, fblank = 0.0f;
};

namespace {

void release(ObitHIPBeam *h) {
  if (!h) return;
  void *ptrs[] = {h->d_freq, h->d_antType, h->d_jmap, h->d_pos, h->d_amp, h->d_g, h->d_jones, h->d_coh,
                  h->d_images, h->d_bfreq, h->d_cpos, h->d_imgs, h->d_tab};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete h;
}

// no component with flux: the data are left as they are (the reference
// kernels' gotSome)
bool beam_active(const ObitHIPBeam *h) {
  return h->desc.ncomp > 0 && (h->d_tab ? h->anyTab : h->anyFlux);
}

// The coherency products of the current Jones table at parallactic angle
// parAng (k_beam_coh), on stream.
int beam_coh(ObitHIPBeam *h, float parAng, hipStream_t stream) {
  using namespace obit;
  const ObitHIPBeamDesc &D = h->desc;
  const double c2 = std::cos(2.0 * (double)parAng), s2 = std::sin(2.0 * (double)parAng);
  const long long ncoh = (long long)D.ntype * D.ntype * D.ncomp * D.njchan;
  hipLaunchKernelGGL(beam::k_beam_coh, dim3((unsigned)((ncoh + 255) / 256)), dim3(256), 0, stream,
                     h->d_jones, h->d_coh, D.ntype, D.ncomp, D.njchan, D.stokType, (float)c2, (float)s2);
  OBIT_HIP_TRY(hipGetLastError());
  return kOK;
}

// The model applied to nrow rows at d_vis from the coherency products
// (k_beam_dft), on stream.
int beam_dft(ObitHIPBeam *h, float *d_vis, long long nrow, hipStream_t stream) {
  using namespace obit;
  const ObitHIPBeamDesc &D = h->desc;
  beam::Desc d;
  d.vis = d_vis;
  d.nrow = nrow;
  d.lrec = D.lrec; d.nrparm = D.nrparm;
  d.ilocu = D.ilocu; d.ilocv = D.ilocv; d.ilocw = D.ilocw;
  d.ilocb = D.ilocb; d.iloca1 = D.iloca1; d.iloca2 = D.iloca2;
  d.nif = D.nif; d.nchan = D.nchan; d.incf = D.incf; d.incif = D.incif;
  const int cpt = D.ncorr == 2 ? beam::kCPT2 : beam::kCPT4;
  d.nchunk = (D.nchan + cpt - 1) / cpt;
  const long long nrowblk = (nrow + beam::kRows - 1) / beam::kRows;
  if (nrowblk > 0x7fffffffLL) return fail(kInvalid, "ObitHIPBeamDFTCalc: too many rows");
  d.nrowblk = (int)nrowblk;
  d.nitems = nrowblk * D.nif * d.nchunk;
  d.ncomp = D.ncomp; d.ntype = D.ntype; d.nant = D.nant; d.njchan = D.njchan;
  d.npair = D.ntype * D.ntype;
  d.fIF = h->d_freq; d.dfIF = h->d_freq + D.nif;
  d.cx = h->d_pos; d.cy = h->d_pos + D.ncomp; d.cz = h->d_pos + 2 * (size_t)D.ncomp;
  d.camp = h->d_amp; d.cg = h->d_g;
  d.antType = h->d_antType; d.jmap = h->d_jmap;
  d.coh = h->d_coh;
  d.tab = h->d_tab;
  d.op = D.opType;
  for (int i = 0; i < 4; ++i) d.sf[i] = D.stokFact[i];
  d.chanb = D.chanb; d.chane = D.chane; d.IFb = D.IFb; d.IFe = D.IFe;
  long long grid = (long long)std::max(h->ncu, 1) * beam::kBlocksPerCU;
  if (grid > d.nitems) grid = d.nitems;
  void *args[] = {&d};
  const void *fn;
  const bool sp = h->d_tab != nullptr;
  if (D.ncorr == 2) {
    if (h->gauss) fn = sp ? (const void *)beam::k_beam_dft<2, beam::kCPT2, true, true> : (const void *)beam::k_beam_dft<2, beam::kCPT2, true, false>;
    else fn = sp ? (const void *)beam::k_beam_dft<2, beam::kCPT2, false, true> : (const void *)beam::k_beam_dft<2, beam::kCPT2, false, false>;
  } else {
    if (h->gauss) fn = sp ? (const void *)beam::k_beam_dft<4, beam::kCPT4, true, true> : (const void *)beam::k_beam_dft<4, beam::kCPT4, true, false>;
    else fn = sp ? (const void *)beam::k_beam_dft<4, beam::kCPT4, false, true> : (const void *)beam::k_beam_dft<4, beam::kCPT4, false, false>;
  }
  OBIT_HIP_TRY(hipLaunchKernel(fn, dim3((unsigned)grid), dim3(beam::kRows), args, 0, stream));
  return kOK;
}

}  // namespace

extern "C" {

int ObitHIPBeamDFTCreate(ObitHIPBeam **out, const ObitHIPBeamDesc *desc, const double *fIF,
                         const double *dfIF, const int *antType, const int *visJChan,
                         void *stream) {
  using namespace obit;
  if (!out || !desc || !fIF || !dfIF || !antType || !visJChan)
    return fail(kInvalid, "ObitHIPBeamDFTCreate: null argument");
  *out = nullptr;
  ObitHIPBeamDesc d = *desc;
  if (d.lrec < 1 || d.nrparm < 0 || d.nif < 1 || d.nchan < 1 || d.incf < 1 || d.incif < 1)
    return fail(kInvalid, "ObitHIPBeamDFTCreate: bad record layout");
  if (d.ncorr != 2 && d.ncorr != 4) return fail(kInvalid, "ObitHIPBeamDFTCreate: ncorr must be 2 or 4");
  const long long last = (long long)d.nrparm + (long long)(d.nif - 1) * d.incif +
                         (long long)(d.nchan - 1) * d.incf + 3LL * (d.ncorr - 1) + 2;
  if (last >= d.lrec) return fail(kInvalid, "ObitHIPBeamDFTCreate: correlations exceed the %d-float row", d.lrec);
  auto inrow = [&](int i) { return i >= 0 && i < d.lrec; };
  if (!inrow(d.ilocu) || !inrow(d.ilocv) || !inrow(d.ilocw))
    return fail(kInvalid, "ObitHIPBeamDFTCreate: u, v, w outside the row");
  if (d.ilocb < 0 && (!inrow(d.iloca1) || !inrow(d.iloca2)))
    return fail(kInvalid, "ObitHIPBeamDFTCreate: no baseline or antenna parameters");
  if (d.ilocb >= d.lrec) return fail(kInvalid, "ObitHIPBeamDFTCreate: baseline outside the row");
  if (d.ncomp < 0 || d.compLen < 4) return fail(kInvalid, "ObitHIPBeamDFTCreate: bad component table shape");
  if (d.ntype < 1 || d.ntype > beam::kMaxTypes || d.nant < 1 || d.njchan < 1)
    return fail(kInvalid, "ObitHIPBeamDFTCreate: bad antenna types / beam channels");
  if (d.stokType < 1 || d.stokType > 4) return fail(kInvalid, "ObitHIPBeamDFTCreate: stokType must be 1..4");
  if (d.opType < 0 || d.opType > 2) return fail(kInvalid, "ObitHIPBeamDFTCreate: bad opType");
  // a negative end selects through the last channel / IF
  if (d.chane < 0) d.chane = d.nchan - 1;
  if (d.IFe < 0) d.IFe = d.nif - 1;
  if (d.chanb < 0 || d.chane >= d.nchan || d.chanb > d.chane || d.IFb < 0 || d.IFe >= d.nif || d.IFb > d.IFe)
    return fail(kInvalid, "ObitHIPBeamDFTCreate: channel %d..%d / IF %d..%d selection outside %d x %d", d.chanb,
                d.chane, d.IFb, d.IFe, d.nchan, d.nif);
  for (int i = 0; i < d.nant; ++i)
    if (antType[i] < 0 || antType[i] >= d.ntype)
      return fail(kInvalid, "ObitHIPBeamDFTCreate: antenna %d has type %d of %d", i, antType[i], d.ntype);
  for (long long i = 0; i < (long long)d.nif * d.nchan; ++i)
    if (visJChan[i] < 0 || visJChan[i] >= d.njchan)
      return fail(kInvalid, "ObitHIPBeamDFTCreate: beam channel %d of entry %lld outside 0..%d", visJChan[i], i,
                  d.njchan - 1);
  ObitHIPBeam *h = new ObitHIPBeam();
  h->desc = d;
  h->stream = (hipStream_t)stream;
  int rc = kOK;
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    OBIT_HIP_TRY(hipMalloc(dst, std::max<size_t>(bytes, 16)));
    if (src && bytes) OBIT_HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return kOK;
  };
  std::vector<double> f(2 * (size_t)d.nif);
  for (int i = 0; i < d.nif; ++i) { f[i] = fIF[i]; f[d.nif + i] = dfIF[i]; }
  int dev = 0;
  if (!rc) rc = (hipGetDevice(&dev) == hipSuccess) ? kOK : fail(kHipError, "hipGetDevice failed");
  if (!rc && hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    rc = fail(kHipError, "hipDeviceGetAttribute failed");
  if (!rc) rc = up((void **)&h->d_freq, f.data(), f.size() * sizeof(double));
  if (!rc) rc = up((void **)&h->d_antType, antType, sizeof(int) * d.nant);
  if (!rc) rc = up((void **)&h->d_jmap, visJChan, sizeof(int) * (size_t)d.nif * d.nchan);
  const size_t nc = (size_t)std::max(d.ncomp, 1);
  if (!rc) rc = up((void **)&h->d_pos, nullptr, 3 * nc * sizeof(double));
  if (!rc) rc = up((void **)&h->d_amp, nullptr, nc * sizeof(float));
  if (!rc) rc = up((void **)&h->d_g, nullptr, 3 * nc * sizeof(float));
  if (!rc) rc = up((void **)&h->d_jones, nullptr, (size_t)d.ntype * nc * d.njchan * 8 * sizeof(float));
  if (!rc) rc = up((void **)&h->d_coh, nullptr, (size_t)d.ntype * d.ntype * nc * d.njchan * 4 * sizeof(float2));
  if (!rc) rc = ObitHIPBeamDFTSetJones(h, nullptr);
  if (rc) { release(h); return rc; }
  *out = h;
  return kOK;
}

int ObitHIPBeamDFTSetMod(ObitHIPBeam *h, const float *comps) {
  using namespace obit;
  if (!h || (!comps && h->desc.ncomp > 0)) return fail(kInvalid, "ObitHIPBeamDFTSetMod: null argument");
  // a previous call's kernels may still read the tables being replaced
  OBIT_HIP_TRY(hipStreamSynchronize(h->stream));
  const int n = h->desc.ncomp, L = h->desc.compLen;
  std::vector<double> pos(3 * (size_t)std::max(n, 1));
  std::vector<float> amp(std::max(n, 1)), g(3 * (size_t)std::max(n, 1));
  bool gauss = false, any = false;
  const double inv2pi = 1.0 / (2.0 * M_PI);
  for (int k = 0; k < n; ++k) {
    const float *c = comps + (size_t)k * L;
    amp[k] = c[0];
    any = any || c[0] != 0.0f;
    for (int a = 0; a < 3; ++a) pos[(size_t)a * n + k] = (double)c[1 + a] * inv2pi;
    for (int a = 0; a < 3; ++a) {
      g[(size_t)a * n + k] = L >= 7 ? c[4 + a] : 0.0f;
      gauss = gauss || g[(size_t)a * n + k] != 0.0f;
    }
  }
  if (n > 0) {
    OBIT_HIP_TRY(hipMemcpy(h->d_pos, pos.data(), 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    OBIT_HIP_TRY(hipMemcpy(h->d_amp, amp.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    OBIT_HIP_TRY(hipMemcpy(h->d_g, g.data(), 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  }
  h->gauss = gauss;
  h->anyFlux = any;
  h->haveMod = true;
  return kOK;
}

int ObitHIPBeamDFTSetJones(ObitHIPBeam *h, const float *jones) {
  using namespace obit;
  if (!h) return fail(kInvalid, "ObitHIPBeamDFTSetJones: null handle");
  OBIT_HIP_TRY(hipStreamSynchronize(h->stream));
  const size_t n = (size_t)h->desc.ntype * std::max(h->desc.ncomp, 1) * h->desc.njchan;
  if (jones) {
    OBIT_HIP_TRY(hipMemcpy(h->d_jones, jones, n * 8 * sizeof(float), hipMemcpyHostToDevice));
  } else {
    std::vector<float> unit(n * 8, 0.0f);
    for (size_t i = 0; i < n; ++i) { unit[8 * i] = 1.0f; unit[8 * i + 6] = 1.0f; }
    OBIT_HIP_TRY(hipMemcpy(h->d_jones, unit.data(), n * 8 * sizeof(float), hipMemcpyHostToDevice));
  }
  return kOK;
}

int ObitHIPBeamDFTCalc(ObitHIPBeam *h, float *d_vis, long long nrow, float parAng) {
  using namespace obit;
  if (!h || (!d_vis && nrow > 0) || nrow < 0) return fail(kInvalid, "ObitHIPBeamDFTCalc: bad arguments");
  if (!h->haveMod) return fail(kState, "ObitHIPBeamDFTCalc: no component table (ObitHIPBeamDFTSetMod)");
  if (!beam_active(h) || nrow == 0) return kOK;
  int rc = beam_coh(h, parAng, h->stream);
  if (!rc) rc = beam_dft(h, d_vis, nrow, h->stream);
  return rc;
}

int ObitHIPBeamDFTSetAmps(ObitHIPBeam *h, const float *amp) {
  using namespace obit;
  if (!h) return fail(kInvalid, "ObitHIPBeamDFTSetAmps: null handle");
  OBIT_HIP_TRY(hipStreamSynchronize(h->stream));
  if (!amp) {
    if (h->d_tab) OBIT_HIP_TRY(hipFree(h->d_tab));
    h->d_tab = nullptr;
    return kOK;
  }
  const size_t n = (size_t)std::max(h->desc.ncomp, 1) * h->desc.nif * h->desc.nchan;
  if (!h->d_tab) OBIT_HIP_TRY(hipMalloc(&h->d_tab, n * sizeof(float)));
  OBIT_HIP_TRY(hipMemcpy(h->d_tab, amp, n * sizeof(float), hipMemcpyHostToDevice));
  bool any = false;
  for (size_t i = 0; i < n && !any; ++i) any = amp[i] != 0.0f;
  h->anyTab = any;
  return kOK;
}

int ObitHIPBeamDFTGetJones(ObitHIPBeam *h, float *jones) {
  using namespace obit;
  if (!h || !jones) return fail(kInvalid, "ObitHIPBeamDFTGetJones: null argument");
  const size_t n = (size_t)h->desc.ntype * std::max(h->desc.ncomp, 1) * h->desc.njchan;
  OBIT_HIP_TRY(hipStreamSynchronize(h->stream));
  OBIT_HIP_TRY(hipMemcpy(jones, h->d_jones, n * 8 * sizeof(float), hipMemcpyDeviceToHost));
  return kOK;
}

int ObitHIPBeamDFTSetBeams(ObitHIPBeam *h, const ObitHIPBeamImageDesc *desc, const float *images,
                           const float *beamFreq, float refSizeFact, float fblank) {
  using namespace obit;
  if (!h || !desc || !images || !beamFreq) return fail(kInvalid, "ObitHIPBeamDFTSetBeams: null argument");
  OBIT_HIP_TRY(hipStreamSynchronize(h->stream));
  const int nt = h->desc.ntype, nch = h->desc.njchan;
  std::vector<beam::BeamImg> bi(nt);
  long long off = 0;
  for (int t = 0; t < nt; ++t) {
    const ObitHIPBeamImageDesc &d = desc[t];
    if (d.nchan != nch)
      return fail(kInvalid, "ObitHIPBeamDFTSetBeams: type %d has %d beam planes, the Jones table %d", t, d.nchan, nch);
    if (d.hwidth < 1 || d.hwidth > 4) return fail(kInvalid, "ObitHIPBeamDFTSetBeams: hwidth must be 1..4");
    if (d.nx < 2 * d.hwidth + 1 || d.ny < 2 * d.hwidth + 1)
      return fail(kInvalid, "ObitHIPBeamDFTSetBeams: type %d image smaller than the interpolation kernel", t);
    if (d.proj < 0 || d.proj > 3)
      return fail(kInvalid, "ObitHIPBeamDFTSetBeams: projection %d not supported (linear, SIN, TAN, ARC)", d.proj);
    beam::BeamImg &b = bi[t];
    b.nx = d.nx; b.ny = d.ny; b.proj = d.proj; b.hwidth = d.hwidth;
    b.crval[0] = d.crval[0]; b.crval[1] = d.crval[1];
    b.crpix[0] = d.crpix[0]; b.crpix[1] = d.crpix[1];
    b.cdelt[0] = d.cdelt[0]; b.cdelt[1] = d.cdelt[1];
    const int iwid = 2 * d.hwidth + 1;
    for (int j = 0; j < 10; ++j) b.denom[j] = 0.0f;
    for (int j = 1; j <= iwid; ++j) {
      float prod = 1.0f;
      for (int i = 1; i <= iwid; ++i)
        if (i != j) prod = prod * (float)(j - i);
      b.denom[j - 1] = 1.0f / prod;
    }
    b.off = off;
    off += 8LL * nch * d.nx * d.ny;
  }
  void *ptrs[] = {h->d_images, h->d_bfreq, h->d_imgs};
  for (void *p : ptrs)
    if (p) OBIT_HIP_TRY(hipFree(p));
  h->d_images = h->d_bfreq = nullptr;
  h->d_imgs = nullptr;
  OBIT_HIP_TRY(hipMalloc(&h->d_images, (size_t)off * sizeof(float)));
  OBIT_HIP_TRY(hipMemcpy(h->d_images, images, (size_t)off * sizeof(float), hipMemcpyHostToDevice));
  OBIT_HIP_TRY(hipMalloc(&h->d_bfreq, (size_t)nch * sizeof(float)));
  OBIT_HIP_TRY(hipMemcpy(h->d_bfreq, beamFreq, (size_t)nch * sizeof(float), hipMemcpyHostToDevice));
  OBIT_HIP_TRY(hipMalloc(&h->d_imgs, (size_t)nt * sizeof(beam::BeamImg)));
  OBIT_HIP_TRY(hipMemcpy(h->d_imgs, bi.data(), (size_t)nt * sizeof(beam::BeamImg), hipMemcpyHostToDevice));
  h->refSizeFact = refSizeFact;
  h->fblank = fblank;
  return kOK;
}

int ObitHIPBeamDFTJonesFromBeams(ObitHIPBeam *h, const float *pos, float parAng) {
  using namespace obit;
  if (!h || (!pos && h->desc.ncomp > 0)) return fail(kInvalid, "ObitHIPBeamDFTJonesFromBeams: null argument");
  if (!h->d_images) return fail(kState, "ObitHIPBeamDFTJonesFromBeams: no beam images (ObitHIPBeamDFTSetBeams)");
  const int n = h->desc.ncomp;
  if (n == 0) return kOK;
  OBIT_HIP_TRY(hipStreamSynchronize(h->stream));
  if (!h->d_cpos) OBIT_HIP_TRY(hipMalloc(&h->d_cpos, (size_t)n * 2 * sizeof(float)));
  OBIT_HIP_TRY(hipMemcpy(h->d_cpos, pos, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(beam::k_beam_jones, dim3((unsigned)((n + 255) / 256), (unsigned)h->desc.ntype), dim3(256), 0,
                     h->stream, h->d_images, h->d_imgs, h->d_cpos, h->d_bfreq, h->refSizeFact, n, h->desc.njchan,
                     parAng, h->fblank, h->d_jones);
  OBIT_HIP_TRY(hipGetLastError());
  return kOK;
}

int ObitHIPBeamDFTDestroy(ObitHIPBeam *h) {
  if (h && h->stream) (void)hipStreamSynchronize(h->stream);
  else if (h) (void)hipDeviceSynchronize();
  release(h);
  return obit::kOK;
}

}  // extern "C"

// ---------------- Obit's beam-model entry points (ObitHIPBeamModel.h) ----------------
//
// ObitCUDABeamModelDFTInit / SetMod / Calc / Shutdown over the handle
// above: the GPUBeamInfo the reference's host class fills
// (src/ObitGPUBeamModel.c:212-896) is read into a descriptor, the
// component rows into the handle's tables, and Calc streams the caller's
// pinned rows through its d_data[] buffers on its streams.

namespace {

struct BeamAbi {
  ObitHIPBeam *h = nullptr;
  ObitHIPBeamDesc desc{};
  std::vector<double> fIF, dfIF;
  std::vector<int> antType, visJChan;
  std::vector<float> fscale;    // d_freqScale as the caller holds it
  int kincf = 1, kincif = 1;
  std::vector<float> pos;       // [ncomp][2] beam offsets dX, dY (deg)
  const void *beamsFrom = nullptr;  // d_BeamInterp the beam images were read from
  hipEvent_t cohDone = nullptr;
};

std::mutex g_abiMu;
std::unordered_map<const GPUBeamInfo *, BeamAbi *> g_abi;

// the reference's checkCudaErrors reports and exits with EXIT_FAILURE
[[noreturn]] void abi_exit(const char *where) {
  std::fprintf(stderr, "obit_hip: %s: %s\n", where, ObitHIPLastError());
  std::fflush(stderr);
  std::exit(EXIT_FAILURE);
}

BeamAbi *abi_find(const GPUBeamInfo *g) {
  std::lock_guard<std::mutex> lk(g_abiMu);
  auto it = g_abi.find(g);
  return it == g_abi.end() ? nullptr : it->second;
}

template <class T>
int d2h(std::vector<T> &out, const void *src, size_t n, const char *what) {
  using namespace obit;
  if (!src && n) return fail(kInvalid, "%s: null device pointer", what);
  out.resize(n);
  if (n) OBIT_HIP_TRY(hipMemcpy(out.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
  return kOK;
}

// The record layout, frequency lines and antenna tables of gpuInfo.
int abi_init(const GPUBeamInfo *g, BeamAbi *st) {
  using namespace obit;
  if (!g || !g->h_visInfo || !g->h_antInfo) return fail(kInvalid, "null gpuInfo, h_visInfo or h_antInfo");
  const GPUBeamVisInfo &V = *g->h_visInfo;
  const GPUBeamAntInfo &A = *g->h_antInfo;
  const int nif = std::max(V.nIF, 1);
  if (V.nchan < 1 || V.nchan % nif) return fail(kInvalid, "nchan %d is not a multiple of nIF %d", V.nchan, nif);
  const int nch = V.nchan / nif;
  if (V.incs != 3) return fail(kInvalid, "Stokes increment %d (3 expected)", V.incs);
  if (V.nstok < 2) return fail(kInvalid, "%d Stokes (XX, YY at least expected)", V.nstok);
  st->kincf = V.kincf;
  st->kincif = V.kincif;
  auto fidx = [&](int c, int i) { return (long long)c * V.kincf + (long long)i * V.kincif; };
  if (fidx(nch - 1, nif - 1) >= (long long)nch * nif || fidx(nch - 1, 0) < 0 || fidx(0, nif - 1) < 0)
    return fail(kInvalid, "frequency table increments %d, %d outside %d x %d", V.kincf, V.kincif, nch, nif);
  int rc = d2h(st->fscale, V.d_freqScale, (size_t)nch * nif, "d_freqScale");
  if (rc) return rc;
  // one line per IF through its first and last channel
  st->fIF.assign(nif, 0.0);
  st->dfIF.assign(nif, 0.0);
  for (int i = 0; i < nif; ++i) {
    const double f0 = st->fscale[fidx(0, i)], f1 = st->fscale[fidx(nch - 1, i)];
    st->fIF[i] = f0;
    st->dfIF[i] = nch > 1 ? (f1 - f0) / (nch - 1) : 0.0;
    for (int c = 0; c < nch; ++c) {
      const double f = st->fscale[fidx(c, i)], lin = f0 + c * st->dfIF[i];
      if (!(std::fabs(f - lin) <= 1e-6 * std::fabs(lin)))
        return fail(kInvalid, "channel %d of IF %d: frequency factor %.9g is %.3g off the IF's line", c, i, f,
                    f - lin);
    }
  }
  rc = d2h(st->antType, A.d_AntType, (size_t)std::max(A.numAnt, 0), "d_AntType");
  if (!rc) rc = d2h(st->visJChan, A.d_visJChann, (size_t)nch * nif, "d_visJChann");
  if (rc) return rc;
  ObitHIPBeamDesc &d = st->desc;
  d = ObitHIPBeamDesc{};
  d.lrec = V.lenvis; d.nrparm = V.nrparm;
  d.ilocu = V.ilocu; d.ilocv = V.ilocv; d.ilocw = V.ilocw;
  d.ilocb = A.ilocb; d.iloca1 = A.iloca1; d.iloca2 = A.iloca2;
  d.nif = nif; d.nchan = nch; d.incf = V.incf; d.incif = V.incif;
  d.ncorr = V.nstok >= 4 ? 4 : 2;
  d.ntype = A.numAntType; d.nant = A.numAnt; d.njchan = A.numJChan;
  d.chanb = V.chanb; d.chane = V.chane; d.IFb = V.IFb; d.IFe = V.IFe;
  return kOK;
}

// The component rows of h_modelInfo into a new handle.
int abi_setmod(const GPUBeamInfo *g, BeamAbi *st) {
  using namespace obit;
  if (!g->h_modelInfo) return fail(kInvalid, "null h_modelInfo");
  const GPUBeamModelInfo &M = *g->h_modelInfo;
  const GPUBeamVisInfo &V = *g->h_visInfo;
  const bool gauss = M.type == GPUBeamModTypeGaussTSpec;
  if (M.type != GPUBeamModTypePointTSpec && !gauss)
    return fail(kInvalid, "model type %d not supported (PointTSpec, GaussTSpec: the rows ObitSkyModelVMBeamMF loads)",
                (int)M.type);
  const int base = gauss ? 10 : 7;
  const int n = M.nmodel, L = M.size, nspec = M.nSpec;
  if (n < 0 || nspec < 1 || L < base + 2 * nspec)
    return fail(kInvalid, "%d components of %d floats with %d subbands (%d floats needed)", n, L, nspec,
                base + 2 * nspec);
  if (M.stokType < 1 || M.stokType > 4 || M.opType < 0 || M.opType > 2)
    return fail(kInvalid, "stokType %d / opType %d", M.stokType, (int)M.opType);
  const int nif = st->desc.nif, nch = st->desc.nchan;
  std::vector<float> rows, frat;
  std::vector<int> spec;
  int rc = d2h(rows, M.d_model, (size_t)n * L, "d_model");
  if (!rc) rc = d2h(spec, M.d_specIndex, (size_t)nch * nif, "d_specIndex");
  if (!rc) rc = d2h(frat, V.d_freqRat, (size_t)nspec, "d_freqRat");
  if (rc) return rc;
  for (size_t f = 0; f < spec.size(); ++f)
    if (spec[f] < 0 || spec[f] >= nspec)
      return fail(kInvalid, "subband %d of frequency entry %zu outside 0..%d", spec[f], f, nspec - 1);
  const int clen = gauss ? 7 : 4;
  const size_t nc = (size_t)std::max(n, 1);
  std::vector<float> comps(nc * clen, 0.0f), amps(nc * nif * nch, 0.0f);
  st->pos.assign(2 * nc, 0.0f);
  for (int k = 0; k < n; ++k) {
    const float *r = rows.data() + (size_t)k * L;
    float *c = comps.data() + (size_t)k * clen;
    c[0] = 1.0f;  // the amplitude is the per-channel table below
    c[1] = r[4]; c[2] = r[5]; c[3] = r[6];
    if (gauss) { c[4] = r[7]; c[5] = r[8]; c[6] = r[9]; }
    st->pos[2 * k] = r[2];
    st->pos[2 * k + 1] = r[3];
    for (int i = 0; i < nif; ++i)
      for (int ch = 0; ch < nch; ++ch) {
        const long long f = (long long)ch * st->kincf + (long long)i * st->kincif;
        const int s = spec[f];
        const double flux = r[base + 2 * s], si = r[base + 2 * s + 1];
        amps[((size_t)k * nif + i) * nch + ch] =
            flux == 0.0 ? 0.0f : (float)(flux * std::pow((double)st->fscale[f] * (double)frat[s], si));
      }
  }
  ObitHIPBeamDesc d = st->desc;
  d.ncomp = n; d.compLen = clen;
  d.stokType = M.stokType; d.opType = (int)M.opType;
  for (int i = 0; i < 4; ++i) d.stokFact[i] = M.stokFact[i];
  if (st->h) {
    (void)ObitHIPBeamDFTDestroy(st->h);
    st->h = nullptr;
  }
  st->beamsFrom = nullptr;
  hipStream_t s0 = (g->nstream > 0 && g->stream) ? (hipStream_t)g->stream[0] : nullptr;
  rc = ObitHIPBeamDFTCreate(&st->h, &d, st->fIF.data(), st->dfIF.data(), st->antType.data(),
                            st->visJChan.data(), (void *)s0);
  if (!rc) rc = ObitHIPBeamDFTSetMod(st->h, comps.data());
  if (!rc) rc = ObitHIPBeamDFTSetAmps(st->h, amps.data());
  if (!rc) st->desc = d;
  return rc;
}

// Beam images and geometry from the caller's interpolators
// (CUDABeamInterp objects made by ObitGPUCUDABeamInterpCreate,
// src/ObitGPUBeamInterp.c:232-295): 8 per antenna type, the first one's
// image descriptor placing the components.
int abi_beams(const GPUBeamInfo *g, BeamAbi *st) {
  using namespace obit;
  const GPUBeamAntInfo &A = *g->h_antInfo;
  const int nt = st->desc.ntype, nch = st->desc.njchan;
  if (!A.d_BeamInterp) return fail(kState, "no beam interpolators (d_BeamInterp)");
  std::vector<CUDABeamInterp *> ptrs;
  int rc = d2h(ptrs, A.d_BeamInterp, (size_t)nt * 8, "d_BeamInterp");
  if (rc) return rc;
  std::vector<ObitHIPBeamImageDesc> desc(nt);
  std::vector<float> images;
  float fblank = 0.0f;
  static const char *kProj[8] = {"-SIN", "-TAN", "-ARC", "-NCP", "-GLS", "-MER", "-AIT", "-STG"};
  for (int t = 0; t < nt; ++t) {
    for (int k = 0; k < 8; ++k) {
      std::vector<CUDABeamInterp> bi;
      std::vector<CUDAFArray> fa;
      rc = d2h(bi, ptrs[t * 8 + k], 1, "beam interpolator");
      if (!rc) rc = d2h(fa, bi[0].d_myArray, 1, "beam array");
      if (rc) return rc;
      const int nx = fa[0].naxis[0], ny = fa[0].ndim > 1 ? fa[0].naxis[1] : 1;
      const int nz = fa[0].ndim > 2 ? fa[0].naxis[2] : 1;
      if (k == 0) {
        std::vector<CUDAImageDesc> id;
        rc = d2h(id, bi[0].d_myDesc, 1, "beam image descriptor");
        if (rc) return rc;
        ObitHIPBeamImageDesc &b = desc[t];
        b.nx = nx; b.ny = ny; b.nchan = nz;
        b.hwidth = bi[0].hwidth;
        b.proj = 0;
        for (int j = 0; j < 8; ++j)
          if (std::equal(kProj[j], kProj[j] + 4, &id[0].ctype[0][4])) b.proj = j + 1;
        if (b.proj > 3) return fail(kInvalid, "type %d beam projection %s not supported", t, kProj[b.proj - 1]);
        b.crval[0] = id[0].crval[0]; b.crval[1] = id[0].crval[1];
        b.crpix[0] = id[0].crpix[0]; b.crpix[1] = id[0].crpix[1];
        b.cdelt[0] = id[0].cdelt[0]; b.cdelt[1] = id[0].cdelt[1];
        if (nz != nch) return fail(kInvalid, "type %d beam has %d planes, numJChan %d", t, nz, nch);
        if (t == 0) {
          fblank = fa[0].fblank;
          images.reserve((size_t)nt * 8 * nch * nx * ny);
        }
      } else if (nx != desc[t].nx || ny != desc[t].ny || nz != nch) {
        return fail(kInvalid, "type %d beam image %d is %d x %d x %d, image 0 %d x %d x %d", t, k, nx, ny, nz,
                    desc[t].nx, desc[t].ny, nch);
      }
      std::vector<float> pix;
      rc = d2h(pix, fa[0].d_array, (size_t)nx * ny * nz, "beam pixels");
      if (rc) return rc;
      images.insert(images.end(), pix.begin(), pix.end());
    }
  }
  std::vector<float> bfreq;
  rc = d2h(bfreq, A.d_BeamFreq, (size_t)nch, "d_BeamFreq");
  if (!rc) rc = ObitHIPBeamDFTSetBeams(st->h, desc.data(), images.data(), bfreq.data(), A.refSizeFact, fblank);
  if (!rc) st->beamsFrom = A.d_BeamInterp;
  return rc;
}

int abi_calc(const GPUBeamInfo *g, BeamAbi *st, int doJones, int doUnit, float parAng, const long visRange[2]) {
  using namespace obit;
  ObitHIPBeam *h = st->h;
  if (!h) return fail(kState, "no model (ObitCUDABeamModelDFTSetMod)");
  if (!visRange) return fail(kInvalid, "null visRange");
  if (g->nstream < 1 || !g->stream || !g->d_data) return fail(kInvalid, "no streams or device buffers");
  const GPUBeamAntInfo &A = *g->h_antInfo;
  hipStream_t s0 = (hipStream_t)g->stream[0];
  if (h->stream != s0) {
    OBIT_HIP_TRY(hipStreamSynchronize(h->stream));
    h->stream = s0;
  }
  const size_t njones = (size_t)st->desc.ntype * std::max(st->desc.ncomp, 1) * st->desc.njchan * 8 * sizeof(float);
  int rc = kOK;
  if (doJones) {
    if (A.d_BeamInterp != st->beamsFrom) rc = abi_beams(g, st);
    if (!rc) rc = ObitHIPBeamDFTJonesFromBeams(h, st->pos.data(), parAng);
    if (!rc && A.d_Jones) OBIT_HIP_TRY(hipMemcpyAsync(A.d_Jones, h->d_jones, njones, hipMemcpyDeviceToDevice, s0));
  } else if (doUnit) {
    rc = ObitHIPBeamDFTSetJones(h, nullptr);
    if (!rc && A.d_Jones) OBIT_HIP_TRY(hipMemcpyAsync(A.d_Jones, h->d_jones, njones, hipMemcpyDeviceToDevice, s0));
  } else if (A.d_Jones) {
    // the table as the caller's d_Jones holds it
    OBIT_HIP_TRY(hipMemcpyAsync(h->d_jones, A.d_Jones, njones, hipMemcpyDeviceToDevice, s0));
  }
  if (rc) return rc;
  const long long r0 = visRange[0], r1 = visRange[1];
  if (r1 < r0 || !beam_active(h)) return kOK;
  if (r0 < 0) return fail(kInvalid, "visRange %lld..%lld", r0, r1);
  if (!g->h_data) return fail(kInvalid, "null h_data");
  const long long lenvis = st->desc.lrec;
  // rows of one d_data[] buffer ((nvis + 5) lenvis floats over nstream, as
  // the host class allocates them)
  const long long cap = (long long)(((size_t)(g->nvis + 5) * lenvis * sizeof(float) / g->nstream) /
                                    ((size_t)lenvis * sizeof(float)));
  if (cap < 1) return fail(kInvalid, "device buffers hold no row (nvis %d, %d streams)", g->nvis, g->nstream);
  rc = beam_coh(h, parAng, s0);
  if (rc) return rc;
  if (!st->cohDone) OBIT_HIP_TRY(hipEventCreateWithFlags(&st->cohDone, hipEventDisableTiming));
  OBIT_HIP_TRY(hipEventRecord(st->cohDone, s0));
  const int ns = g->nstream;
  int used = 0;
  for (long long a = r0, k = 0; a <= r1; a += cap, ++k) {
    const long long nrow = std::min(cap, r1 - a + 1);
    const int si = (int)(k % ns);
    hipStream_t s = (hipStream_t)g->stream[si];
    float *dv = g->d_data[si];
    float *hv = g->h_data + a * lenvis;
    const size_t bytes = (size_t)nrow * lenvis * sizeof(float);
    if (!dv) return fail(kInvalid, "d_data[%d] is null", si);
    OBIT_HIP_TRY(hipStreamWaitEvent(s, st->cohDone, 0));
    OBIT_HIP_TRY(hipMemcpyAsync(dv, hv, bytes, hipMemcpyHostToDevice, s));
    rc = beam_dft(h, dv, nrow, s);
    if (rc) return rc;
    OBIT_HIP_TRY(hipMemcpyAsync(hv, dv, bytes, hipMemcpyDeviceToHost, s));
    used = std::max(used, si + 1);
  }
  for (int i = 0; i < used; ++i) OBIT_HIP_TRY(hipStreamSynchronize((hipStream_t)g->stream[i]));
  return kOK;
}

}  // namespace

extern "C" {

void ObitCUDABeamModelDFTInit(GPUBeamInfo *gpuInfo) {
  BeamAbi *st = abi_find(gpuInfo);
  const bool fresh = st == nullptr;
  if (fresh) st = new BeamAbi();
  if (abi_init(gpuInfo, st)) {
    if (fresh) delete st;
    abi_exit("ObitCUDABeamModelDFTInit");
  }
  if (fresh) {
    std::lock_guard<std::mutex> lk(g_abiMu);
    g_abi[gpuInfo] = st;
  }
}

void ObitCUDABeamModelDFTSetMod(GPUBeamInfo *gpuInfo) {
  BeamAbi *st = abi_find(gpuInfo);
  if (!st) {
    obit::fail(obit::kState, "no layout (ObitCUDABeamModelDFTInit)");
    abi_exit("ObitCUDABeamModelDFTSetMod");
  }
  if (abi_setmod(gpuInfo, st)) abi_exit("ObitCUDABeamModelDFTSetMod");
}

void ObitCUDABeamModelDFTCalc(GPUBeamInfo *gpuInfo, int doJones, int doUnit, float parAng, long visRange[2]) {
  BeamAbi *st = abi_find(gpuInfo);
  if (!st) {
    obit::fail(obit::kState, "no layout (ObitCUDABeamModelDFTInit)");
    abi_exit("ObitCUDABeamModelDFTCalc");
  }
  if (abi_calc(gpuInfo, st, doJones, doUnit, parAng, visRange)) abi_exit("ObitCUDABeamModelDFTCalc");
}

void ObitCUDABeamModelDFTShutdown(GPUBeamInfo *gpuInfo) {
  BeamAbi *st = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_abiMu);
    auto it = g_abi.find(gpuInfo);
    if (it != g_abi.end()) {
      st = it->second;
      g_abi.erase(it);
    }
  }
  if (st) {
    if (st->h) (void)ObitHIPBeamDFTDestroy(st->h);
    if (st->cohDone) (void)hipEventDestroy(st->cohDone);
    delete st;
  }
  (void)hipDeviceSynchronize();
}

}  // extern "C"
