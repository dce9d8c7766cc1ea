/*
 * oracle.c -- CPU oracle for the visibility-domain hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Reference paths below are
 * relative to ObitSystem/Obit/ of bill-cotton/Obit.
 *
 * Floating-point contract: compiled with -ffp-contract=off and no -march,
 * so every float operation is an IEEE single operation in SSE registers,
 * as in Obit's stock "-O3 -msse" build (configure.ac:38, m4/msse.m4:15).
 * C promotion rules are kept exactly: where the reference mixes ofloat
 * and odouble the same casts appear here.
 */
#include "oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* FP32 vector math as a stock HAVE_SSE build evaluates it             */
/* ------------------------------------------------------------------ */

typedef union { float f; uint32_t u; int32_t i; } fbits;

/* Cephes single-precision constants as the SSE header stores them
 * (float-rounded); include/sse_mathfun.h:200-212, 292-301.  These are
 * the published Cephes sinf/cosf/expf coefficients. */
static const float C_FOPI = 1.27323954473516f;
static const float C_MDP1 = -0.78515625f;
static const float C_MDP2 = -2.4187564849853515625e-4f;
static const float C_MDP3 = -3.77489497744594108e-8f;
static const float C_SIN0 = -1.9515295891E-4f;
static const float C_SIN1 = 8.3321608736E-3f;
static const float C_SIN2 = -1.6666654611E-1f;
static const float C_COS0 = 2.443315711809948E-005f;
static const float C_COS1 = -1.388731625493765E-003f;
static const float C_COS2 = 4.166664568298827E-002f;
static const float C_EXPHI = 88.3762626647949f;
static const float C_EXPLO = -88.3762626647949f;
static const float C_LOG2EF = 1.44269504088896341f;
static const float C_EXPC1 = 0.693359375f;
static const float C_EXPC2 = -2.12194440e-4f;
static const float C_EXPP0 = 1.9875691500E-4f;
static const float C_EXPP1 = 1.3981999507E-3f;
static const float C_EXPP2 = 8.3334519073E-3f;
static const float C_EXPP3 = 4.1665795894E-2f;
static const float C_EXPP4 = 1.6666665459E-1f;
static const float C_EXPP5 = 5.0000001201E-1f;

/* One lane of sincos_ps (include/sse_mathfun.h:568-711): octant
 * reduction by 4/pi with j=(j+1)&~1, three-part Cody-Waite subtraction,
 * degree-3 polynomials in z=x^2, selection by octant, sign by bit tricks.
 * Pinned by: the cephes error bound (|err| < 2^-21 vs libm on |x|<1e4)
 * and the golden points in tests/test_oracle_math.py. */
void orc_sincos_cephes(float xin, float *s, float *c)
{
  fbits xb, sb, yb;
  xb.f = xin;
  uint32_t sign_sin = xb.u & 0x80000000u;
  xb.u &= 0x7fffffffu;
  float x = xb.f;
  float y = x * C_FOPI;
  int32_t j = (int32_t)y;              /* cvttps: truncation */
  j = (j + 1) & ~1;
  y = (float)j;
  uint32_t swap_sin = ((uint32_t)(j & 4)) << 29;
  int poly_mask = ((j & 2) == 0);
  float t1 = y * C_MDP1, t2 = y * C_MDP2, t3 = y * C_MDP3;
  x = x + t1;
  x = x + t2;
  x = x + t3;
  uint32_t sign_cos = ((uint32_t)(~(j - 2) & 4)) << 29;
  sign_sin ^= swap_sin;
  float z = x * x;
  float yc = C_COS0;
  yc = yc * z; yc = yc + C_COS1;
  yc = yc * z; yc = yc + C_COS2;
  yc = yc * z; yc = yc * z;
  float tmp = z * 0.5f;
  yc = yc - tmp;
  yc = yc + 1.0f;
  float ys = C_SIN0;
  ys = ys * z; ys = ys + C_SIN1;
  ys = ys * z; ys = ys + C_SIN2;
  ys = ys * z; ys = ys * x;
  ys = ys + x;
  /* select: ysin2 = mask&ys, ysin1 = ~mask&yc, then the same add/sub */
  float ysin2 = poly_mask ? ys : 0.0f;
  float ysin1 = poly_mask ? 0.0f : yc;
  float ys_r = ys - ysin2;
  float yc_r = yc - ysin1;
  sb.f = ysin1 + ysin2;
  yb.f = yc_r + ys_r;
  sb.u ^= sign_sin;
  yb.u ^= sign_cos;
  *s = sb.f;
  *c = yb.f;
}

/* One lane of exp_ps (include/sse_mathfun.h:214-289): clamp, n=floor(x
 * log2e + 1/2), two-part ln2 subtraction, degree-5 polynomial, scale by
 * 2^n built in the exponent field.  Pinned by: |err| vs libm exp and the
 * golden points in tests/test_oracle_math.py. */
float orc_exp_cephes(float xin)
{
  float x = xin;
  /* minps/maxps with the constant as second operand */
  x = (x < C_EXPHI) ? x : C_EXPHI;
  x = (x > C_EXPLO) ? x : C_EXPLO;
  float fx = x * C_LOG2EF;
  fx = fx + 0.5f;
  int32_t e0 = (int32_t)fx;
  float tmp = (float)e0;
  float mask = (tmp > fx) ? 1.0f : 0.0f;
  fx = tmp - mask;
  tmp = fx * C_EXPC1;
  float z = fx * C_EXPC2;
  x = x - tmp;
  x = x - z;
  z = x * x;
  float y = C_EXPP0;
  y = y * x; y = y + C_EXPP1;
  y = y * x; y = y + C_EXPP2;
  y = y * x; y = y + C_EXPP3;
  y = y * x; y = y + C_EXPP4;
  y = y * x; y = y + C_EXPP5;
  y = y * z;
  y = y + x;
  y = y + 1.0f;
  int32_t e = (int32_t)fx;
  e = e + 0x7f;
  fbits p; p.u = ((uint32_t)e) << 23;
  return y * p.f;
}

#ifdef ORC_SSE4LANE
/* The same two functions four lanes at a time in SSE2 registers, as the
 * reference's sincos_ps / exp_ps evaluate them (include/sse_mathfun.h:
 * 214-289, 568-711): the identical operation sequence per lane, so every
 * lane is bit-identical to orc_sincos_cephes / orc_exp_cephes.  Built only
 * into the CPU-baseline library (liboracle_cpu.so), which times Obit's
 * CPU path as a stock "-O3 -msse" build runs it; the checker keeps the
 * scalar form.  Tested bit for bit against the scalar lanes. */
#include <emmintrin.h>

static void sincos4(const float *in, float *so, float *co)
{
  const __m128i signbit = _mm_set1_epi32((int)0x80000000u);
  __m128 xin = _mm_loadu_ps(in);
  __m128 sign_sin = _mm_and_ps(xin, _mm_castsi128_ps(signbit));
  __m128 x = _mm_andnot_ps(_mm_castsi128_ps(signbit), xin);
  __m128 y = _mm_mul_ps(x, _mm_set1_ps(C_FOPI));
  __m128i j = _mm_cvttps_epi32(y);
  j = _mm_add_epi32(j, _mm_set1_epi32(1));
  j = _mm_and_si128(j, _mm_set1_epi32(~1));
  y = _mm_cvtepi32_ps(j);
  __m128i swap_sin = _mm_slli_epi32(_mm_and_si128(j, _mm_set1_epi32(4)), 29);
  __m128i poly = _mm_cmpeq_epi32(_mm_and_si128(j, _mm_set1_epi32(2)), _mm_setzero_si128());
  __m128 t1 = _mm_mul_ps(y, _mm_set1_ps(C_MDP1));
  __m128 t2 = _mm_mul_ps(y, _mm_set1_ps(C_MDP2));
  __m128 t3 = _mm_mul_ps(y, _mm_set1_ps(C_MDP3));
  x = _mm_add_ps(x, t1);
  x = _mm_add_ps(x, t2);
  x = _mm_add_ps(x, t3);
  __m128i sign_cos = _mm_slli_epi32(_mm_andnot_si128(_mm_sub_epi32(j, _mm_set1_epi32(2)),
                                                     _mm_set1_epi32(4)), 29);
  sign_sin = _mm_xor_ps(sign_sin, _mm_castsi128_ps(swap_sin));
  __m128 z = _mm_mul_ps(x, x);
  __m128 yc = _mm_set1_ps(C_COS0);
  yc = _mm_mul_ps(yc, z); yc = _mm_add_ps(yc, _mm_set1_ps(C_COS1));
  yc = _mm_mul_ps(yc, z); yc = _mm_add_ps(yc, _mm_set1_ps(C_COS2));
  yc = _mm_mul_ps(yc, z); yc = _mm_mul_ps(yc, z);
  __m128 tmp = _mm_mul_ps(z, _mm_set1_ps(0.5f));
  yc = _mm_sub_ps(yc, tmp);
  yc = _mm_add_ps(yc, _mm_set1_ps(1.0f));
  __m128 ys = _mm_set1_ps(C_SIN0);
  ys = _mm_mul_ps(ys, z); ys = _mm_add_ps(ys, _mm_set1_ps(C_SIN1));
  ys = _mm_mul_ps(ys, z); ys = _mm_add_ps(ys, _mm_set1_ps(C_SIN2));
  ys = _mm_mul_ps(ys, z); ys = _mm_mul_ps(ys, x);
  ys = _mm_add_ps(ys, x);
  __m128 pm = _mm_castsi128_ps(poly);
  __m128 ysin2 = _mm_and_ps(pm, ys);
  __m128 ysin1 = _mm_andnot_ps(pm, yc);
  __m128 ys_r = _mm_sub_ps(ys, ysin2);
  __m128 yc_r = _mm_sub_ps(yc, ysin1);
  __m128 sn = _mm_add_ps(ysin1, ysin2);
  __m128 cs = _mm_add_ps(yc_r, ys_r);
  _mm_storeu_ps(so, _mm_xor_ps(sn, sign_sin));
  _mm_storeu_ps(co, _mm_xor_ps(cs, _mm_castsi128_ps(sign_cos)));
}

static void exp4(const float *in, float *out)
{
  __m128 x = _mm_loadu_ps(in);
  x = _mm_min_ps(x, _mm_set1_ps(C_EXPHI));
  x = _mm_max_ps(x, _mm_set1_ps(C_EXPLO));
  __m128 fx = _mm_mul_ps(x, _mm_set1_ps(C_LOG2EF));
  fx = _mm_add_ps(fx, _mm_set1_ps(0.5f));
  __m128 tmp = _mm_cvtepi32_ps(_mm_cvttps_epi32(fx));
  __m128 mask = _mm_and_ps(_mm_cmpgt_ps(tmp, fx), _mm_set1_ps(1.0f));
  fx = _mm_sub_ps(tmp, mask);
  tmp = _mm_mul_ps(fx, _mm_set1_ps(C_EXPC1));
  __m128 z = _mm_mul_ps(fx, _mm_set1_ps(C_EXPC2));
  x = _mm_sub_ps(x, tmp);
  x = _mm_sub_ps(x, z);
  z = _mm_mul_ps(x, x);
  __m128 y = _mm_set1_ps(C_EXPP0);
  y = _mm_mul_ps(y, x); y = _mm_add_ps(y, _mm_set1_ps(C_EXPP1));
  y = _mm_mul_ps(y, x); y = _mm_add_ps(y, _mm_set1_ps(C_EXPP2));
  y = _mm_mul_ps(y, x); y = _mm_add_ps(y, _mm_set1_ps(C_EXPP3));
  y = _mm_mul_ps(y, x); y = _mm_add_ps(y, _mm_set1_ps(C_EXPP4));
  y = _mm_mul_ps(y, x); y = _mm_add_ps(y, _mm_set1_ps(C_EXPP5));
  y = _mm_mul_ps(y, z);
  y = _mm_add_ps(y, x);
  y = _mm_add_ps(y, _mm_set1_ps(1.0f));
  __m128i e = _mm_add_epi32(_mm_cvttps_epi32(fx), _mm_set1_epi32(0x7f));
  __m128 p2n = _mm_castsi128_ps(_mm_slli_epi32(e, 23));
  _mm_storeu_ps(out, _mm_mul_ps(y, p2n));
}
#endif

/* ObitSinCosVec, stock SSE build (src/ObitSinCos.c:105-186): a single
 * element goes to libm sinf/cosf (include/ObitSinCos.h:59-62); otherwise
 * each lane is sincos_ps (zero-padding of the tail does not change the
 * lanes that are used). */
void orc_sincos_vec(int n, const float *ang, float *s, float *c)
{
  if (n <= 0) return;
  if (n == 1) { s[0] = sinf(ang[0]); c[0] = cosf(ang[0]); return; }
#ifdef ORC_SSE4LANE
  int i = 0;
  for (; i + 4 <= n; i += 4) sincos4(ang + i, s + i, c + i);
  if (i < n) {
    float a[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ss[4], cc[4];
    for (int k = i; k < n; k++) a[k - i] = ang[k];
    sincos4(a, ss, cc);
    for (int k = i; k < n; k++) { s[k] = ss[k - i]; c[k] = cc[k - i]; }
  }
#else
  for (int i = 0; i < n; i++) orc_sincos_cephes(ang[i], &s[i], &c[i]);
#endif
}

/* ObitExpVec, SSE branch (src/ObitExp.c:106-129): every element through
 * exp_ps, including a single one. */
void orc_exp_vec(int n, const float *arg, float *out)
{
#ifdef ORC_SSE4LANE
  int i = 0;
  for (; i + 4 <= n; i += 4) exp4(arg + i, out + i);
  if (i < n) {
    float a[4] = {0.0f, 0.0f, 0.0f, 0.0f}, o[4];
    for (int k = i; k < n; k++) a[k - i] = arg[k];
    exp4(a, o);
    for (int k = i; k < n; k++) out[k] = o[k - i];
  }
#else
  for (int i = 0; i < n; i++) out[i] = orc_exp_cephes(arg[i]);
#endif
}

/* ------------------------------------------------------------------ */
/* DFT: restatement of ThreadSkyModelFTDFT / ThreadSkyModelMFFTDFT     */
/* ------------------------------------------------------------------ */

static int count_mcomp(const orc_dft_desc *d)
{
  int mcomp = 0;
  for (int k = 0; k < d->ncomp; k++)
    if (d->comps[(size_t)k * d->lcomp] != 0.0f) mcomp = k + 1;
  return mcomp;
}

/* Model application shared by both classes: src/ObitSkyModel.c:2794-2844,
 * src/ObitSkyModelMF.c:2377-2425. */
static void apply_model(const orc_dft_desc *d, float *visData, int offsetChannel,
                        double sumReal, double sumImag)
{
  float modReal, modImag, wt = 0.0f, temp;
  if (d->doFlip) { modReal = -(float)sumImag; modImag = (float)sumReal; }
  else           { modReal =  (float)sumReal; modImag = (float)sumImag; }
  if (d->doDivide) {
    wt = modReal * modReal + modImag * modImag;
    modReal /= wt;
    modImag /= wt;
    wt = (float)sqrt((double)wt);
  }
  for (int iStoke = d->startPoln; iStoke < d->startPoln + d->numPoln; iStoke++) {
    int offset = offsetChannel + iStoke * d->incs;
    /* a flagged Stokes is skipped before the factor update, as in the
     * reference, so the next Stokes sees the unscaled model */
    if ((visData[offset + 2] <= 0.0f) && !d->doReplace) continue;
    if (d->doDivide) {
      temp = modReal * visData[offset] + modImag * visData[offset + 1];
      visData[offset + 1] = modReal * visData[offset + 1] - modImag * visData[offset];
      visData[offset] = temp;
      visData[offset + 2] *= wt;
    } else if (d->doReplace) {
      visData[offset] = modReal;
      visData[offset + 1] = modImag;
    } else {
      visData[offset] -= modReal;
      visData[offset + 1] -= modImag;
    }
    modReal *= d->stokFactor;
    modImag *= d->stokFactor;
  }
}

#define FAZ_BASE 100   /* src/ObitSkyModel.c:2389 */
#define FAZ_MF   256   /* src/ObitSkyModelMF.c:1993 */

/* src/ObitSkyModel.c:2354-2857, model types Point, PointSpec, Gauss,
 * GaussSpec.  Pinned by: FP64 closed form (tests), UVSub golden master. */
static void dft_rows_base(const orc_dft_desc *d, float *buffer, int loVis, int hiVis)
{
  float AmpArr[FAZ_BASE], FazArr[FAZ_BASE], CosArr[FAZ_BASE], SinArr[FAZ_BASE];
  float ExpArg[FAZ_BASE], ExpVal[FAZ_BASE], ExpArg2[FAZ_BASE], ExpVal2[FAZ_BASE];
  const int lcomp = d->lcomp, nterm = d->nSpecTerm;
  const int mcomp = count_mcomp(d);
  if (d->ncomp <= 0) return;
  const double SMRefFreq = 1.0 / d->smRefFreq;
  for (int iVis = loVis; iVis < hiVis; iVis++) {
    float *visData = buffer + (size_t)iVis * d->lrec;
    for (int iIF = d->startIF; iIF < d->startIF + d->numIF; iIF++) {
      int offsetIF = d->nrparm + iIF * d->incif;
      for (int iChannel = d->startChan; iChannel < d->startChan + d->numChan; iChannel++) {
        int offsetChannel = offsetIF + iChannel * d->incf;
        int OK = 0;
        for (int iStoke = d->startPoln; iStoke < d->startPoln + d->numPoln; iStoke++) {
          int offset = offsetChannel + iStoke * d->incs;
          OK = OK || (visData[offset + 2] > 0.0f);
        }
        if (!OK && !d->doReplace) continue;
        int ifq = iIF * d->kincif + iChannel * d->kincf;
        float freqFact = d->fscale[ifq];
        double specFreqFact = d->freqArr[ifq] * SMRefFreq;
        double lnspecFreqFact = log(specFreqFact);
        double u = (double)visData[d->ilocu] * freqFact;
        double v = (double)visData[d->ilocv] * freqFact;
        double w = (double)visData[d->ilocw] * freqFact;
        double sumReal = 0.0, sumImag = 0.0;
        const float *ccData = d->comps;
        int it, itcnt, lim, iComp, jt, iterm;
        float amp, arg, ll, lll, specFact;
        double tx, ty, tz;
        switch (d->modType) {
        case ORC_PointMod:
          for (it = 0; it < mcomp; it += FAZ_BASE) {
            itcnt = 0;
            lim = (mcomp < it + FAZ_BASE) ? mcomp : it + FAZ_BASE;
            for (iComp = it; iComp < lim; iComp++) {
              AmpArr[itcnt] = ccData[0];
              tx = ccData[1] * u; ty = ccData[2] * v; tz = ccData[3] * w;
              FazArr[itcnt] = (float)(tx + ty + tz);
              ccData += lcomp; itcnt++;
            }
            orc_sincos_vec(itcnt, FazArr, SinArr, CosArr);
            for (jt = 0; jt < itcnt; jt++) {
              sumReal += AmpArr[jt] * CosArr[jt];
              sumImag += AmpArr[jt] * SinArr[jt];
            }
          }
          break;
        case ORC_PointModSpec:
          for (it = 0; it < mcomp; it += FAZ_BASE) {
            itcnt = 0;
            lim = (mcomp < it + FAZ_BASE) ? mcomp : it + FAZ_BASE;
            for (iComp = it; iComp < lim; iComp++) {
              if (ccData[0] != 0.0f) {
                tx = ccData[1] * u; ty = ccData[2] * v; tz = ccData[3] * w;
                lll = ll = (float)lnspecFreqFact;
                arg = 0.0f;
                for (iterm = 0; iterm < nterm; iterm++) { arg += ccData[4 + iterm] * lll; lll *= ll; }
                ExpArg2[itcnt] = arg;
                AmpArr[itcnt] = ccData[0];
                FazArr[itcnt] = (float)(tx + ty + tz);
              } else { ExpArg2[itcnt] = 0.0f; FazArr[itcnt] = 0.0f; AmpArr[itcnt] = 0.0f; }
              ccData += lcomp; itcnt++;
            }
            orc_sincos_vec(itcnt, FazArr, SinArr, CosArr);
            orc_exp_vec(itcnt, ExpArg2, ExpVal2);
            for (jt = 0; jt < itcnt; jt++) {
              AmpArr[jt] *= ExpVal2[jt];
              sumReal += AmpArr[jt] * CosArr[jt];
              sumImag += AmpArr[jt] * SinArr[jt];
            }
          }
          break;
        case ORC_GaussMod:
          for (it = 0; it < mcomp; it += FAZ_BASE) {
            lim = (mcomp < it + FAZ_BASE) ? mcomp : it + FAZ_BASE;
            itcnt = 0;
            for (iComp = it; iComp < lim; iComp++) {
              if (ccData[0] != 0.0f) {
                amp = ccData[0];
                arg = (float)(ccData[4] * u * u + ccData[5] * v * v + ccData[6] * u * v);
                tx = ccData[1] * u; ty = ccData[2] * v; tz = ccData[3] * w;
                ExpArg[itcnt] = arg;
                FazArr[itcnt] = (float)(tx + ty + tz);
                AmpArr[itcnt] = amp;
              } else { ExpArg[itcnt] = 0.0f; FazArr[itcnt] = 0.0f; AmpArr[itcnt] = 0.0f; }
              ccData += lcomp; itcnt++;
            }
            orc_sincos_vec(itcnt, FazArr, SinArr, CosArr);
            orc_exp_vec(itcnt, ExpArg, ExpVal);
            for (jt = 0; jt < itcnt; jt++) {
              amp = AmpArr[jt] * ExpVal[jt];
              sumReal += amp * CosArr[jt];
              sumImag += amp * SinArr[jt];
            }
          }
          break;
        case ORC_GaussModSpec:
          for (it = 0; it < mcomp; it += FAZ_BASE) {
            itcnt = 0;
            lim = (mcomp < it + FAZ_BASE) ? mcomp : it + FAZ_BASE;
            for (iComp = it; iComp < lim; iComp++) {
              if (ccData[0] != 0.0f) {
                lll = ll = (float)lnspecFreqFact;
                arg = 0.0f;
                for (iterm = 0; iterm < nterm; iterm++) { arg += ccData[7 + iterm] * lll; lll *= ll; }
                specFact = (float)exp((double)arg);
                amp = specFact * ccData[0];
                arg = (float)(ccData[4] * u * u + ccData[5] * v * v + ccData[6] * u * v);
                tx = ccData[1] * u; ty = ccData[2] * v; tz = ccData[3] * w;
                ExpArg[itcnt] = arg;
                FazArr[itcnt] = (float)(tx + ty + tz);
                AmpArr[itcnt] = amp;
              } else { ExpArg[itcnt] = 0.0f; FazArr[itcnt] = 0.0f; AmpArr[itcnt] = 0.0f; }
              ccData += lcomp; itcnt++;
            }
            orc_sincos_vec(itcnt, FazArr, SinArr, CosArr);
            orc_exp_vec(itcnt, ExpArg, ExpVal);
            for (jt = 0; jt < itcnt; jt++) {
              amp = AmpArr[jt] * ExpVal[jt];
              sumReal += amp * CosArr[jt];
              sumImag += amp * SinArr[jt];
            }
          }
          break;
        default:
          return;  /* other types are outside this path */
        }
        apply_model(d, visData, offsetChannel, sumReal, sumImag);
      }
    }
  }
}

/* SkyModel2DDot scalar tail (src/ObitSkyModelMF.c:1835-1901): the scalars
 * arrive as ofloat, products in double, result rounded to float. */
static void mf_dot3(int n, const float *v1, const float *v2, const float *v3,
                    float s1, float s2, float s3, float *ov)
{
  double ds1 = s1, ds2 = s2, ds3 = s3;
  for (int i = 0; i < n; i++) ov[i] = (float)(v1[i] * ds1 + v2[i] * ds2 + v3[i] * ds3);
}
/* SkyModelDot scalar tail (src/ObitSkyModelMF.c:1795-1820): FP32 sum. */
static float mf_dot(int n, const float *v1, const float *v2)
{
  float out = 0.0f;
  for (int i = 0; i < n; i++) out += v1[i] * v2[i];
  return out;
}

/* src/ObitSkyModelMF.c:1971-2440, model types Point, PointTSpec, Gauss,
 * GaussTSpec.  The MF Gaussian branch negates the exponent argument
 * (src/ObitSkyModelMF.c:2183); this restatement keeps the base-class sign
 * (src/ObitSkyModel.c:2570) so that both classes model the same sky. */
static void dft_rows_mf(const orc_dft_desc *d, float *buffer, int loVis, int hiVis)
{
  float AmpArr[FAZ_MF], FazArr[FAZ_MF], CosArr[FAZ_MF], SinArr[FAZ_MF];
  float ExpArg[FAZ_MF], ExpVal[FAZ_MF], ExpArg2[FAZ_MF], ExpVal2[FAZ_MF];
  float VecL[FAZ_MF], VecM[FAZ_MF], VecN[FAZ_MF];
  const int lcomp = d->lcomp;
  const int mcomp = count_mcomp(d);
  if (d->ncomp <= 0) return;
  for (int iVis = loVis; iVis < hiVis; iVis++) {
    float *visData = buffer + (size_t)iVis * d->lrec;
    for (int iIF = d->startIF; iIF < d->startIF + d->numIF; iIF++) {
      int offsetIF = d->nrparm + iIF * d->incif;
      for (int iChannel = d->startChan; iChannel < d->startChan + d->numChan; iChannel++) {
        int offsetChannel = offsetIF + iChannel * d->incf;
        int OK = 0;
        for (int iStoke = d->startPoln; iStoke < d->startPoln + d->numPoln; iStoke++) {
          int offset = offsetChannel + iStoke * d->incs;
          OK = OK || (visData[offset + 2] > 0.0f);
        }
        if (!OK && !d->doReplace) continue;
        int ifq = iIF * d->kincif + iChannel * d->kincf;
        float freqFact = d->fscale[ifq];
        int itab = d->specIndex ? d->specIndex[ifq] : 0;
        float logNuONu0 = d->specFreq ? (float)log(d->freqArr[ifq] / d->specFreq[itab]) : 0.0f;
        double u = (double)visData[d->ilocu] * freqFact;
        double v = (double)visData[d->ilocv] * freqFact;
        double w = (double)visData[d->ilocw] * freqFact;
        double sumReal = 0.0, sumImag = 0.0;
        const float *ccData = d->comps;
        int it, itcnt, lim, iComp;
        float amp, arg;
        switch (d->modType) {
        case ORC_PointMod:
          for (it = 0; it < mcomp; it += FAZ_MF) {
            itcnt = 0;
            lim = (mcomp < it + FAZ_MF) ? mcomp : it + FAZ_MF;
            for (iComp = it; iComp < lim; iComp++) {
              if (ccData[0] != 0.0f) {
                AmpArr[itcnt] = ccData[0];
                VecL[itcnt] = ccData[1]; VecM[itcnt] = ccData[2]; VecN[itcnt] = ccData[3];
                itcnt++;
              }
              ccData += lcomp;
            }
            mf_dot3(itcnt, VecL, VecM, VecN, (float)u, (float)v, (float)w, FazArr);
            orc_sincos_vec(itcnt, FazArr, SinArr, CosArr);
            sumReal += mf_dot(itcnt, AmpArr, CosArr);
            sumImag += mf_dot(itcnt, AmpArr, SinArr);
          }
          break;
        case ORC_PointModTSpec: {
          int jtab = 5 + itab;
          for (it = 0; it < mcomp; it += FAZ_MF) {
            itcnt = 0;
            lim = (mcomp < it + FAZ_MF) ? mcomp : it + FAZ_MF;
            for (iComp = it; iComp < lim; iComp++) {
              if (ccData[jtab] != 0.0f) {
                VecL[itcnt] = ccData[1]; VecM[itcnt] = ccData[2]; VecN[itcnt] = ccData[3];
                ExpArg2[itcnt] = logNuONu0 * ccData[4];
                AmpArr[itcnt] = ccData[jtab];
                itcnt++;
              }
              ccData += lcomp;
            }
            mf_dot3(itcnt, VecL, VecM, VecN, (float)u, (float)v, (float)w, FazArr);
            orc_sincos_vec(itcnt, FazArr, SinArr, CosArr);
            orc_exp_vec(itcnt, ExpArg2, ExpVal2);
            for (int i = 0; i < itcnt; i++) AmpArr[i] = AmpArr[i] * ExpVal2[i];
            sumReal += mf_dot(itcnt, AmpArr, CosArr);
            sumImag += mf_dot(itcnt, AmpArr, SinArr);
          }
          break; }
        case ORC_GaussMod:
          for (it = 0; it < mcomp; it += FAZ_MF) {
            itcnt = 0;
            lim = (mcomp < it + FAZ_MF) ? mcomp : it + FAZ_MF;
            for (iComp = it; iComp < lim; iComp++) {
              if (ccData[0] != 0.0f) {
                amp = ccData[0];
                arg = (float)(ccData[4] * u * u + ccData[5] * v * v + ccData[6] * u * v);
                ExpArg[itcnt] = arg;
                AmpArr[itcnt] = amp;
                VecL[itcnt] = ccData[1]; VecM[itcnt] = ccData[2]; VecN[itcnt] = ccData[3];
                itcnt++;
              }
              ccData += lcomp;
            }
            mf_dot3(itcnt, VecL, VecM, VecN, (float)u, (float)v, (float)w, FazArr);
            orc_sincos_vec(itcnt, FazArr, SinArr, CosArr);
            orc_exp_vec(itcnt, ExpArg, ExpVal);
            for (int i = 0; i < itcnt; i++) AmpArr[i] = AmpArr[i] * ExpVal[i];
            sumReal += mf_dot(itcnt, AmpArr, CosArr);
            sumImag += mf_dot(itcnt, AmpArr, SinArr);
          }
          break;
        case ORC_GaussModTSpec: {
          int jtab = 8 + itab;
          for (it = 0; it < mcomp; it += FAZ_MF) {
            itcnt = 0;
            lim = (mcomp < it + FAZ_MF) ? mcomp : it + FAZ_MF;
            for (iComp = it; iComp < lim; iComp++) {
              if (ccData[jtab] != 0.0f) {
                arg = (float)(ccData[4] * u * u + ccData[5] * v * v + ccData[6] * u * v);
                amp = ccData[jtab];
                ExpArg[itcnt] = arg;
                AmpArr[itcnt] = amp;
                ExpArg2[itcnt] = logNuONu0 * ccData[7];
                VecL[itcnt] = ccData[1]; VecM[itcnt] = ccData[2]; VecN[itcnt] = ccData[3];
                itcnt++;
              }
              ccData += lcomp;
            }
            mf_dot3(itcnt, VecL, VecM, VecN, (float)u, (float)v, (float)w, FazArr);
            orc_sincos_vec(itcnt, FazArr, SinArr, CosArr);
            orc_exp_vec(itcnt, ExpArg, ExpVal);
            for (int i = 0; i < itcnt; i++) AmpArr[i] = AmpArr[i] * ExpVal[i];
            orc_exp_vec(itcnt, ExpArg2, ExpVal2);
            for (int i = 0; i < itcnt; i++) AmpArr[i] = AmpArr[i] * ExpVal2[i];
            sumReal += mf_dot(itcnt, AmpArr, CosArr);
            sumImag += mf_dot(itcnt, AmpArr, SinArr);
          }
          break; }
        default:
          return;
        }
        apply_model(d, visData, offsetChannel, sumReal, sumImag);
      }
    }
  }
}

void orc_dft_obit(const orc_dft_desc *d, float *buffer, int row0, int row1)
{
  if (d->isMF) dft_rows_mf(d, buffer, row0, row1);
  else         dft_rows_base(d, buffer, row0, row1);
}

typedef struct { const orc_dft_desc *d; float *buf; int lo, hi; } dft_job;
static void *dft_job_run(void *a)
{
  dft_job *j = (dft_job *)a;
  orc_dft_obit(j->d, j->buf, j->lo, j->hi);
  return NULL;
}

/* ObitSkyModelFTDFT (src/ObitSkyModel.c:2287-2331) and the MF driver
 * (src/ObitSkyModelMF.c:1711-1752): contiguous row ranges, the last thread
 * takes the remainder. */
void orc_dft_obit_threaded(const orc_dft_desc *d, float *buffer, int nthreads)
{
  int nvis = d->nrow;
  int thresh = d->isMF ? 100 : 1000;
  int nThreads = (nvis < thresh) ? 1 : (nthreads < 1 ? 1 : nthreads);
  int per = nvis / nThreads; if (per < 1) per = 1;
  pthread_t th[256];
  dft_job jobs[256];
  if (nThreads > 256) nThreads = 256;
  int lo = 0, hi = per < nvis ? per : nvis;
  for (int i = 0; i < nThreads; i++) {
    if (i == nThreads - 1) hi = nvis;
    jobs[i].d = d; jobs[i].buf = buffer; jobs[i].lo = lo; jobs[i].hi = hi;
    lo += per; hi += per; if (hi > nvis) hi = nvis;
  }
  if (nThreads == 1) { dft_job_run(&jobs[0]); return; }
  for (int i = 0; i < nThreads; i++) pthread_create(&th[i], NULL, dft_job_run, &jobs[i]);
  for (int i = 0; i < nThreads; i++) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------ */
/* FP64 model oracle                                                   */
/* ------------------------------------------------------------------ */

/* Closed form V(u,v,w) = sum_k A_k(nu) G_k(u,v) exp(i (x_k u + y_k v + z_k w))
 * with u,v,w = uvw_f32 * nu/nu_ref in FP64 and the Obit table conventions
 * (src/ObitSkyModel.c:1939-2043 table layout; :2528-2589 amplitude terms;
 * src/ObitSkyModelMF.c:2143-2260 tabulated spectra).  The positive
 * spectral sign is the CPU's.  Pinned by: analytic special cases in
 * tests/test_oracle_dft.py (origin component, conjugate symmetry,
 * linearity, Gaussian FWHM->0 limit) and brute-force agreement with the
 * FP32 restatement within its own error bound. */
typedef struct { const orc_dft_desc *d; const float *buf; const double *fr;
                 double *out; int lo, hi; } m64_job;

static void model_fp64_rows(const orc_dft_desc *d, const float *buffer,
                            const double *freqRatio, double *out, int lo, int hi)
{
  const int lcomp = d->lcomp, nterm = d->nSpecTerm;
  const int mcomp = count_mcomp(d);
  const int nch = d->nchan, nif = d->nif;
  for (int r = lo; r < hi; r++) {
    const float *vis = buffer + (size_t)r * d->lrec;
    double *o = out + (size_t)r * nif * nch * 2;
    memset(o, 0, sizeof(double) * 2 * nif * nch);
    for (int iIF = d->startIF; iIF < d->startIF + d->numIF; iIF++) {
      for (int ic = d->startChan; ic < d->startChan + d->numChan; ic++) {
        int ifq = iIF * d->kincif + ic * d->kincf;
        double fr = freqRatio[ifq];
        double u = (double)vis[d->ilocu] * fr;
        double v = (double)vis[d->ilocv] * fr;
        double w = (double)vis[d->ilocw] * fr;
        double L = 0.0, Ltab = 0.0;
        int itab = 0;
        if (d->modType == ORC_PointModSpec || d->modType == ORC_GaussModSpec)
          L = log(d->freqArr[ifq] / d->smRefFreq);
        if (d->modType == ORC_PointModTSpec || d->modType == ORC_GaussModTSpec) {
          itab = d->specIndex[ifq];
          Ltab = log(d->freqArr[ifq] / d->specFreq[itab]);
        }
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < mcomp; k++) {
          const float *cc = d->comps + (size_t)k * lcomp;
          double amp = cc[0];
          double G = 0.0;
          switch (d->modType) {
          case ORC_PointMod: break;
          case ORC_PointModSpec: {
            double a = 0.0, Lp = L;
            for (int t = 0; t < nterm; t++) { a += (double)cc[4 + t] * Lp; Lp *= L; }
            amp *= exp(a); break; }
          case ORC_GaussMod:
            G = (double)cc[4] * u * u + (double)cc[5] * v * v + (double)cc[6] * u * v; break;
          case ORC_GaussModSpec: {
            double a = 0.0, Lp = L;
            for (int t = 0; t < nterm; t++) { a += (double)cc[7 + t] * Lp; Lp *= L; }
            amp *= exp(a);
            G = (double)cc[4] * u * u + (double)cc[5] * v * v + (double)cc[6] * u * v; break; }
          case ORC_PointModTSpec:
            amp = (double)cc[5 + itab] * exp((double)cc[4] * Ltab); break;
          case ORC_GaussModTSpec:
            amp = (double)cc[8 + itab] * exp((double)cc[7] * Ltab);
            G = (double)cc[4] * u * u + (double)cc[5] * v * v + (double)cc[6] * u * v; break;
          default: break;
          }
          if (amp == 0.0) continue;
          if (G != 0.0) amp *= exp(G);
          double ph = (double)cc[1] * u + (double)cc[2] * v + (double)cc[3] * w;
          sr += amp * cos(ph);
          si += amp * sin(ph);
        }
        double *oo = o + ((size_t)iIF * nch + ic) * 2;
        if (d->doFlip) { oo[0] = -si; oo[1] = sr; }
        else           { oo[0] = sr;  oo[1] = si; }
      }
    }
  }
}
static void *m64_run(void *a)
{
  m64_job *j = (m64_job *)a;
  model_fp64_rows(j->d, j->buf, j->fr, j->out, j->lo, j->hi);
  return NULL;
}
void orc_dft_model_fp64(const orc_dft_desc *d, const float *buffer,
                        const double *freqRatio, double *out,
                        int row0, int row1, int nthreads)
{
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  int n = row1 - row0;
  if (n <= 0) return;
  if (nthreads > n) nthreads = n;
  pthread_t th[64]; m64_job jb[64];
  int per = n / nthreads, lo = row0;
  for (int i = 0; i < nthreads; i++) {
    jb[i].d = d; jb[i].buf = buffer; jb[i].fr = freqRatio; jb[i].out = out;
    jb[i].lo = lo; jb[i].hi = (i == nthreads - 1) ? row1 : lo + per;
    lo += per;
  }
  if (nthreads == 1) { m64_run(&jb[0]); return; }
  for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, m64_run, &jb[i]);
  for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------ */
/* Gridding convolution function                                       */
/* ------------------------------------------------------------------ */

/* Schwab's rational approximations of the 0-order prolate spheroidal
 * wave functions as tabulated by AIPS and Obit (src/ObitUVGrid.c:1440-
 * 1597).  Only support widths 6 (two branches) are needed by ConvFunc(5)
 * since it passes im=6; the table rows are indexed by alpha = 0..2 in
 * steps of 1/2 (ialf 1..5).  Pinned by: psi(0)/psi(0.5) values, symmetry,
 * zero at eta=1, continuity at the branch point, tests/test_oracle_grid.py. */
static const float P6L[5][5] = {
  {5.613913e-2f, -3.019847e-1f, 6.256387e-1f, -6.324887e-1f, 3.303194e-1f},
  {6.843713e-2f, -3.342119e-1f, 6.302307e-1f, -5.829747e-1f, 2.765700e-1f},
  {8.203343e-2f, -3.644705e-1f, 6.278660e-1f, -5.335581e-1f, 2.312756e-1f},
  {9.675562e-2f, -3.922489e-1f, 6.197133e-1f, -4.857470e-1f, 1.934013e-1f},
  {1.124069e-1f, -4.172349e-1f, 6.069622e-1f, -4.405326e-1f, 1.618978e-1f}};
static const float Q6L[5][2] = {
  {9.077644e-1f, 2.535284e-1f}, {8.626056e-1f, 2.291400e-1f},
  {8.212018e-1f, 2.078043e-1f}, {7.831755e-1f, 1.890848e-1f},
  {7.481828e-1f, 1.726085e-1f}};
static const float P6U[5][5] = {
  {8.531865e-4f, -1.616105e-2f, 6.888533e-2f, -1.109391e-1f, 7.747182e-2f},
  {2.060760e-3f, -2.558954e-2f, 8.595213e-2f, -1.170228e-1f, 7.094106e-2f},
  {4.028559e-3f, -3.697768e-2f, 1.021332e-1f, -1.201436e-1f, 6.412774e-2f},
  {6.887946e-3f, -4.994202e-2f, 1.168451e-1f, -1.207733e-1f, 5.744210e-2f},
  {1.071895e-2f, -6.404749e-2f, 1.297386e-1f, -1.194208e-1f, 5.112822e-2f}};
static const float Q6U[5][2] = {
  {1.101270e+0f, 3.858544e-1f}, {1.025431e+0f, 3.337648e-1f},
  {9.599102e-1f, 2.918724e-1f}, {9.025276e-1f, 2.575336e-1f},
  {8.517470e-1f, 2.289667e-1f}};
static const float ALPHA[5] = {0.0f, 0.5f, 1.0f, 1.5f, 2.0f};

float orc_sphfn(int ialf, int im, int iflag, float eta)
{
  float psi = 0.0f, eta2, x;
  if (ialf < 1 || ialf > 5) return -999.9f;
  if (im != 6) return -999.9f;       /* only the support used by ConvFunc(5) */
  if (fabs(eta) > 1.0) return -999.9f;
  eta2 = eta * eta;
  int j = ialf - 1;
  /* float numerator, double denominator (the literal 1.0 is double) */
  if (fabs(eta) <= 0.75) {
    x = eta2 - 0.5625f;
    psi = (float)((P6L[j][0] + x * (P6L[j][1] + x * (P6L[j][2] + x * (P6L[j][3] + x * P6L[j][4])))) /
                  (1.0 + x * (Q6L[j][0] + x * Q6L[j][1])));
  } else {
    x = eta2 - 1.0f;
    psi = (float)((P6U[j][0] + x * (P6U[j][1] + x * (P6U[j][2] + x * (P6U[j][3] + x * P6U[j][4])))) /
                  (1.0 + x * (Q6U[j][0] + x * Q6U[j][1])));
  }
  if ((iflag > 0) || (ialf == 1) || (eta == 0.0f)) return psi;
  /* the reference tests abs(eta) with the integer abs: only eta == 1 hits */
  if ((int)eta == 1 || (int)eta == -1) psi = 0.0f;
  else psi = (float)(pow((1.0 - eta2), ALPHA[ialf - 1]) * psi);
  return psi;
}

/* ConvFunc(in,5) (src/ObitUVGrid.c:1352-1415): convWidth 7, 100 points per
 * cell, parm[0] = 7/2 = 3 (integer division), so nmax=300, bias=350,
 * ialf=3, im=6.  convgfn[j], j = i/100 + (i%100)*7, plus a shifted copy
 * of the first row at 700..705 and a zero at 706. */
void orc_convfunc_sph(float *convfn, float *convgfn)
{
  const int convWidth = 7, convNperCell = 100;
  const int size = convWidth * convNperCell + 1 + convWidth;  /* 708 */
  float parm0 = (float)(convWidth / 2), parm1 = 1.0f;
  memset(convfn, 0, sizeof(float) * size);
  memset(convgfn, 0, sizeof(float) * size);
  int nmax = (int)(parm0 * convNperCell + 0.1);
  int bias = (convNperCell / 2) * convWidth;
  int ialf = (int)(2.0 * parm1 + 1.1);
  int im = (int)(2.0 * parm0 + 0.1);
  if (im < 4) im = 4;
  if (im > 8) im = 8;
  for (int i = 0; i < nmax; i++) {
    float eta = (float)i / (float)(nmax - 1);
    convfn[bias + i] = orc_sphfn(ialf, im, 0, eta);
  }
  for (int i = 1; i <= bias - 1; i++) convfn[bias - i] = convfn[bias + i];
  int lim = convWidth * convNperCell + 1, j = 0;
  for (int i = 0; i < lim; i++) {
    j = (i / convNperCell) + (i % convNperCell) * convWidth;
    convgfn[j] = convfn[i];
  }
  for (int i = 1; i < convWidth; i++) {
    j = convNperCell * convWidth + i - 1;
    convgfn[j] = convgfn[i];
  }
  convgfn[j + 1] = 0.0f;
}

/* ------------------------------------------------------------------ */
/* Gridding restatement                                                */
/* ------------------------------------------------------------------ */

/* _lroundf (src/ObitThreadGrid.c:132-138): argument arrives as ofloat,
 * the +/-0.5 is a double operation, then truncation. */
static int lroundf_obit(float in)
{
  if (in >= 0.0f) return (int)(in + 0.5);
  return (int)(in - 0.5);
}

/* fast_rot + fast_prep_grid for one row (src/ObitThreadGrid.c:2045-2177,
 * 1263-1387).  The reference reads the rotated visibility from slot
 * (ichan-bChan) and writes the weighted one to slot ichan of the same
 * work array; the same aliasing is kept by using one array. */
void orc_grid_prep_row(const orc_grid_desc *g, const float *buffer, int row,
                       int *iuarr, int *ivarr, int *itu, int *itv, float *visout)
{
  const int halfWidth = g->convWidth / 2;
  const int convNperCell = g->convNperCell;
  int bChan = g->bChan, eChan = g->eChan;
  if (eChan < bChan + 1) eChan = bChan + 1;
  const int nwork = 2 * (g->nchanIF + 16);
  float *fwork1 = (float *)calloc((size_t)nwork, sizeof(float));
  float *fwork2 = (float *)calloc((size_t)nwork, sizeof(float));
  float *fwork3 = (float *)calloc((size_t)nwork, sizeof(float));
  const float *vis_in = buffer;
  int ivis = row * g->lrec;
  const float *rot = g->rot;
  int halfv = g->ny / 2;
  float maxBL2 = g->maxBL * g->maxBL;
  float minBL2 = g->minBL * g->minBL;
  float bmTaper = g->bmTaper;
  int doTape = (fabs(bmTaper) > 0.0) || ((g->sigma1 != NULL) && (g->sigma1[0] != 0.0f));
  float guardu = (float)fabs(g->guardu / g->uscale);
  float guardv = (float)fabs(g->guardv / g->vscale);
  float u = vis_in[ivis], v = vis_in[ivis + 1], w = vis_in[ivis + 2];
  float uu = u * rot[0] + v * rot[1] + w * rot[2];
  float vv = u * rot[3] + v * rot[4] + w * rot[5];
  float ww = u * rot[6] + v * rot[7] + w * rot[8];
  float phaseSign, freqFact;
  /* ---- fast_rot ---- */
  if (g->isBeam) {
    for (int ichan = bChan; ichan < eChan; ichan++) {
      fwork1[(ichan - bChan) * 2] = 1.0f;
      fwork1[(ichan - bChan) * 2 + 1] = 0.0f;
    }
  } else {
    const double twopi = 2 * 3.14159265358979323846;
    const double itwopi = 1.0 / (2 * 3.14159265358979323846);
    double dshift[3] = {(double)g->shift[0], (double)g->shift[1], (double)g->shift[2]};
    phaseSign = (uu <= 0.0f) ? -1.0f : 1.0f;
    int num = 0;
    for (int ichan = bChan; ichan < eChan; ichan++) {
      freqFact = phaseSign * g->freqArr[ichan];
      float uf = uu * freqFact, vf = vv * freqFact, wf = ww * freqFact;
      double dphase = (uf * dshift[0] + vf * dshift[1] + wf * dshift[2]);
      int iphase = (int)(dphase * itwopi);
      fwork1[ichan - bChan] = (float)(dphase - iphase * twopi);
      num++;
    }
    orc_sincos_vec(num, fwork1, fwork2, fwork3);
    for (int ichan = bChan; ichan < eChan; ichan++) {
      int jvis = ivis + g->nrparm + ichan * 3;
      float vvr = vis_in[jvis];
      float vvi = vis_in[jvis + 1] * phaseSign;
      float vr = vvr * fwork3[ichan - bChan] - vvi * fwork2[ichan - bChan];
      float vi = vvr * fwork2[ichan - bChan] + vvi * fwork3[ichan - bChan];
      fwork1[(ichan - bChan) * 2] = vr;
      fwork1[(ichan - bChan) * 2 + 1] = vi;
    }
  }
  /* ---- fast_prep_grid ---- */
  phaseSign = (uu <= 0.0f) ? -1.0f : 1.0f;
  for (int ichan = bChan; ichan < eChan; ichan++) {
    freqFact = phaseSign * g->freqArr[ichan];
    float uf = uu * freqFact, vf = vv * freqFact;
    int jvis = ivis + g->nrparm + ichan * 3;
    float vr = fwork1[(ichan - bChan) * 2];
    float vi = fwork1[(ichan - bChan) * 2 + 1];
    float vw = vis_in[jvis + 2];
    int want = ((vw > 0.0f) && (uf < guardu) && (fabs(vf) < guardv));
    float BL2 = uf * uf + vf * vf;
    if ((maxBL2 > 0) && want) want = want && maxBL2 > BL2;
    if ((minBL2 > 0) && want) want = want && minBL2 < BL2;
    if (want) {
      if (doTape) {
        if ((g->sigma1 != NULL) && (g->sigma1[ichan] != 0.0f)) {
          float ftemp = uf * uf * (g->sigma2[ichan] + bmTaper) +
                        vf * vf * (g->sigma1[ichan] + bmTaper) +
                        uf * vf * (g->sigma3[ichan]);
          vw *= expf(ftemp);
        } else if (fabs(bmTaper) > 0.0) {
          vw *= expf(bmTaper * BL2);
        }
      }
      fwork1[2 * ichan] = vr * vw;
      fwork1[2 * ichan + 1] = vi * vw;
      uf *= g->uscale;
      vf *= g->vscale;
      int iu = (int)(uf + 0.5);
      int iv = lroundf_obit(vf);
      itu[ichan] = convNperCell + lroundf_obit((float)(convNperCell * (iu - uf - 0.5)));
      itv[ichan] = convNperCell + lroundf_obit((float)(convNperCell * (iv - vf - 0.5)));
      iuarr[ichan] = iu;
      ivarr[ichan] = iv - halfWidth + halfv;
    } else {
      fwork1[2 * ichan] = 0.0f;
      fwork1[2 * ichan + 1] = 0.0f;
      iuarr[ichan] = halfWidth;
      ivarr[ichan] = halfv;
      itu[ichan] = 0;
      itv[ichan] = 0;
    }
  }
  for (int ichan = bChan; ichan < eChan; ichan++) {
    visout[2 * ichan] = fwork1[2 * ichan];
    visout[2 * ichan + 1] = fwork1[2 * ichan + 1];
  }
  free(fwork1); free(fwork2); free(fwork3);
}

void orc_grid_prep(const orc_grid_desc *g, const float *buffer, int row0, int row1,
                   int *iu, int *iv, int *itu, int *itv, float *vis)
{
  const size_t n = (size_t)g->nchanIF;
  for (int r = row0; r < row1; r++) {
    size_t o = (size_t)(r - row0) * n;
    orc_grid_prep_row(g, buffer, r, iu + o, iv + o, itu + o, itv + o, vis + 2 * o);
  }
}

/* fast_grid, C path (src/ObitThreadGrid.c:2002-2027): skip (0,0) vis and
 * rows with zero v weight; cfn = cu*cv; grid += vis*cfn (no FMA). */
static void fast_grid_c(float *grid, const float *vis, int iu, int iv, int lrow,
                        int nconv, const float *cu, const float *cv)
{
  if (nconv <= 0) return;
  if ((vis[0] == 0.0f) && (vis[1] == 0.0f)) return;
  for (int jv = 0; jv < nconv; jv++) {
    float cvv = cv[jv];
    if (cvv != 0.0f) {
      int addr = (iv + jv) * lrow + iu * 2;
      for (int ju = 0; ju < nconv; ju++) {
        float cfn = cu[ju] * cvv;
        grid[addr] += vis[0] * cfn; addr++;
        grid[addr] += vis[1] * cfn; addr++;
      }
    }
  }
}

/* ThreadGrid (src/ObitThreadGrid.c:960-991). */
void orc_grid_obit(const orc_grid_desc *g, const float *buffer, int row0, int row1,
                   float *grid)
{
  const int hw = g->convWidth / 2, fw = g->convWidth;
  const int lrow = 2 * (1 + g->nx / 2 + hw);
  int bChan = g->bChan, eChan = g->eChan;
  if (eChan < bChan + 1) eChan = bChan + 1;
  const int n = g->nchanIF;
  int *iu = malloc(sizeof(int) * n), *iv = malloc(sizeof(int) * n);
  int *itu = malloc(sizeof(int) * n), *itv = malloc(sizeof(int) * n);
  float *vis = malloc(sizeof(float) * 2 * n);
  for (int r = row0; r < row1; r++) {
    orc_grid_prep_row(g, buffer, r, iu, iv, itu, itv, vis);
    for (int ic = bChan; ic < eChan; ic++)
      fast_grid_c(grid, &vis[2 * ic], iu[ic], iv[ic], lrow, fw,
                  g->convgfn + fw * itu[ic], g->convgfn + fw * itv[ic]);
  }
  free(iu); free(iv); free(itu); free(itv); free(vis);
}

/* FP64 values on the CPU's sample indices and the same tabulated GCF:
 * products and sums in double.  Pinned by: tests compare its sum of
 * weights against the analytic sum of GCF taps and its indices against
 * orc_grid_prep. */
void orc_grid_fp64(const orc_grid_desc *g, const float *buffer, int row0, int row1,
                   double *grid)
{
  const int hw = g->convWidth / 2, fw = g->convWidth;
  const int lrow = 2 * (1 + g->nx / 2 + hw);
  int bChan = g->bChan, eChan = g->eChan;
  if (eChan < bChan + 1) eChan = bChan + 1;
  const int n = g->nchanIF;
  int *iu = malloc(sizeof(int) * n), *iv = malloc(sizeof(int) * n);
  int *itu = malloc(sizeof(int) * n), *itv = malloc(sizeof(int) * n);
  float *vis = malloc(sizeof(float) * 2 * n);
  for (int r = row0; r < row1; r++) {
    orc_grid_prep_row(g, buffer, r, iu, iv, itu, itv, vis);
    for (int ic = bChan; ic < eChan; ic++) {
      double vr = vis[2 * ic], vi = vis[2 * ic + 1];
      if (vr == 0.0 && vi == 0.0) continue;
      const float *cu = g->convgfn + fw * itu[ic];
      const float *cv = g->convgfn + fw * itv[ic];
      for (int jv = 0; jv < fw; jv++) {
        double cvv = cv[jv];
        if (cvv == 0.0) continue;
        size_t addr = (size_t)(iv[ic] + jv) * lrow + iu[ic] * 2;
        for (int ju = 0; ju < fw; ju++) {
          double cfn = (double)cu[ju] * cvv;
          grid[addr++] += vr * cfn;
          grid[addr++] += vi * cfn;
        }
      }
    }
  }
  free(iu); free(iv); free(itu); free(itv); free(vis);
}

/* ThreadFlip (src/ObitThreadGrid.c:998-1045): rows 1..ny/2 paired with
 * ny-row; the hw+1 columns around u=0 are folded into the conjugate row,
 * reading both sides before writing (the in-place order is kept). */
#define FLIP_BODY(T)                                                         \
  const int halfWidth = convWidth / 2;                                       \
  const int lrow = 2 * (1 + nx / 2 + halfWidth);                             \
  const int ny2 = ny / 2;                                                    \
  for (int vrow = 1; vrow <= ny2; vrow++) {                                  \
    int vc = ny - vrow;                                                      \
    T *gci = grid + (size_t)vc * lrow + halfWidth * 2;                       \
    T *gxo = grid + (size_t)vrow * lrow + halfWidth * 2;                     \
    T *gxi = gxo, *gco = gci;                                                \
    for (int iu = 0; iu <= halfWidth; iu++) {                                \
      T cji0 = gci[0], cji1 = gci[1], cjo0 = gco[0], cjo1 = gco[1];          \
      T xxi0 = gxi[0], xxi1 = gxi[1], xxo0 = gxo[0], xxo1 = gxo[1];          \
      gxo[0] = xxo0 + cji0; gxo[1] = xxo1 - cji1;                            \
      gco[0] = cjo0 + xxi0; gco[1] = cjo1 - xxi1;                            \
      gxo += 2; gco += 2; gxi -= 2; gci -= 2;                                \
    }                                                                        \
  }
void orc_grid_flip(float *grid, int nx, int ny, int convWidth) { FLIP_BODY(float) }
void orc_grid_flip_fp64(double *grid, int nx, int ny, int convWidth) { FLIP_BODY(double) }

/* ThreadMerge (src/ObitThreadGrid.c:1052-1105): drop the hw leading
 * columns, swap v halves ("centre at edges"), add into out. */
void orc_grid_merge(const float *grid, int nx, int ny, int convWidth, float *out)
{
  const int halfWidth = convWidth / 2;
  const int ilrow = 2 * (1 + nx / 2 + halfWidth);
  const int olrow = 2 * (1 + nx / 2);
  int vs = ny / 2;
  for (int vrow = 0; vrow < ny / 2; vrow++) {
    const float *gi = grid + (size_t)vrow * ilrow + halfWidth * 2;
    float *go = out + (size_t)vs * olrow;
    for (int iu = 0; iu < olrow; iu += 2) { go[0] += gi[0]; go[1] += gi[1]; go += 2; gi += 2; }
    vs++;
  }
  vs = 0;
  for (int vrow = ny / 2; vrow < ny; vrow++) {
    const float *gi = grid + (size_t)vrow * ilrow + halfWidth * 2;
    float *go = out + (size_t)vs * olrow;
    for (int iu = 0; iu < olrow; iu += 2) { go[0] += gi[0]; go[1] += gi[1]; go += 2; gi += 2; }
    vs++;
  }
}

typedef struct { const orc_grid_desc *g; const float *buf; int lo, hi; float *grid; } gjob;
static void *gjob_run(void *a)
{
  gjob *j = (gjob *)a;
  orc_grid_obit(j->g, j->buf, j->lo, j->hi, j->grid);
  orc_grid_flip(j->grid, j->g->nx, j->g->ny, j->g->convWidth);
  return NULL;
}

/* ObitThreadGridSetupBase (src/ObitThreadGrid.c:244-253, 346-347):
 * nGpI private grids for one image, grid j takes rows [j*nvis/nGpI,
 * (j+1)*nvis/nGpI) (the last takes the rest); each is gridded and flipped
 * (ThreadGrid, ThreadFlip), then merged in order j=0..nGpI-1. */
void orc_grid_obit_threaded(const orc_grid_desc *g, const float *buffer,
                            int nGpI, int nthreads, float *out)
{
  if (nGpI < 1) nGpI = 1;
  if (nGpI > 64) nGpI = 64;
  const int hw = g->convWidth / 2;
  const size_t size = (size_t)2 * (1 + hw + g->nx / 2) * g->ny;
  int nvis = g->nrow, per = nvis / nGpI;
  gjob jb[64]; pthread_t th[64];
  for (int j = 0; j < nGpI; j++) {
    jb[j].g = g; jb[j].buf = buffer;
    jb[j].lo = j * per; jb[j].hi = (j == nGpI - 1) ? nvis : (j + 1) * per;
    jb[j].grid = (float *)calloc(size, sizeof(float));
  }
  if (nthreads < 1) nthreads = 1;
  for (int j0 = 0; j0 < nGpI; j0 += nthreads) {
    int nd = (nGpI - j0 < nthreads) ? nGpI - j0 : nthreads;
    if (nd == 1) { gjob_run(&jb[j0]); continue; }
    for (int i = 0; i < nd; i++) pthread_create(&th[i], NULL, gjob_run, &jb[j0 + i]);
    for (int i = 0; i < nd; i++) pthread_join(th[i], NULL);
  }
  memset(out, 0, sizeof(float) * 2 * (size_t)(1 + g->nx / 2) * g->ny);
  for (int j = 0; j < nGpI; j++) {
    orc_grid_merge(jb[j].grid, g->nx, g->ny, g->convWidth, out);
    free(jb[j].grid);
  }
}

/* ------------------------------------------------------------------ */
/* Arithmetic identities the device gridder relies on                  */
/* ------------------------------------------------------------------ */

/* The device forms the CPU's GCF tap fraction (float)(N * ((double)du -
 * 0.5)) (src/ObitThreadGrid.c:1366-1376) as fmaf(N, du, -N/2).  Both are
 * roundings of the same real number: this counts the floats du in
 * (-0.5, 0.5] whose bit patterns lie in [lo, hi) for which they differ. */
typedef struct { uint64_t lo, hi; float N; long bad; } fma_job;
static void *fma_run(void *a)
{
  fma_job *j = (fma_job *)a;
  long bad = 0;
  for (uint64_t u = j->lo; u < j->hi; u++) {
    fbits b; b.u = (uint32_t)u;
    float du = b.f;
    if (!(du > -0.5f && du <= 0.5f)) continue;
    float x = (float)((double)j->N * ((double)du - 0.5));
    float y = fmaf(j->N, du, -0.5f * j->N);
    fbits bx, by; bx.f = x; by.f = y;
    if (bx.u != by.u) bad++;
  }
  j->bad = bad;
  return NULL;
}
long orc_check_tap_fraction(float N, int nthreads)
{
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  pthread_t th[64]; fma_job jb[64];
  const uint64_t total = 0x100000000ULL, per = total / nthreads;
  for (int i = 0; i < nthreads; i++) {
    jb[i].lo = i * per; jb[i].hi = (i == nthreads - 1) ? total : (i + 1) * per;
    jb[i].N = N; jb[i].bad = 0;
    pthread_create(&th[i], NULL, fma_run, &jb[i]);
  }
  long bad = 0;
  for (int i = 0; i < nthreads; i++) { pthread_join(th[i], NULL); bad += jb[i].bad; }
  return bad;
}
