// stokes.hip -- Stokes I records for the gridder, formed on the device.
//
// The CPU forms them while it reads (ObitUVCalSelect, reference
// src/ObitUVCalSelect.c:540-695, set up for Stokes 'I' by
// ObitUVCalSelectIInit, :697-760).  Here the whole resident data set is
// converted in one pass: per (row, IF, channel), from the first two
// correlations (RR, LL or XX, YY):
//   w1 = max(0, wt1) / 2,  w2 = max(0, wt2) / 2,  wt = w1 + w2;
//   V  = (w1 > 0 ? V1 / 2 : 0) + (w2 > 0 ? V2 / 2 : 0), doubled when one
//        of the two weights is not positive;
//   with both hands required (formal I) a sample missing either is zero.
// A single correlation, or data that are already Stokes I, pass through
// the same arithmetic with factors (1, 0).  Arithmetic is FP32 in the
// CPU's order with FP contraction off, so results are bit-identical.
#include "common.h"
#include "ObitHIPVis.h"

#include <algorithm>

namespace obit {
namespace stokes {

struct Desc {
  const float *in;
  float *out;
  long long nvis;
  int lenvis, nrparm, nif, nchan, incs, incf, incif;
  int dual;      // 1: first two correlations at factors 1/2; 0: first one at factor 1
  int bothCorr;  // formal I: both hands required
};

__global__ __launch_bounds__(64) void k_stokes_i(Desc d) {
#pragma clang fp contract(off)
  const int lout = d.nrparm + 3 * d.nif * d.nchan;
  const int ip2 = d.dual ? d.incs : 0;
  const float f1 = d.dual ? 0.5f : 1.0f, f2 = d.dual ? 0.5f : 0.0f;
  const int nk = d.nif * d.nchan;
  for (long long r = blockIdx.x; r < d.nvis; r += gridDim.x) {
    const float *__restrict__ rin = d.in + r * (long long)d.lenvis;
    float *__restrict__ rout = d.out + r * (long long)lout;
    for (int s = threadIdx.x; s < d.nrparm; s += blockDim.x) rout[s] = rin[s];
    // four channels per thread: every load is issued before any store (a
    // channel past the end reads the last one and stores nothing)
    for (int k0 = threadIdx.x; k0 < nk; k0 += 4 * blockDim.x) {
      float a[4][6];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = min(k0 + u * (int)blockDim.x, nk - 1);
        const int iif = k / d.nchan, ch = k - iif * d.nchan;
        const float *p = rin + d.nrparm + iif * d.incif + ch * d.incf;
        a[u][0] = p[0]; a[u][1] = p[1]; a[u][2] = p[2];
        a[u][3] = p[ip2]; a[u][4] = p[ip2 + 1]; a[u][5] = p[ip2 + 2];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + u * (int)blockDim.x;
        // MAX (0.0, w) * 0.5 as the CPU writes it: a NaN weight passes through
        const float w1 = ((0.0f > a[u][2]) ? 0.0f : a[u][2]) * 0.5f;
        const float w2 = ((0.0f > a[u][5]) ? 0.0f : a[u][5]) * 0.5f;
        const bool one = (w1 <= 0.0f) || (w2 <= 0.0f);
        const bool take = !(d.bothCorr && one);
        const float wt = take ? w1 + w2 : 0.0f;
        const bool live = take && (wt > 0.0f);
        const float x1 = (w1 > 0.0f) ? f1 : 0.0f;
        const float x2 = (w2 > 0.0f) ? f2 : 0.0f;
        float vr = x1 * a[u][0] + x2 * a[u][3];
        float vi = x1 * a[u][1] + x2 * a[u][4];
        if (one) {
          vr = 2.0f * vr;
          vi = 2.0f * vi;
        }
        if (k < nk) {
          float *q = rout + d.nrparm + 3 * k;
          q[0] = live ? vr : 0.0f;
          q[1] = live ? vi : 0.0f;
          q[2] = wt;
        }
      }
    }
  }
}

}  // namespace stokes
}  // namespace obit

extern "C" {

int ObitHIPStokesI(const float *d_in, float *d_out, long long nvis, int lenvis, int nrparm,
                   int nif, int nchan, int nstok, int incs, int incf, int incif, int kstoke0,
                   int formalI, void *stream) {
  using namespace obit;
  if (!d_in || !d_out || nvis < 0 || nrparm < 0 || nif < 1 || nchan < 1 || nstok < 1 ||
      incs < 1 || incf < 1 || incif < 1)
    return fail(kInvalid, "ObitHIPStokesI: bad arguments");
  // the correlations a record holds must lie inside it
  const long long last = (long long)nrparm + (long long)(nif - 1) * incif +
                         (long long)(nchan - 1) * incf + (long long)(nstok - 1) * incs + 2;
  if (last >= lenvis) return fail(kInvalid, "ObitHIPStokesI: increments exceed lenvis");
  // ObitUVCalSelectIInit's availability test
  bool missing = (kstoke0 > 1) || (kstoke0 < -2);
  if (kstoke0 <= -5) missing = kstoke0 < -6;
  if (formalI && kstoke0 < 0)
    missing = missing || nstok < 2 || (kstoke0 != -1 && kstoke0 != -5);
  if (missing)
    return fail(kInvalid, "ObitHIPStokesI: Stokes %sI not available (first Stokes %d, %d correlations)",
                formalI ? "formal " : "", kstoke0, nstok);
  if (nvis == 0) return kOK;
  stokes::Desc d;
  d.in = d_in;
  d.out = d_out;
  d.nvis = nvis;
  d.lenvis = lenvis;
  d.nrparm = nrparm;
  d.nif = nif;
  d.nchan = nchan;
  d.incs = incs;
  d.incf = incf;
  d.incif = incif;
  // ObitUVCalSelectIInit's branches: true Stokes or one correlation take
  // the first at factor 1; otherwise the first two at factors 1/2
  d.dual = (kstoke0 <= 0 && nstok != 1) ? 1 : 0;
  d.bothCorr = d.dual ? (formalI ? 1 : 0) : 0;
  int dev = 0, ncu = 0;
  OBIT_HIP_TRY(hipGetDevice(&dev));
  OBIT_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // one wave per row, 32 waves per CU
  const long long nb = std::min<long long>(nvis, (long long)std::max(ncu, 1) * 32);
  hipLaunchKernelGGL(stokes::k_stokes_i, dim3((unsigned)nb), dim3(64), 0, (hipStream_t)stream, d);
  OBIT_HIP_TRY(hipGetLastError());
  return kOK;
}

}  // extern "C"
