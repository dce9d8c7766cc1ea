/*
 * host_grid.c -- the gridding host class: fills CUDAGridInfo from a uv-data
 * description and per-facet gridding parameters and drives the
 * ObitCUDAGrid* entry points, as Obit's ObitGPUGrid.c does (reference
 * src/ObitGPUGrid.c:255-323 Create, 378-571 SetGPUStruct, 580-609
 * 2GPU/Grid, 616-633 Flip, 643-675 2Host, 683-721 Shutdown).
 *
 * The spheroidal GCF table comes from ObitHostConvFuncSph, this library's
 * statement of ConvFunc(in,5) (src/ObitUVGrid.c:1352-1415): F. Schwab's
 * rational approximation of the prolate spheroidal function as AIPS
 * tabulates it (support 6 cells, alpha = 1, 100 points per cell, embedded
 * in a 7-cell table).
 */
#include "host.h"
#include "ObitCUDAGrid.h"
#include "ObitCUDAUtil.h"
#include "ObitHIPVis.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DG2RAD (3.14159265358979323846 / 180.0)

/* Schwab's coefficients for support 6, alpha = 1 (row ialf=3 of the AIPS
 * tables for im=6; src/ObitUVGrid.c:1471-1490). */
static const float SPH_P6L[5] = {8.203343e-2f, -3.644705e-1f, 6.278660e-1f, -5.335581e-1f, 2.312756e-1f};
static const float SPH_Q6L[2] = {8.212018e-1f, 2.078043e-1f};
static const float SPH_P6U[5] = {4.028559e-3f, -3.697768e-2f, 1.021332e-1f, -1.201436e-1f, 6.412774e-2f};
static const float SPH_Q6U[2] = {9.599102e-1f, 2.918724e-1f};

/* Gridding form (iflag <= 0) for alpha = 1: psi(eta) * (1 - eta^2); the
 * numerator polynomial is FP32, the denominator is evaluated in double
 * because its leading constant is a double literal in the reference. */
static float sph_alpha1_support6(float eta)
{
  float psi, eta2 = eta * eta, x;
  if (fabs(eta) <= 0.75) {
    x = eta2 - 0.5625f;
    psi = (float)((SPH_P6L[0] + x * (SPH_P6L[1] + x * (SPH_P6L[2] + x * (SPH_P6L[3] + x * SPH_P6L[4])))) /
                  (1.0 + x * (SPH_Q6L[0] + x * SPH_Q6L[1])));
  } else {
    x = eta2 - 1.0f;
    psi = (float)((SPH_P6U[0] + x * (SPH_P6U[1] + x * (SPH_P6U[2] + x * (SPH_P6U[3] + x * SPH_P6U[4])))) /
                  (1.0 + x * (SPH_Q6U[0] + x * SPH_Q6U[1])));
  }
  if (eta == 0.0f) return psi;
  if (eta == 1.0f || eta == -1.0f) return 0.0f;
  return (float)(pow(1.0 - eta2, 1.0) * psi);
}

void ObitHostConvFuncSph(float *convfn)
{
  const int W = 7, N = 100, bias = (N / 2) * W, nmax = 300;
  int i;
  memset(convfn, 0, sizeof(float) * (W * N + 1 + W));
  for (i = 0; i < nmax; i++) {
    float eta = (float)i / (float)(nmax - 1);
    convfn[bias + i] = sph_alpha1_support6(eta);
  }
  for (i = 1; i <= bias - 1; i++) convfn[bias - i] = convfn[bias + i];
}

/* GridCorrFn (src/ObitUVGrid.c:1612-1662): the gridding correction for one
 * image axis of n pixels centred on 1-relative pixel icent, the inverse of
 * the cosine transform of the tabulated convolving function, normalised
 * to 1 at pixel n/2.  The transform is evaluated as the reference does: an
 * FP32 complex phase ramp applied to every table entry once per pixel. */
void ObitHostGridCorrFn(int n, int icent, const float *convfn, int convWidth, int convNperCell,
                        float *out)
{
  const int size = convWidth * convNperCell + 1;
  const int bias = convWidth * (convNperCell / 2) + 1;
  /* the reference forms these in double (double literals) and stores floats */
  const float p4 = (float)(2.0 * 3.14159265358979323846 / (convNperCell * (double)n));
  const float p1 = (float)((double)p4 * (1.0 - (double)icent) * (1.0 - (double)bias));
  const float p2 = (float)((double)p4 * (1.0 - (double)icent));
  const float p3 = (float)((double)p4 * (1.0 - (double)bias));
  float *data = (float *)malloc(sizeof(float) * 4 * (size_t)size);
  float *ramp = data + 2 * size;
  int i, j;
  if (!data) return;
  for (i = 0; i < size; i++) {
    float phase = p1 + (float)i * p2;
    data[2 * i] = (float)((double)convfn[i] * cos((double)phase));
    data[2 * i + 1] = (float)((double)convfn[i] * sin((double)phase));
    phase = p3 + (float)i * p4;
    ramp[2 * i] = (float)cos(phase);
    ramp[2 * i + 1] = (float)sin(phase);
  }
  for (i = 0; i < n; i++) {
    float sumre = 0.0f;
    for (j = 0; j < 2 * size; j += 2) {
      float tr, ti;
      sumre += data[j];
      tr = data[j] * ramp[j] - data[j + 1] * ramp[j + 1];
      ti = data[j] * ramp[j + 1] + data[j + 1] * ramp[j];
      data[j] = tr;
      data[j + 1] = ti;
    }
    out[i] = sumre;
  }
  {
    const float c = (float)(1.0 / (double)out[n / 2]);  /* ObitFArraySMul */
    for (i = 0; i < n; i++) out[i] *= c;
    for (i = 0; i < n; i++) out[i] = 1.0f / out[i];      /* ObitFArraySDiv */
  }
  free(data);
}

void ObitHostPlanesFromBins(int nchanIF, int nbin, const int *bChan, const int *eChan,
                            int cpuParity, int *planeOfChan)
{
  int c, b;
  for (c = 0; c < nchanIF; c++) planeOfChan[c] = nbin + 100;
  for (b = 0; b < nbin; b++) {
    int hi = cpuParity ? eChan[b] - 1 : eChan[b];
    if (cpuParity && hi < bChan[b]) hi = bChan[b];  /* at least one channel */
    for (c = bChan[b]; c <= hi && c < nchanIF; c++) planeOfChan[c] = b;
  }
}

/* ObitGPUGridCreate (src/ObitGPUGrid.c:255-323). */
CUDAGridInfo *ObitHostGridCreate(int nfacet, int nGPU, const int *GPUs, const ObitHostUVDesc *uv,
                                 int nplane, int doBuff, int nVisPIO)
{
  int i, iGPU, mGPU = nfacet < nGPU ? nfacet : nGPU;
  CUDAGridInfo *c;
  if (nfacet <= 0 || nGPU <= 0 || !uv) return NULL;
  c = (CUDAGridInfo *)calloc(1, sizeof(CUDAGridInfo));
  c->nGPU = mGPU;
  c->cuda_device = (long *)calloc(2 * mGPU, sizeof(long));
  for (i = 0; i < mGPU; i++) c->cuda_device[i] = GPUs ? GPUs[i] : i;
  c->FacetGPU = (long *)calloc(nfacet, sizeof(long));
  iGPU = -1;
  for (i = 0; i < nfacet; i++) { iGPU = (iGPU + 1) % mGPU; c->FacetGPU[i] = iGPU; }
  if (nVisPIO <= 0) nVisPIO = 8192;  /* GPU_NVISPIO, include/ObitGPUGrid.h:92 */
  if (uv->nvis > 0 && nVisPIO > uv->nvis) nVisPIO = uv->nvis;
  for (iGPU = 0; iGPU < mGPU; iGPU++) {
    if (ObitCUDACheckGPU((int)c->cuda_device[iGPU]) < 0) { free(c->cuda_device); free(c->FacetGPU); free(c); return NULL; }
    ObitCUDASetGPU((int)c->cuda_device[iGPU]);
    ObitCUDAGridAlloc(c, nfacet, uv->nchan, uv->hasIF ? uv->nif : 1, nVisPIO, uv->lrec, nplane, iGPU,
                      doBuff);
  }
  c->h_gridInfo->GPU_device_no = (int *)ObitCUDAUtilAllocHost((int)((mGPU + 1) * sizeof(int)));
  for (i = 0; i < mGPU; i++) c->h_gridInfo->GPU_device_no[i] = (int)c->cuda_device[i];
  c->gpu_memory = (size_t *)calloc(mGPU, sizeof(size_t));
  return c;
}

/* ObitGPUGridSetGPUStruct (src/ObitGPUGrid.c:378-571). */
int ObitHostGridSetFacet(CUDAGridInfo *c, int ifacet, const ObitHostFacet *f,
                         const ObitHostUVDesc *uv, const int *planeOfChan,
                         const float *sigma1, const float *sigma2, const float *sigma3)
{
  int i, nchanIF;
  GridInfo *gi;
  FacetInfo *fi;
  if (!c || !f || !uv || ifacet < 0 || ifacet >= c->nfacet) return OBIT_HIP_EINVAL;
  gi = c->h_gridInfo;
  nchanIF = uv->nchan * (uv->hasIF ? uv->nif : 1);
  if (ifacet == 0) {
    float convfn[708];
    gi->nfacet = (int)c->nfacet;
    gi->nchan = nchanIF;
    gi->nif = 1;
    gi->nstok = uv->nstok;
    gi->nrparm = uv->nrparm;
    gi->lenvis = uv->lrec;
    gi->convWidth = 7;
    gi->convNperCell = 100;
    gi->guardu = 0.4f;
    gi->guardv = 0.4f;
    for (i = 0; i < nchanIF; i++) gi->h_freqArr[i] = uv->fscale[i];
    ObitHostConvFuncSph(convfn);
    for (i = 0; i < 7 * 100; i++) gi->h_convfn[i] = convfn[i];
    for (i = 0; i < nchanIF; i++) gi->h_freqPlane[i] = planeOfChan ? planeOfChan[i] : 0;
    for (i = 0; i < nchanIF; i++) {
      gi->h_sigma1[i] = sigma1 ? sigma1[i] : 0.0f;
      gi->h_sigma2[i] = sigma2 ? sigma2[i] : 0.0f;
      gi->h_sigma3[i] = sigma3 ? sigma3[i] : 0.0f;
    }
  }
  fi = c->h_facetInfo[ifacet];
  fi->GPU_num = (int)c->FacetGPU[ifacet];
  fi->nx = f->nx;
  fi->ny = f->ny;
  /* grid size with hw extra columns plus slop (ObitGPUGrid.c:520-522) */
  fi->sizeGrid = 2 * (f->nx / 2 + 1 + gi->convWidth / 2) * f->ny + f->nx;
  /* scales (ObitUVGrid.c:275-280) */
  fi->uscale = (float)(f->nx * fabs((float)(DG2RAD * f->cellx)));
  fi->vscale = (float)(-f->ny * fabs((float)(DG2RAD * f->celly)));
  fi->wscale = 1.0f;
  for (i = 0; i < 3; i++) fi->shift[i] = f->doBeam ? 0.0f : f->shift[i];
  for (i = 0; i < 9; i++) fi->rotUV[i] = f->rotUV[i];
  fi->doBeam = f->doBeam;
  fi->maxBL = f->blmax;
  if (f->blmax <= 0.0f) fi->maxBL = 1.0e12f;
  fi->minBL = f->blmin;
  fi->bmTaper = f->bmTaperUV;
  /* last facet(s): allocate device grids */
  if (ifacet >= c->nfacet - c->nGPU) ObitCUDAGridInit(c, c->FacetGPU[ifacet]);
  return OBIT_HIP_OK;
}

/* ObitGPUGrid2GPU + ObitGPUGridGrid over nVisPIO-row pieces
 * (src/ObitGPUGrid.c:580-609; buffer loop ObitUVGrid.c:605-636). */
int ObitHostGridBuffer(CUDAGridInfo *c, const float *buffer, long long nvis)
{
  long long done = 0;
  int iGPU, first = 1;
  GridInfo *gi;
  if (!c || !buffer) return OBIT_HIP_EINVAL;
  gi = c->h_gridInfo;
  if (!gi->h_data_in) return OBIT_HIP_ESTATE;
  while (done < nvis) {
    long long n = nvis - done;
    if (n > c->nVisPIO) n = c->nVisPIO;
    memcpy(gi->h_data_in, buffer + done * gi->lenvis, (size_t)n * gi->lenvis * sizeof(float));
    gi->nvis = (int)n;
    for (iGPU = 0; iGPU < c->nGPU; iGPU++) {
      ObitCUDASetGPU(gi->GPU_device_no[iGPU]);
      ObitCUDAGridGrid(c, iGPU, first);
    }
    first = 0;
    done += n;
  }
  return OBIT_HIP_OK;
}

int ObitHostGridFlip(CUDAGridInfo *c)
{
  int iGPU;
  if (!c) return OBIT_HIP_EINVAL;
  for (iGPU = 0; iGPU < c->nGPU; iGPU++) {
    ObitCUDASetGPU(c->h_gridInfo->GPU_device_no[iGPU]);
    ObitCUDAGridFlip(c, iGPU);
  }
  return OBIT_HIP_OK;
}

/* ObitGPUGrid2Host (src/ObitGPUGrid.c:643-675): copy back, drop the hw
 * guard columns, swap v halves (ObitCArray2DCenter, ObitCArray.c:1549-1581). */
int ObitHostGrid2Host(CUDAGridInfo *c, int ifacet, float *out)
{
  GridInfo *gi;
  FacetInfo *fi;
  int iplane, iy, hw, lorow, lirow, ny;
  if (!c || !out || ifacet < 0 || ifacet >= c->nfacet) return OBIT_HIP_EINVAL;
  gi = c->h_gridInfo;
  fi = c->h_facetInfo[ifacet];
  ObitCUDASetGPU(gi->GPU_device_no[c->FacetGPU[ifacet]]);
  ObitCUDAGrid2CPU(c, ifacet);
  hw = gi->convWidth / 2;
  lorow = fi->nx / 2 + 1;
  lirow = lorow + hw;
  ny = fi->ny;
  for (iplane = 0; iplane < gi->nplane; iplane++) {
    const float *src = fi->h_grid + (size_t)iplane * fi->sizeGrid;
    float *dst = out + (size_t)iplane * ny * lorow * 2;
    for (iy = 0; iy < ny; iy++) {
      int oy = (iy + ny / 2) % ny;  /* centre at the edges */
      memcpy(dst + (size_t)oy * lorow * 2, src + 2 * ((size_t)iy * lirow + hw),
             sizeof(float) * 2 * lorow);
    }
  }
  return OBIT_HIP_OK;
}

/* ObitGPUGridShutdown (src/ObitGPUGrid.c:683-721). */
void ObitHostGridShutdown(CUDAGridInfo *c)
{
  int iGPU;
  if (!c) return;
  for (iGPU = 0; iGPU < c->nGPU; iGPU++) {
    if (c->h_gridInfo && c->h_gridInfo->GPU_device_no) ObitCUDASetGPU(c->h_gridInfo->GPU_device_no[iGPU]);
    ObitCUDAGridShutdown(c, iGPU);
  }
  free(c->gpu_memory);
  free(c->FacetGPU);
  free(c->cuda_device);
  free(c);
}
