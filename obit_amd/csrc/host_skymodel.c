/*
 * host_skymodel.c -- the DFT host class: fills GPUInfo from a uv-data
 * description and a sky model and drives the ObitCUDASkyModel* entry
 * points, as Obit's ObitGPUSkyModel.c does (reference src/ObitGPUSkyModel.c:
 * 207-403 DFTInit, 416-491 DFTSetMod, 500-520 DFTCalc, 530-579 Shutdown).
 *
 * Two deliberate differences from that file, both to follow Obit's CPU DFT:
 *  - stokFact[0] = 1 and stokFact[1] = stokFactor: the reference sets every
 *    entry to 'factor' (ObitGPUSkyModel.c:464-467), which the component
 *    table already carries (ObitSkyModel.c:1965);
 *  - the FP64 channel frequencies are handed to the library
 *    (ObitHIPSkyModelDFTSetFreq) so the phase is formed from them.
 */
#include "host.h"
#include "ObitCUDASkyModel.h"
#include "ObitCUDAUtil.h"
#include "ObitHIPVis.h"

#include <stdlib.h>
#include <string.h>

/* Obit component types (include/ObitSkyModel.h:131-163). */
enum { PointMod = 0, GaussMod = 1, PointModSpec = 10, GaussModSpec = 11,
       PointModTSpec = 20, GaussModTSpec = 21 };

GPUInfo *ObitHostDFTCreate(void)
{
  GPUInfo *g = (GPUInfo *)calloc(1, sizeof(GPUInfo));
  return g;
}

/* ObitGPUSkyModelDFTInit (src/ObitGPUSkyModel.c:207-403). */
int ObitHostDFTInit(GPUInfo *g, const ObitHostUVDesc *uv, const ObitHostSkyModel *sm, int nVisPIO)
{
  int i, kincf, kincif, nchan, nif, memsize;
  if (!g || !uv || !sm) return OBIT_HIP_EINVAL;
  g->cuda_device = sm->cuda_device;
  ObitCUDASetGPU(g->cuda_device);

  if (nVisPIO <= 0) nVisPIO = 10000;
  if (nVisPIO > uv->nvis && uv->nvis > 0) nVisPIO = uv->nvis;
  g->bufferSize = nVisPIO * uv->lrec;
  g->h_data = ObitCUDAUtilAllocHost(g->bufferSize * (int)sizeof(float));
  g->nvis = nVisPIO;
  g->nchan = uv->nchan;
  g->nrparm = uv->nrparm;
  g->lenvis = uv->lrec;
  g->nstream = STREAM_COUNT;

  nchan = uv->nchan;
  nif = uv->hasIF ? uv->nif : 1;
  memsize = nchan * nif * (int)sizeof(float);
  g->d_freq = ObitCUDAUtilAllocGPU(memsize);
  ObitCUDAUtilHost2GPU(g->d_freq, (float *)uv->fscale, memsize, NULL);

  /* increments in the frequency tables (ObitGPUSkyModel.c:265-277) */
  if (uv->hasIF) {
    if (uv->freqBeforeIF) { kincf = 1; kincif = nchan; }
    else { kincif = 1; kincf = nif; }
  } else { kincif = 1; kincf = 1; }

  g->d_visInfo = (GPUVisInfo *)ObitCUDAUtilAllocGPU(sizeof(GPUVisInfo));
  g->h_visInfo = (GPUVisInfo *)ObitCUDAUtilAllocHost(sizeof(GPUVisInfo));
  memset(g->h_visInfo, 0, sizeof(GPUVisInfo));
  GPUVisInfo *vi = g->h_visInfo;
  vi->freqScale = g->d_freq;
  vi->nvis = g->nvis;
  vi->nprod = uv->ncorr;
  vi->nrparm = uv->nrparm;
  vi->lenvis = uv->lrec;
  vi->nchan = nchan;
  vi->ilocu = uv->ilocu; vi->ilocv = uv->ilocv; vi->ilocw = uv->ilocw;
  vi->iloct = uv->iloct; vi->ilocb = uv->ilocb; vi->ilocsu = uv->ilocsu;
  vi->chanb = sm->startChannel - 1;
  vi->chane = sm->startChannel + sm->numberChannel - 2;
  vi->incf = uv->incf / 3;
  vi->nIF = nif;
  vi->IFb = sm->startIF - 1;
  vi->IFe = sm->startIF + sm->numberIF - 2;
  vi->incif = uv->incif / 3;
  vi->nstok = uv->nstok;
  /* Stokes range: the CPU's startPoln/numberPoln (ObitSkyModel.c:3802-3808);
   * the reference host hard-codes 0 and min(2, nstok)-1 */
  vi->stokb = sm->startPoln - 1;
  vi->stoke = sm->startPoln + sm->numberPoln - 2;
  vi->incs = uv->incs / 3;
  vi->kincif = kincif;
  vi->kincf = kincf;
  if (uv->crvalStokes > 0.0f) vi->stype = 0;
  else if (uv->crvalStokes < -4.0f) vi->stype = 2;
  else vi->stype = 1;
  /* ratio of uv reference frequency to the spectrum reference(s) */
  if (sm->isMF && sm->nSpec > 0) {
    vi->freqRat = ObitCUDAUtilAllocHost(sm->nSpec * (int)sizeof(float));
    for (i = 0; i < sm->nSpec; i++) vi->freqRat[i] = (float)(uv->freq / sm->specFreq[i]);
    vi->d_freqRat = ObitCUDAUtilAllocGPU(sm->nSpec * (int)sizeof(float));
    ObitCUDAUtilHost2GPU(vi->d_freqRat, vi->freqRat, sm->nSpec * (int)sizeof(float), NULL);
  } else {
    vi->freqRat = ObitCUDAUtilAllocHost(sizeof(float));
    vi->freqRat[0] = 1.0f;
    if (sm->imgRefFreq > 0.0) vi->freqRat[0] = (float)(uv->freq / sm->imgRefFreq);
    vi->d_freqRat = ObitCUDAUtilAllocGPU(sizeof(float));
    ObitCUDAUtilHost2GPU(vi->d_freqRat, vi->freqRat, sizeof(float), NULL);
  }
  ObitCUDAUtilHost2GPU((float *)g->d_visInfo, (float *)vi, sizeof(GPUVisInfo), NULL);

  g->d_modelInfo = (GPUModelInfo *)ObitCUDAUtilAllocGPU(sizeof(GPUModelInfo));
  g->h_modelInfo = (GPUModelInfo *)ObitCUDAUtilAllocHost(sizeof(GPUModelInfo));
  memset(g->h_modelInfo, 0, sizeof(GPUModelInfo));
  GPUModelInfo *mi = g->h_modelInfo;
  mi->nterm = sm->nSpecTerm;
  mi->type = GPUModTypePoint;
  if (sm->isMF && sm->specIndex) {
    int n = nchan * nif;
    mi->specIndex = (int *)ObitCUDAUtilAllocHost(n * (int)sizeof(int));
    for (i = 0; i < n; i++) mi->specIndex[i] = sm->specIndex[i];
    mi->d_specIndex = (int *)ObitCUDAUtilAllocGPU(n * (int)sizeof(int));
    ObitCUDAUtilHostAny2GPU(mi->d_specIndex, mi->specIndex, n * (int)sizeof(int), NULL);
  }

  /* per-stream device buffers and streams (ObitGPUSkyModel.c:385-399) */
  g->d_data = (float **)calloc(g->nstream, sizeof(float *));
  memsize = ((g->nvis + 3) * g->lenvis) * (int)sizeof(float) / g->nstream;
  for (i = 0; i < g->nstream; i++) g->d_data[i] = ObitCUDAUtilAllocGPU(memsize);
  g->stream = (int **)calloc(g->nstream, sizeof(int *));
  g->cycleDone = (int **)calloc(g->nstream, sizeof(int *));
  for (i = 0; i < g->nstream; i++) {
    g->stream[i] = ObitCUDAStreamCreate();
    g->cycleDone[i] = ObitCUDAEventCreate();
    ObitCUDAEventRecord(g->cycleDone[i], g->stream[i]);
  }

  ObitCUDASkyModelDFTInit(g);
  if (sm->useFreq64 && uv->freqArr) {
    int rc = ObitHIPSkyModelDFTSetFreq(g, uv->freqArr, nchan * nif, uv->freq);
    if (rc) return rc;
    if (sm->isMF && sm->nSpec > 0 && sm->specFreq) rc = ObitHIPSkyModelDFTSetSpecRef(g, sm->specFreq, sm->nSpec);
    else if (sm->imgRefFreq > 0.0) rc = ObitHIPSkyModelDFTSetSpecRef(g, &sm->imgRefFreq, 1);
    if (rc) return rc;
  }
  return OBIT_HIP_OK;
}

/* ObitGPUSkyModelDFTSetMod (src/ObitGPUSkyModel.c:416-491). */
int ObitHostDFTSetMod(GPUInfo *g, const ObitHostSkyModel *sm, const float *comps, int lcomp, int ncomp)
{
  int i, high, memsize;
  if (!g || !sm || !comps || lcomp < 4 || ncomp < 0) return OBIT_HIP_EINVAL;
  /* last non-zero component (ObitGPUSkyModel.c:423-425) */
  high = 1;
  for (i = 0; i < ncomp; i++) if (comps[(size_t)i * lcomp] != 0.0f) high = i + 1;
  if (ncomp == 0) high = 0;
  if (g->nmodel[0] != high || g->modelSize != lcomp || !g->d_model[0]) {
    g->nmodel[0] = g->nmodel[1] = g->nmodel[2] = g->nmodel[3] = high;
    g->modelSize = lcomp;
    if (g->d_model[0]) ObitCUDAUtilFreeGPU(g->d_model[0]);
    memsize = (high > 0 ? high : 1) * lcomp * (int)sizeof(float);
    g->d_model[0] = ObitCUDAUtilAllocGPU(memsize);
    g->d_model[1] = g->d_model[2] = g->d_model[3] = g->d_model[0];
  }
  GPUModelInfo *mi = g->h_modelInfo;
  mi->nterm = sm->nSpecTerm;
  for (i = 0; i < 4; i++) mi->model[i] = g->d_model[i];
  switch (sm->modType) {
    case PointMod: mi->type = GPUModTypePoint; break;
    case PointModSpec: mi->type = GPUModTypePointSpec; break;
    case PointModTSpec: mi->type = GPUModTypePointTSpec; break;
    case GaussMod: mi->type = GPUModTypeGauss; break;
    case GaussModSpec: mi->type = GPUModTypeGaussSpec; break;
    case GaussModTSpec: mi->type = GPUModTypeGaussTSpec; break;
    default: return OBIT_HIP_EINVAL;  /* the reference g_error()s */
  }
  mi->opType = GPUOpTypeSub;
  if (sm->doDivide) mi->opType = GPUOpTypeDiv;
  if (sm->doReplace) mi->opType = GPUOpTypeRepl;
  /* CPU semantics: the table carries 'factor'; Stokes after the first are
   * scaled by stokFactor (ObitSkyModel.c:2839-2841) */
  mi->stokFact[0] = 1.0f;
  mi->stokFact[1] = sm->stokFactor;
  mi->stokFact[2] = sm->stokFactor;
  mi->stokFact[3] = sm->stokFactor;
  mi->doSwap = sm->doFlip ? 1 : 0;
  if (high > 0)
    ObitCUDAUtilHost2GPU(g->d_model[0], (float *)comps, high * lcomp * (int)sizeof(float), NULL);
  for (i = 0; i < 4; i++) mi->nmodel[i] = g->nmodel[0];
  mi->size = g->modelSize;
  ObitCUDAUtilHost2GPU((float *)g->d_modelInfo, (float *)mi, sizeof(GPUModelInfo), NULL);
  ObitCUDASkyModelDFTSetMod(g);
  return OBIT_HIP_OK;
}

/* ObitGPUSkyModelDFTCalc (src/ObitGPUSkyModel.c:500-520), applied to a
 * buffer of any length in nVisPIO-row pieces, as ObitSkyModelSubUV reads
 * and writes one uv buffer at a time (src/ObitSkyModel.c:1000-1026). */
int ObitHostDFTCalc(GPUInfo *g, float *buffer, long long nvis)
{
  long long done = 0;
  int per;
  if (!g || !buffer || g->lenvis <= 0) return OBIT_HIP_EINVAL;
  per = g->bufferSize / g->lenvis;
  if (per <= 0) return OBIT_HIP_EINVAL;
  while (done < nvis) {
    int n = (int)((nvis - done) < per ? (nvis - done) : per);
    size_t bytes = (size_t)n * g->lenvis * sizeof(float);
    g->nvis = n;
    g->h_visInfo->nvis = n;
    memcpy(g->h_data, buffer + done * g->lenvis, bytes);
    ObitCUDASkyModelDFTCalc(g);
    memcpy(buffer + done * g->lenvis, g->h_data, bytes);
    done += n;
  }
  return OBIT_HIP_OK;
}

/* ObitGPUSkyModelDFTShutdown (src/ObitGPUSkyModel.c:530-579). */
void ObitHostDFTShutdown(GPUInfo *g)
{
  int i;
  if (!g) return;
  ObitCUDASkyModelDFTShutdown(g);
  if (g->h_data) ObitCUDAUtilFreeHost(g->h_data);
  if (g->h_visInfo) {
    if (g->h_visInfo->d_freqRat) ObitCUDAUtilFreeGPU(g->h_visInfo->d_freqRat);
    if (g->h_visInfo->freqRat) ObitCUDAUtilFreeHost(g->h_visInfo->freqRat);
    ObitCUDAUtilFreeHost((float *)g->h_visInfo);
  }
  if (g->h_modelInfo) {
    if (g->h_modelInfo->d_specIndex) ObitCUDAUtilFreeGPU((float *)g->h_modelInfo->d_specIndex);
    if (g->h_modelInfo->specIndex) ObitCUDAUtilFreeHost((float *)g->h_modelInfo->specIndex);
    ObitCUDAUtilFreeHost((float *)g->h_modelInfo);
  }
  if (g->d_visInfo) ObitCUDAUtilFreeGPU((float *)g->d_visInfo);
  if (g->d_modelInfo) ObitCUDAUtilFreeGPU((float *)g->d_modelInfo);
  if (g->d_model[0]) ObitCUDAUtilFreeGPU(g->d_model[0]);
  if (g->d_freq) ObitCUDAUtilFreeGPU(g->d_freq);
  for (i = 0; i < g->nstream; i++) {
    if (g->stream) ObitCUDAStreamDestroy(g->stream[i]);
    if (g->cycleDone) ObitCUDAEventDestroy(g->cycleDone[i]);
    if (g->d_data) ObitCUDAUtilFreeGPU(g->d_data[i]);
  }
  free(g->stream);
  free(g->cycleDone);
  free(g->d_data);
  free(g);
}
