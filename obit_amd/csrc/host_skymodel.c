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

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Obit component types (include/ObitSkyModel.h:131-163). */
enum { PointMod = 0, GaussMod = 1, PointModSpec = 10, GaussModSpec = 11,
       PointModTSpec = 20, GaussModTSpec = 21 };

/* The host object: Obit's GPUInfo first (the pointer handed out), then a
 * second pinned buffer so the copies of one nVisPIO piece overlap the
 * device work of the next (ObitHostDFTCalc). */
typedef struct {
  GPUInfo g;
  float *h_alt;
} HostDFT;

static HostDFT *host_dft(GPUInfo *g) { return (HostDFT *)g; }

GPUInfo *ObitHostDFTCreate(void)
{
  HostDFT *h = (HostDFT *)calloc(1, sizeof(HostDFT));
  return h ? &h->g : NULL;
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
  host_dft(g)->h_alt = ObitCUDAUtilAllocHost(g->bufferSize * (int)sizeof(float));
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

/* memcpy of a large buffer in slices on several threads: one thread's
 * copy into pinned memory runs well below the host's memory bandwidth,
 * and the per-buffer path copies every row in and out of the pinned
 * buffer. */
typedef struct { char *dst; const char *src; size_t n; } copy_slice;
static void *copy_run(void *a)
{
  copy_slice *c = (copy_slice *)a;
  memcpy(c->dst, c->src, c->n);
  return NULL;
}
static void par_memcpy(void *dst, const void *src, size_t n)
{
  enum { kMaxThreads = 8 };
  const size_t minSlice = (size_t)4 << 20;
  int nt = (int)(n / minSlice);
  if (nt > kMaxThreads) nt = kMaxThreads;
  if (nt < 2) { memcpy(dst, src, n); return; }
  pthread_t th[kMaxThreads];
  copy_slice cs[kMaxThreads];
  const size_t per = (n / nt + 63) & ~(size_t)63;
  size_t off = 0;
  int started = 0;
  for (int i = 0; i < nt && off < n; i++) {
    cs[i].dst = (char *)dst + off;
    cs[i].src = (const char *)src + off;
    cs[i].n = (n - off < per) ? n - off : per;
    off += cs[i].n;
    if (i == nt - 1 || off >= n) { copy_run(&cs[i]); break; }  /* the last slice on this thread */
    if (pthread_create(&th[i], NULL, copy_run, &cs[i]) != 0) { copy_run(&cs[i]); continue; }
    started |= 1 << i;
  }
  for (int i = 0; i < nt; i++)
    if (started & (1 << i)) pthread_join(th[i], NULL);
}

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* The copies a helper thread makes while the device works on the other
 * pinned buffer: the previous piece out, then the next piece in. */
typedef struct {
  float *outDst, *outSrc, *inDst, *inSrc;
  size_t outBytes, inBytes;
  double seconds;
} copy_job;
static void *copy_job_run(void *a)
{
  copy_job *j = (copy_job *)a;
  double t0 = now_s();
  if (j->outBytes) par_memcpy(j->outDst, j->outSrc, j->outBytes);
  if (j->inBytes) par_memcpy(j->inDst, j->inSrc, j->inBytes);
  j->seconds = now_s() - t0;
  return NULL;
}

/* ObitGPUSkyModelDFTCalc (src/ObitGPUSkyModel.c:500-520), applied to a
 * buffer of any length in nVisPIO-row pieces, as ObitSkyModelSubUV reads
 * and writes one uv buffer at a time (src/ObitSkyModel.c:1000-1026).  The
 * reference copies a piece into the pinned buffer, runs the DFT and copies
 * it back, in turn; here the pieces alternate between two pinned buffers
 * and a helper thread copies the previous piece out and the next one in
 * while ObitCUDASkyModelDFTCalc works on the current one. */
int ObitHostDFTCalcTimed(GPUInfo *g, float *buffer, long long nvis, double times[3])
{
  float *bufs[2];
  long long r0 = 0, prev0 = 0;
  int per, k, nprev = 0;
  double t0, tc, calc = 0.0, copies = 0.0, total;
  if (!g || !buffer || g->lenvis <= 0) return OBIT_HIP_EINVAL;
  per = g->bufferSize / g->lenvis;
  if (per <= 0) return OBIT_HIP_EINVAL;
  if (nvis <= 0) return OBIT_HIP_OK;
  bufs[0] = g->h_data;
  bufs[1] = host_dft(g)->h_alt ? host_dft(g)->h_alt : g->h_data;
  t0 = now_s();
  {
    int n = (int)(nvis < per ? nvis : per);
    par_memcpy(bufs[0], buffer, (size_t)n * g->lenvis * sizeof(float));
  }
  copies += now_s() - t0;
  for (k = 0; r0 < nvis; k++) {
    const int n = (int)((nvis - r0) < per ? (nvis - r0) : per);
    const long long r1 = r0 + n;
    const int nnext = (int)((nvis - r1) < per ? (nvis - r1) : per);
    float *cur = bufs[k & 1], *other = bufs[(k + 1) & 1];
    copy_job job;
    pthread_t th;
    int threaded = 0;
    memset(&job, 0, sizeof(job));
    if (other != cur) {
      /* the other buffer: the previous piece out, the next piece in */
      if (nprev) {
        job.outDst = buffer + prev0 * g->lenvis;
        job.outSrc = other;
        job.outBytes = (size_t)nprev * g->lenvis * sizeof(float);
      }
      if (nnext > 0) {
        job.inDst = other;
        job.inSrc = buffer + r1 * g->lenvis;
        job.inBytes = (size_t)nnext * g->lenvis * sizeof(float);
      }
      if (job.outBytes || job.inBytes) threaded = pthread_create(&th, NULL, copy_job_run, &job) == 0;
      if (!threaded) copy_job_run(&job);
    }
    g->h_data = cur;
    g->nvis = n;
    g->h_visInfo->nvis = n;
    tc = now_s();
    ObitCUDASkyModelDFTCalc(g);
    calc += now_s() - tc;
    if (threaded) pthread_join(th, NULL);
    copies += job.seconds;
    if (other == cur) {
      /* one pinned buffer: out, then the next piece in, in turn */
      tc = now_s();
      par_memcpy(buffer + r0 * g->lenvis, cur, (size_t)n * g->lenvis * sizeof(float));
      if (nnext > 0) par_memcpy(cur, buffer + r1 * g->lenvis, (size_t)nnext * g->lenvis * sizeof(float));
      copies += now_s() - tc;
      nprev = 0;
    } else {
      prev0 = r0;
      nprev = n;
    }
    r0 = r1;
  }
  if (nprev) {
    tc = now_s();
    par_memcpy(buffer + prev0 * g->lenvis, bufs[(k - 1) & 1], (size_t)nprev * g->lenvis * sizeof(float));
    copies += now_s() - tc;
  }
  g->h_data = bufs[0];
  total = now_s() - t0;
  if (times) {
    times[0] += total;
    times[1] += calc;
    times[2] += copies;
  }
  return OBIT_HIP_OK;
}

int ObitHostDFTCalc(GPUInfo *g, float *buffer, long long nvis)
{
  return ObitHostDFTCalcTimed(g, buffer, nvis, NULL);
}

/* ObitGPUSkyModelDFTShutdown (src/ObitGPUSkyModel.c:530-579). */
void ObitHostDFTShutdown(GPUInfo *g)
{
  int i;
  if (!g) return;
  ObitCUDASkyModelDFTShutdown(g);
  if (g->h_data) ObitCUDAUtilFreeHost(g->h_data);
  if (host_dft(g)->h_alt) ObitCUDAUtilFreeHost(host_dft(g)->h_alt);
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
  free(host_dft(g));
}

/* ------------------------------------------------------------------ */
/* Component table of one image field (ObitSkyModelLoadComps)          */
/* ------------------------------------------------------------------ */

/* glib's G_PI and Obit's DG2RAD macro (include/ObitTypes.h:43), kept as
 * macros: DG2RAD expands textually, so "a * DG2RAD" is (a * pi) / 180. */
#define OBIT_G_PI 3.1415926535897932384626433832795028841971693993751
#define OBIT_DG2RAD OBIT_G_PI / 180.0

/* ObitSkyGeomShiftSIN (src/ObitSkyGeom.c:177-205). */
static void geom_shift_sin(double ra, double dec, float rotate, double xra, double xdec,
                           float dxyzc[3])
{
  float maprr, xshift, yshift;
  double xxshft, yyshft, dzctmp;
  maprr = rotate * OBIT_DG2RAD;
  xxshft = cos(OBIT_DG2RAD * xdec) * sin(OBIT_DG2RAD * (xra - ra));
  yyshft = sin(OBIT_DG2RAD * xdec) * cos(OBIT_DG2RAD * dec) -
           cos(OBIT_DG2RAD * xdec) * sin(OBIT_DG2RAD * dec) * cos(OBIT_DG2RAD * (xra - ra));
  xshift = (cos(maprr) * xxshft + sin(maprr) * yyshft);
  yshift = (-sin(maprr) * xxshft + cos(maprr) * yyshft);
  dxyzc[0] = xshift;
  dxyzc[1] = yshift;
  dzctmp = 1.0 - (dxyzc[0] * dxyzc[0]) - (dxyzc[1] * dxyzc[1]);
  dzctmp = sqrt(dzctmp);
  dxyzc[0] = 2.0 * OBIT_G_PI * dxyzc[0];
  dxyzc[1] = 2.0 * OBIT_G_PI * dxyzc[1];
  dxyzc[2] = 2.0 * OBIT_G_PI * (dzctmp - 1.0);
}

/* The uv re-projection matrix of ObitUVDescShift3DMatrix for a 3-D image
 * (src/ObitUVDesc.c:882-975, imDesc->do3D branch); returns do3Dmul. */
static int geom_shift3d_umat(double ra, double dec, float urotat, double xra, double xdec,
                             float mrotat, float xPxOff, float yPxOff, float umat[3][3])
{
  double rm[3][3], x[3][3], t[3][3], r, sa, ca, sd, cd, sd0, cd0;
  int i, j, k, do3Dmul;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) rm[i][j] = 0.0;
  rm[2][2] = 1.0;
  do3Dmul = (mrotat != 0.0) || (urotat != 0.0) || (fabs(xra - ra) > 1.0e-10) ||
            (fabs(xdec - dec) > 1.0e-10) || (fabs(xPxOff) > 1.0e-20) || (fabs(yPxOff) > 1.0e-20);
  sa = sin(OBIT_DG2RAD * (xra - ra));
  ca = cos(OBIT_DG2RAD * (xra - ra));
  sd = sin(OBIT_DG2RAD * xdec);
  cd = cos(OBIT_DG2RAD * xdec);
  sd0 = sin(OBIT_DG2RAD * dec);
  cd0 = cos(OBIT_DG2RAD * dec);
  /* map + */
  r = OBIT_DG2RAD * mrotat;
  rm[0][0] = cos(r);
  rm[1][1] = rm[0][0];
  rm[1][0] = -sin(r);
  rm[0][1] = sin(r);
  /* forward matrix */
  x[0][0] = ca;
  x[1][0] = -sd * sa;
  x[2][0] = cd * sa;
  x[0][1] = sd0 * sa;
  x[1][1] = cd * cd0 + sd * sd0 * ca;
  x[2][1] = sd * cd0 - cd * sd0 * ca;
  x[0][2] = -cd0 * sa;
  x[1][2] = cd * sd0 - sd * cd0 * ca;
  x[2][2] = sd * sd0 + cd * cd0 * ca;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      t[j][i] = 0.0;
      for (k = 0; k < 3; k++) t[j][i] += x[k][i] * rm[j][k];
    }
  /* uv - */
  r = OBIT_DG2RAD * urotat;
  rm[0][0] = cos(r);
  rm[1][1] = rm[0][0];
  rm[1][0] = sin(r);
  rm[0][1] = -sin(r);
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      r = 0.0;
      for (k = 0; k < 3; k++) r += rm[k][i] * t[j][k];
      umat[j][i] = r;
    }
  return do3Dmul;
}

int ObitHostLoadCompsPoint(const ObitHostFieldGeom *g, const float *cc, int ncc, int ccStride,
                           float factor, float *comps)
{
  float dxyzc[3], umat[3][3], xp[3], xyz[3], konst, ssrot, ccrot, xxoff, yyoff, zzoff;
  int j, do3Dmul;
  if (!g || !cc || !comps || ncc < 0 || ccStride < 3) return -1;
  if (!g->do3D) return -1;  /* 2-D fields (tangent-point offsets) are not restated */
  /* ObitUVDescShiftPhase for a 3-D image: the shift to the image's
   * reference position (src/ObitUVDesc.c:735-781) */
  geom_shift_sin(g->uvCrval[0], g->uvCrval[1], g->imCrota, g->imCrval[0], g->imCrval[1], dxyzc);
  ssrot = sin(OBIT_DG2RAD * (g->uvCrota - g->imCrota));
  ccrot = cos(OBIT_DG2RAD * (g->uvCrota - g->imCrota));
  xxoff = dxyzc[0] * ccrot + dxyzc[1] * ssrot;
  yyoff = dxyzc[1] * ccrot - dxyzc[0] * ssrot;
  zzoff = dxyzc[2];
  do3Dmul = geom_shift3d_umat(g->uvCrval[0], g->uvCrval[1], g->uvCrota, g->imCrval[0], g->imCrval[1],
                              g->imCrota, g->xPxOff, g->yPxOff, umat);
  konst = OBIT_DG2RAD * 2.0 * OBIT_G_PI;
  for (j = 0; j < ncc; j++) {
    const float *a = cc + (size_t)j * ccStride;
    float *row = comps + (size_t)j * 4;
    row[0] = a[0] * factor;
    xp[0] = (a[1] + 0.0f) * konst;  /* 3-D: no reference-pixel offsets */
    xp[1] = (a[2] + 0.0f) * konst;
    xp[2] = 0.0f;
    if (do3Dmul) {
      xyz[0] = xp[0] * umat[0][0] + xp[1] * umat[1][0];
      xyz[1] = xp[0] * umat[0][1] + xp[1] * umat[1][1];
      xyz[2] = xp[0] * umat[0][2] + xp[1] * umat[1][2];
    } else {
      xyz[0] = ccrot * xp[0] + ssrot * xp[1];
      xyz[1] = ccrot * xp[1] - ssrot * xp[0];
      xyz[2] = xp[2];
    }
    row[1] = xyz[0] + xxoff;
    row[2] = xyz[1] + yyoff;
    row[3] = xyz[2] + zzoff;
  }
  return ncc;
}

/* ------------------------------------------------------------------ */
/* Merged CC rows of one table (ObitTableCCUtilMergeSel)               */
/* ------------------------------------------------------------------ */

/* Entries are compared by position (x, then y), ties in input order, as
 * glib's stable g_qsort_with_data orders them with CCComparePos
 * (src/ObitTableCCUtil.c:3099-3119); then by descending |flux|, ties in
 * position order (CCCompareFlux, :3130-3143). */
typedef struct {
  const float *e;  /* entry: x, y, flux, [z], terms */
  int order;       /* rank in the previous order */
} CCKey;

static int cc_cmp_pos(const void *a, const void *b)
{
  const CCKey *p = (const CCKey *)a, *q = (const CCKey *)b;
  if (p->e[0] < q->e[0]) return -1;
  if (p->e[0] > q->e[0]) return 1;
  if (p->e[1] < q->e[1]) return -1;
  if (p->e[1] > q->e[1]) return 1;
  return (p->order > q->order) - (p->order < q->order);
}

static int cc_cmp_flux(const void *a, const void *b)
{
  const CCKey *p = (const CCKey *)a, *q = (const CCKey *)b;
  const float fp = fabsf(p->e[2]), fq = fabsf(q->e[2]);
  if (fp < fq) return 1;
  if (fp > fq) return -1;
  return (p->order > q->order) - (p->order < q->order);
}

int ObitHostCCMergeWidth(int haveZ, int noParms)
{
  return 3 + (haveZ ? 1 : 0) + (noParms > 4 ? noParms - 4 : 0);
}

int ObitHostCCMergeSel(const float *rows, int nrow, int haveZ, int noParms, const int *status,
                       int startComp, int endComp, float *out, float parms[4])
{
  int fsize, toff, nterms, i, j, k, n = 0, nout = 0, doSpec, doTSpec, rowLen;
  float lparms[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  float *ent;
  CCKey *key;
  if (!rows || nrow < 0 || noParms < 0 || !out || !parms) return -1;
  toff = haveZ ? 4 : 3;
  nterms = noParms > 4 ? noParms - 4 : 0;
  fsize = toff + nterms;
  rowLen = toff + noParms;
  for (i = 0; i < 4; i++) parms[i] = 0.0f;
  if (nrow == 0) return 0;
  if (startComp < 1) startComp = 1;
  if (endComp > nrow) endComp = nrow;
  if (endComp < startComp) return 0;
  ent = (float *)malloc(sizeof(float) * (size_t)fsize * (size_t)(endComp - startComp + 1));
  key = (CCKey *)malloc(sizeof(CCKey) * (size_t)(endComp - startComp + 1));
  if (!ent || !key) { free(ent); free(key); return -1; }
  /* MakeCCSortStructSel (:2889-2975): x, y, flux, [z], parms[4..] */
  for (i = startComp - 1; i < endComp; i++) {
    const float *r = rows + (size_t)i * rowLen;
    float *e = ent + (size_t)n * fsize;
    if (status && status[i] < 0) continue;
    e[0] = r[1];
    e[1] = r[2];
    e[2] = r[0];
    if (haveZ) e[3] = r[3];
    for (j = 0; j < nterms; j++) e[toff + j] = r[toff + 4 + j];
    if (n == 0)
      for (j = 0; j < 4 && j < noParms; j++) lparms[j] = r[toff + j];
    key[n].e = e;
    key[n].order = n;
    n++;
  }
  qsort(key, (size_t)n, sizeof(CCKey), cc_cmp_pos);
  doSpec = (lparms[3] >= 9.9f) && (lparms[3] <= 19.0f);
  doTSpec = (lparms[3] >= 19.9f) && (lparms[3] <= 29.0f);
  /* CCMerge / CCMergeSpec (:3153-3248): runs of one position summed into
   * their first entry (flux; with tabulated spectra the spectral terms
   * too); parameterised spectra are not merged */
  if (!doSpec) {
    i = 0;
    while (i < n) {
      float *a = (float *)key[i].e;
      j = i + 1;
      while (j < n) {
        float *b = (float *)key[j].e;
        if (b[0] != a[0] || b[1] != a[1]) break;
        if (doTSpec)
          for (k = toff; k < fsize; k++) a[k] += b[k];
        a[2] += b[2];
        b[0] = -1.0e20f;
        j++;
      }
      i = j;
    }
  }
  for (i = 0; i < n; i++) key[i].order = i;
  qsort(key, (size_t)n, sizeof(CCKey), cc_cmp_flux);
  for (i = 0; i < n; i++) {
    const float *e = key[i].e;
    float *o = out + (size_t)nout * fsize;
    if (!(e[0] > -1.0e19f)) continue;
    o[0] = e[2];
    o[1] = e[0];
    o[2] = e[1];
    if (haveZ) o[3] = e[3];
    for (j = 0; j < nterms; j++) o[toff + j] = e[toff + j];
    nout++;
  }
  for (i = 0; i < 4 && i < noParms; i++) parms[i] = lparms[i];
  free(ent);
  free(key);
  return nout;
}

/* ObitSkyGeomWorldPos (src/ObitSkyGeom.c:492-515) with ObitSkyGeomWorldPosLM
 * (:535-711) for the -SIN projection; returns nonzero where the projection
 * cannot place the pixel. */
static int geom_world_pos(float xpix, float ypix, double xref, double yref, float xrefpix,
                          float yrefpix, float xinc, float yinc, float rot, double *xpos, double *ypos)
{
  const double cond2r = 1.745329252e-2, twopi = 6.28318530717959;
  double cosr, sinr, dx, dy, temp, ra0, dec0, l, m, sins, coss, dt, dect, rat, raout, cos0, sin0;
  dx = (xpix - xrefpix) * xinc;
  dy = (ypix - yrefpix) * yinc;
  cosr = cos(rot * cond2r);
  sinr = sin(rot * cond2r);
  if (rot != 0.0) {
    temp = dx * cosr - dy * sinr;
    dy = dy * cosr + dx * sinr;
    dx = temp;
  }
  *xpos = xref + dx;
  *ypos = yref + dy;
  ra0 = xref * cond2r;
  dec0 = yref * cond2r;
  l = dx * cond2r;
  m = dy * cond2r;
  sins = l * l + m * m;
  cos0 = cos(dec0);
  sin0 = sin(dec0);
  if (sins > 1.0) return 1;
  coss = sqrt(1.0 - sins);
  dt = sin0 * coss + cos0 * m;
  if ((dt > 1.0) || (dt < -1.0)) return 1;
  dect = asin(dt);
  rat = cos0 * coss - sin0 * m;
  if ((rat == 0.0) && (l == 0.0)) return 1;
  rat = atan2(l, rat) + ra0;
  raout = rat;
  if (raout - ra0 > twopi / 2.0) raout = raout - twopi;
  if (raout - ra0 < -twopi / 2.0) raout = raout + twopi;
  if (raout < 0.0) raout += twopi;
  *xpos = raout / cond2r;
  *ypos = dect / cond2r;
  return 0;
}

/* The uv re-projection matrix of ObitUVDescShift3DMatrix for a 2-D image
 * (src/ObitUVDesc.c:1024-1093): the w term of the field centre's
 * direction, (xra, xdec) its position. */
static void geom_shift2d_umat(double ra, double dec, float urotat, double xra, double xdec,
                              float mrotat, float umat[3][3])
{
  double rm[3][3], x[3][3], t[3][3], r, sa, ca, sd, cd, sd0, cd0, ll, mm, dd;
  int i, j, k;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) rm[i][j] = 0.0;
  rm[2][2] = 1.0;
  sa = sin(OBIT_DG2RAD * (xra - ra));
  ca = cos(OBIT_DG2RAD * (xra - ra));
  sd = sin(OBIT_DG2RAD * xdec);
  cd = cos(OBIT_DG2RAD * xdec);
  sd0 = sin(OBIT_DG2RAD * dec);
  cd0 = cos(OBIT_DG2RAD * dec);
  ll = cd * sa;
  mm = sd * cd0 - cd * sd0 * ca;
  dd = sqrt(1.0 - ll * ll - mm * mm);
  r = OBIT_DG2RAD * mrotat;
  rm[0][0] = cos(r);
  rm[1][1] = rm[0][0];
  rm[1][0] = -sin(r);
  rm[0][1] = sin(r);
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) x[i][j] = 0.0;
  x[0][0] = 1.0;
  x[1][1] = 1.0;
  x[2][2] = -1.0;
  x[0][2] = -ll / dd;
  x[1][2] = -mm / dd;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      t[j][i] = 0.0;
      for (k = 0; k < 3; k++) t[j][i] += x[k][i] * rm[j][k];
    }
  r = OBIT_DG2RAD * urotat;
  rm[0][0] = cos(r);
  rm[1][1] = rm[0][0];
  rm[1][0] = sin(r);
  rm[0][1] = -sin(r);
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      r = 0.0;
      for (k = 0; k < 3; k++) r += rm[k][i] * t[j][k];
      umat[j][i] = r;
    }
}

int ObitHostLoadComps(const ObitHostFieldGeom *g, const float *merged, int nmerged, int haveZ,
                      int nterms, const float parms[4], const ObitHostLoadOpts *o, float *comps,
                      int maxComp)
{
  float dxyzc[3], umat[3][3], xp[3], xyz[3], konst, konst2, ssrot, ccrot, xxoff, yyoff, zzoff;
  float xpoff, ypoff, cpa, spa, xmaj, xmin, gp1 = 0.0f, gp2 = 0.0f, gp3 = 0.0f, inPixel[2];
  double imra, imdec, xra, xdec;
  int j, iterm, do3Dmul, toff, stride, lenEntry, fieldType, tableType, gaussTab, ncomp = 0;
  if (!g || !o || !parms || (!merged && nmerged > 0) || (!comps && nmerged > 0) || nmerged < 0 ||
      nterms < 0 || maxComp < 0)
    return -1;
  fieldType = (int)(parms[3] + 0.5f);
  tableType = o->modType;
  if (fieldType > tableType) return -1;
  if (!((fieldType == tableType && (tableType == 0 || tableType == 1 || tableType == 10 || tableType == 11)) ||
        (fieldType == 0 && (tableType == 1 || tableType == 10)) || (fieldType == 10 && tableType == 11)))
    return -1;
  if ((fieldType == 10 || fieldType == 11) && o->nSpecTerm > nterms) return -1;
  gaussTab = tableType == 1 || tableType == 11;
  lenEntry = 4 + (gaussTab ? 3 : 0) + ((tableType >= 10) ? o->nSpecTerm : 0);
  toff = haveZ ? 4 : 3;
  stride = toff + nterms;
  konst = OBIT_DG2RAD * 2.0 * OBIT_G_PI;
  konst2 = OBIT_DG2RAD * sqrt(2.0) * OBIT_G_PI / 2.35482044;
  /* ObitUVDescShiftPhase (src/ObitUVDesc.c:735-794): the field's reference
   * position, for a 2-D image the position of its reference pixel (or of
   * its centre when shifted) */
  if (g->do3D) {
    imra = g->imCrval[0];
    imdec = g->imCrval[1];
  } else {
    if ((g->xshift != 0.0f) || (g->yshift != 0.0f)) {
      inPixel[0] = 1.0 + g->imNaxis[0] * 0.5;
      inPixel[1] = 1.0 + g->imNaxis[1] * 0.5;
    } else {
      inPixel[0] = g->imCrpix[0];
      inPixel[1] = g->imCrpix[1];
    }
    if (geom_world_pos(inPixel[0], inPixel[1], g->imCrval[0], g->imCrval[1], g->imCrpix[0],
                       g->imCrpix[1], g->imCdelt[0], g->imCdelt[1], g->imCrota, &imra, &imdec))
      return -1;
  }
  geom_shift_sin(g->uvCrval[0], g->uvCrval[1], g->imCrota, imra, imdec, dxyzc);
  ssrot = sin(OBIT_DG2RAD * (g->uvCrota - g->imCrota));
  ccrot = cos(OBIT_DG2RAD * (g->uvCrota - g->imCrota));
  if (g->do3D) {
    xpoff = 0.0f;
    ypoff = 0.0f;
  } else {
    xpoff = g->xPxOff * g->imCdelt[0];
    ypoff = g->yPxOff * g->imCdelt[1];
  }
  xxoff = dxyzc[0] * ccrot + dxyzc[1] * ssrot;
  yyoff = dxyzc[1] * ccrot - dxyzc[0] * ssrot;
  zzoff = dxyzc[2];
  /* ObitUVDescShift3DMatrix (src/ObitUVDesc.c:882-1121) */
  do3Dmul = geom_shift3d_umat(g->uvCrval[0], g->uvCrval[1], g->uvCrota, g->imCrval[0], g->imCrval[1],
                              g->imCrota, g->xPxOff, g->yPxOff, umat);
  if (!g->do3D) {
    inPixel[0] = 1.0 + g->imNaxis[0] * 0.5;
    inPixel[1] = 1.0 + g->imNaxis[1] * 0.5;
    if (geom_world_pos(inPixel[0], inPixel[1], g->imCrval[0], g->imCrval[1], g->imCrpix[0],
                       g->imCrpix[1], g->imCdelt[0], g->imCdelt[1], g->imCrota, &xra, &xdec))
      return -1;
    geom_shift2d_umat(g->uvCrval[0], g->uvCrval[1], g->uvCrota, xra, xdec, g->imCrota, umat);
  }
  /* Gaussian terms (src/ObitSkyModel.c:1939-1947) */
  if (fieldType == 1 || fieldType == 11) {
    cpa = cos(OBIT_DG2RAD * parms[2]);
    spa = sin(OBIT_DG2RAD * parms[2]);
    xmaj = parms[0] * konst2;
    xmin = parms[1] * konst2;
    gp1 = -(((cpa * xmaj) * (cpa * xmaj)) + (spa * xmin) * (spa * xmin));
    gp2 = -(((spa * xmaj) * (spa * xmaj)) + (cpa * xmin) * (cpa * xmin));
    gp3 = -2.0 * cpa * spa * (xmaj * xmaj - xmin * xmin);
  }
  for (j = 0; j < nmerged; j++) {
    const float *a = merged + (size_t)j * stride;
    float *table;
    int want;
    if (o->noNeg && (a[0] <= 0.0f)) break;
    want = (fabsf(a[0]) > o->minDFT);
    want = want && (fabsf(a[0]) > o->minFlux);
    want = want && (ncomp < maxComp);
    if (!want) continue;
    table = comps + (size_t)ncomp * lenEntry;
    table[0] = a[0] * o->factor;
    xp[0] = (a[1] + xpoff) * konst;
    xp[1] = (a[2] + ypoff) * konst;
    xp[2] = haveZ ? a[3] * konst : 0.0f;
    if (do3Dmul) {
      xyz[0] = xp[0] * umat[0][0] + xp[1] * umat[1][0];
      xyz[1] = xp[0] * umat[0][1] + xp[1] * umat[1][1];
      xyz[2] = xp[0] * umat[0][2] + xp[1] * umat[1][2];
    } else {
      xyz[0] = ccrot * xp[0] + ssrot * xp[1];
      xyz[1] = ccrot * xp[1] - ssrot * xp[0];
      xyz[2] = xp[2];
    }
    table[1] = xyz[0] + xxoff;
    table[2] = xyz[1] + yyoff;
    table[3] = xyz[2] + zzoff;
    for (iterm = 4; iterm < lenEntry; iterm++) table[iterm] = 0.0f;
    if (gaussTab && (fieldType == 1 || fieldType == 11)) {
      table[4] = gp1;
      table[5] = gp2;
      table[6] = gp3;
    }
    if (fieldType >= 10)
      for (iterm = 0; iterm < o->nSpecTerm; iterm++) table[iterm + (gaussTab ? 7 : 4)] = a[iterm + toff];
    ncomp++;
  }
  return ncomp;
}
