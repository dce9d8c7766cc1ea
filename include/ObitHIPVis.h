/*
 * ObitHIPVis.h -- extensions of the drop-in C ABI for whole-data-set
 * residency in HBM and batched launches.  These have no counterpart in
 * Obit's CUDA layer: the reference boundary moves one uv buffer (<= 10,000
 * rows, ObitGPUSkyModel.c:236-245) per call.  A caller that keeps every
 * visibility of a field resident (Obit's MFImage/Imager major cycle) calls
 * these instead of ObitCUDASkyModelDFTCalc / ObitCUDAGridGrid per buffer.
 * Every function returns 0 on success or a status below, with a message
 * from ObitHIPLastError().
 */
#ifndef OBITHIPVIS_H
#define OBITHIPVIS_H
#include <stddef.h>
#include "ObitCUDASkyModelInfoDef.h"
#include "ObitCUDAGridInfoDef.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OBIT_HIP_OK        0
#define OBIT_HIP_EINVAL    1   /* bad argument or layout */
#define OBIT_HIP_EHIP      2   /* HIP runtime error */
#define OBIT_HIP_ENOMEM    3   /* device allocation failed */
#define OBIT_HIP_ESTATE    4   /* call out of order (e.g. no model set) */

const char *ObitHIPLastError(void);

/* ---- memory helpers with size_t sizes (the reference's int sizes stop at
 * 2 GiB, ObitCUDAUtil.h:99-152) ---- */
int ObitHIPMalloc(void **ptr, size_t bytes);
int ObitHIPFree(void *ptr);
int ObitHIPMemcpyH2D(void *dst, const void *src, size_t bytes);
int ObitHIPMemcpyD2H(void *dst, const void *src, size_t bytes);
int ObitHIPMemcpyD2D(void *dst, const void *src, size_t bytes);
int ObitHIPDeviceSynchronize(void);
float ObitHIPEventElapsedMs(int *start, int *stop);
int ObitHIPDeviceInfo(int device, char *name, int nameLen, int *nCU, int *clockKHz,
                      long long *memBytes);
/* Plain device-to-device copy kernel on stream (bytes a multiple of 16):
 * the known-size reference against which HBM traffic counters are
 * calibrated. */
int ObitHIPCopyKernel(void *dst, const void *src, size_t bytes, void *stream);

/* ---- DFT sky model ---- */

/* Channel frequencies in FP64 (Hz, in the freqScale order of the data,
 * nchan*nIF values) and the uv reference frequency.  Without this call the
 * library fits each IF's FP32 freqScale with a line in FP64; with it, the
 * phase 2*pi*(ul+vm+wn)*nu/nu0 is formed from these exact values. */
int ObitHIPSkyModelDFTSetFreq(GPUInfo *gpuInfo, const double *freqArr, int nfreq,
                              double uvRefFreq);
/* Kernel the last DFT launch used: *path = 0 the split-precision kernel,
 * 1 the all-FP64 kernel.  Without ObitHIPSkyModelDFTSetFreq the channel
 * frequencies come from the FP32 freqScale; the split kernel walks each
 * IF's channels along a line fitted to them, and a table whose values sit
 * off that line by more than 4e-7 (relative; 1e-12 with FP64 frequencies)
 * runs on the FP64 kernel instead, with a one-time message on stderr.
 * freqScale is read again whenever its pointer changes or a table has been
 * uploaded through ObitCUDAUtilHost2GPU / ObitCUDAUtilHostAny2GPU since. */
int ObitHIPSkyModelDFTGetPath(GPUInfo *gpuInfo, int *path);
/* Shape of the last ObitCUDASkyModelDFTCalc launch: channels per thread
 * and the number of component parts whose sums a second kernel adds (1:
 * none).  A buffer of few rows makes few (256-row, IF, channel chunk)
 * work items, so DFTCalc also cuts the component list into parts to give
 * every CU an equal share. */
int ObitHIPSkyModelDFTGetBufferShape(GPUInfo *gpuInfo, int *cpt, int *nsplit);
/* CalcResident's operation on nvis rows at d_vis (the caller's device
 * buffer) with the component table d_model (the caller's device memory,
 * ncomp rows of h_modelInfo->size floats), without allocating anything.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The layout,
 * selection, model type and operation are those DFTInit / DFTSetMod set up;
 * the per-IF frequency lines and the prepared component table go to a
 * static device scratch of 32 MiB per GPU (838,809 components at most;
 * more is refused with OBIT_HIP_EINVAL, as are spectral model types, more
 * than 64 selected IFs and channel tables that are not a line within each
 * IF).  Asynchronous on ObitHIPSkyModelDFTStream; calls that overlap on one
 * GPU share the scratch and race. */
int ObitHIPSkyModelDFTCalcNoAlloc(GPUInfo *gpuInfo, const float *d_model, int ncomp, float *d_vis,
                                  long long nvis);
/* Reference frequencies of the component spectra in FP64 (Hz): one value
 * (parametrised spectra: the image reference frequency) or nSpec values
 * (tabulated spectra: the coarse-bin frequencies).  Without this call they
 * are derived from h_visInfo->freqRat. */
int ObitHIPSkyModelDFTSetSpecRef(GPUInfo *gpuInfo, const double *refFreq, int n);

/* Resident visibilities: copy nvis rows (lenvis floats each) to HBM, or
 * attach a device buffer the caller owns.  CalcResident applies the model
 * to every resident row in one persistent launch. */
int ObitHIPSkyModelDFTSetVis(GPUInfo *gpuInfo, const float *h_vis, long long nvis);
int ObitHIPSkyModelDFTAttachVis(GPUInfo *gpuInfo, float *d_vis, long long nvis);
int ObitHIPSkyModelDFTGetVis(GPUInfo *gpuInfo, float *h_vis, long long nvis);
int ObitHIPSkyModelDFTCalcResident(GPUInfo *gpuInfo);
/* The same operation with the plain all-FP64 kernel (FP64 phase, FP64
 * sincos per evaluation, FP64 sums): the precision/throughput comparator. */
int ObitHIPSkyModelDFTCalcFP64(GPUInfo *gpuInfo);
/* CalcResident captured once into a hipGraph and replayed. */
int ObitHIPSkyModelDFTGraphLaunch(GPUInfo *gpuInfo);
/* Kernel shape: channels per thread (16 or 32), component tile (32, 64
 * or 128), workgroups per CU of the persistent grid (1..8); cpt 64 selects
 * the matrix-core kernel (point models with one channel step for every
 * IF, in groups of 64 selected channels, a ragged last group formed in
 * full and only its selected channels written; for a subtraction with a
 * Gaussian table on such channels, its components without Gaussian terms,
 * the rest on the VALU kernel; other launches run the 32-channel VALU
 * kernel), which ignores tile and wgPerCU.  Autotune times
 * every shape on a copy of the first resident rows and keeps the fastest:
 * sampleRows rows, raised where needed so that the widest grid gets at
 * least 8 rounds of work items, and capped at the resident rows. */
int ObitHIPSkyModelDFTSetVariant(GPUInfo *gpuInfo, int cpt, int tile, int wgPerCU);
int ObitHIPSkyModelDFTGetVariant(GPUInfo *gpuInfo, int *cpt, int *tile, int *wgPerCU);
int ObitHIPSkyModelDFTAutotune(GPUInfo *gpuInfo, long long sampleRows, int reps);
/* Sin/cos policy: 0 hardware sin/cos for the centre phasor and the
 * per-channel step; 1 hardware for the centre phasor, Taylor polynomial
 * for the step (default); 2 polynomial for both.  The matrix-core kernel
 * reads its centre phasors P and its steps W, V the same way. */
int ObitHIPSkyModelDFTSetPrecision(GPUInfo *gpuInfo, int prec);
/* hipStream_t on which the library launches for this gpuInfo. */
void *ObitHIPSkyModelDFTStream(GPUInfo *gpuInfo);

/* ---- Gridding ---- */

/* Grid nvis resident rows (device pointer, lenvis floats each) into every
 * facet on iGPU, in one binning pass and one tile-gridding launch.  At
 * most 2^31 - 1 rows a call (OBIT_HIP_EINVAL beyond: the track lists hold
 * 32-bit row indices); a larger data set is gridded in several calls. */
int ObitHIPGridGridDevice(CUDAGridInfo *in, long iGPU, const float *d_vis, long long nvis);
/* The same with the track lists of the previous pass on iGPU: for rows
 * with the same u, v, w (a major cycle regridding residuals).  The lists
 * depend on uvw and the facet geometry only, so the binning pass is
 * skipped.  Refused (OBIT_HIP_ESTATE) when the previous pass had a
 * different row count, the facets were set up again since, or the rows'
 * u, v, w differ from the previous pass's (an order-independent checksum
 * of their bit patterns, computed on the device). */
int ObitHIPGridGridDeviceReuse(CUDAGridInfo *in, long iGPU, const float *d_vis, long long nvis);
/* Trim the hw guard columns and swap the v halves of every plane of a
 * facet (ObitGPUGrid2Host + ObitCArray2DCenter) on the device:
 * out[nplane][ny][nx/2+1][2]. */
int ObitHIPGrid2Device(CUDAGridInfo *in, long ifacet, float *d_out);
/* The same into host memory (out as above). */
int ObitHIPGrid2HostPlanes(CUDAGridInfo *in, long ifacet, float *h_out);
/* Per-sample gridding indices of one facet, computed by the device code
 * the gridder uses: for row r and channel c, index r*nchan+c of iu, iv
 * (first of the 7 rows), itu, itv (GCF block 0..100) and vis[2*(..)]. */
int ObitHIPGridPrepDebug(CUDAGridInfo *in, long ifacet, const float *h_vis, long long nvis,
                         int *iu, int *iv, int *itu, int *itv, float *vis);
/* ObitUVGridFFT2Im (src/ObitUVGrid.c:673-834) on the device for every plane
 * of facet ifacet, after ObitCUDAGridFlip: the C2R FFT (unnormalised, as
 * FFTW's backward transform) of the gridder's padded half planes, the
 * v-half swap (as a (-1)^y factor), the quadrant swap of
 * ObitFArray2DCenter, the gridding correction of GridCorrFn applied as
 * ObitFArrayMulColRow does, and the normalisation.  Sides of 1024 and 2048
 * take two passes (a transform of every column in place, then of every row
 * with the centring, correction and normalisation written to d_image);
 * other sides a batched 2-D hipFFT C2R and one correcting pass.  isBeam=1: beamNorm[p] receives each plane's centre value
 * x corrections and normalises it; isBeam=0: beamNorm[p] (from the beam)
 * normalises the image.  d_image: device [nplane][ny][nx] floats.  Facets
 * must be square with a side a multiple of 8 (the reference's correction
 * indexes rows and columns alike); at most 1024 planes.  The transform
 * consumes the facet's grid planes (as the reference's transforms its
 * grid): read them first if they are wanted.  Returns after the images
 * are written. */
int ObitHIPGridFFT2Image(CUDAGridInfo *in, long ifacet, int isBeam, float *beamNorm, float *d_image);
/* Statistics of the last gridding pass on iGPU (stats has 8 entries):
 * [0] tile lists, [1] non-empty lists, [2] list entries (track x tile),
 * [3] work items, [4] longest list, [5] lists split into several work
 * items, [6] samples accumulated into tiles (sample x tile), [7] moves of
 * the accumulating base cell among them (tile read-modify-writes). */
int ObitHIPGridListStats(CUDAGridInfo *in, long iGPU, long long *stats);
void *ObitHIPGridStream(CUDAGridInfo *in, long iGPU);

/* ---- Stokes I records ---- */

/* ObitUVCalSelect (src/ObitUVCalSelect.c:540-695) for a Stokes 'I'
 * selection as ObitUVCalSelectIInit sets it up (:697-760), over a whole
 * resident data set: d_in holds nvis rows of lenvis floats (nrparm random
 * parameters, then correlations at float increments incs (Stokes), incf
 * (channel), incif (IF), each [re, im, wt]); kstoke0 is the first
 * correlation's Stokes code (-1 RR, -5 XX, 1 I; the CRVAL of the Stokes
 * axis) and nstok their number; formalI requires both hands.  d_out
 * receives nvis single-Stokes rows of nrparm + 3*nif*nchan floats (random
 * parameters copied, then [IF][chan][re, im, wt]), bit-identical to the
 * CPU's.  Returns OBIT_HIP_EINVAL, with ObitUVCalSelectIInit's message,
 * when the data cannot give Stokes I.  Launched on stream (NULL: the
 * default stream); returns without waiting. */
int ObitHIPStokesI(const float *d_in, float *d_out, long long nvis, int lenvis, int nrparm,
                   int nif, int nchan, int nstok, int incs, int incf, int incif, int kstoke0,
                   int formalI, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* OBITHIPVIS_H */
