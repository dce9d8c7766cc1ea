/*
 * host.h -- C host classes over the drop-in ABI.
 *
 * These play the role of Obit's ObitGPUSkyModel.c and ObitGPUGrid.c: they
 * fill GPUInfo / CUDAGridInfo from a uv-data description and a sky model
 * or facet list, then drive the ObitCUDA* entry points exactly as Obit's
 * host classes do (reference src/ObitGPUSkyModel.c:207-579,
 * src/ObitGPUGrid.c:255-721).  The Obit objects (ObitUV, ObitSkyModel,
 * ObitUVGrid) are replaced by the plain structs below, which carry the
 * fields those classes read.
 */
#ifndef OBIT_HIP_HOST_H
#define OBIT_HIP_HOST_H
#include "ObitCUDASkyModelInfoDef.h"
#include "ObitCUDAGridInfoDef.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The ObitUVDesc fields the GPU host classes read (include/ObitUVDesc.h). */
typedef struct {
  int nvis, lrec, nrparm;
  int ilocu, ilocv, ilocw, iloct, ilocb, ilocsu;
  int nchan, nif, nstok;      /* inaxes of FREQ, IF (1 if absent), STOKES */
  int freqBeforeIF;           /* 1: FREQ axis precedes IF (AIPS order) */
  int hasIF;                  /* 1: the data have an IF axis */
  int incs, incf, incif;      /* increments in floats (ObitUVDesc.c:380-392) */
  int ncorr;
  float crvalStokes;          /* crval of the STOKES axis (<0: correlations) */
  double freq;                /* uv reference frequency (Hz) */
  const double *freqArr;      /* [nchan*nif] channel frequencies (Hz) */
  const float *fscale;        /* [nchan*nif] freqArr/freq as FP32 */
} ObitHostUVDesc;

/* The ObitSkyModel / ObitSkyModelMF fields the DFT host class reads. */
typedef struct {
  int modType;                /* ObitSkyModelCompType (include/ObitSkyModel.h:131) */
  int nSpecTerm;
  int doDivide, doReplace, doFlip;
  float factor, stokFactor;
  int startChannel, numberChannel, startIF, numberIF;  /* 1-relative selection */
  int startPoln, numberPoln;  /* 1-relative Stokes selection */
  int isMF;                   /* tabulated spectra (ObitSkyModelMF) */
  int nSpec;
  const double *specFreq;     /* [nSpec] coarse-bin frequencies (Hz) */
  const int *specIndex;       /* [nchan*nif] coarse bin per channel */
  double imgRefFreq;          /* image reference frequency (Hz), 0 if none */
  int cuda_device;
  int useFreq64;              /* pass FP64 frequencies to the library */
} ObitHostSkyModel;

/* The ObitUVGrid fields the gridding host class reads, per facet. */
typedef struct {
  int nx, ny;
  float cellx, celly;         /* cell size (deg) */
  float shift[3];             /* dxc, dyc, dzc */
  float rotUV[9];             /* URot3D */
  float blmax, blmin;
  float bmTaperUV;
  int doBeam;
} ObitHostFacet;

/* ---- DFT (ObitGPUSkyModel) ---- */
GPUInfo *ObitHostDFTCreate(void);
int ObitHostDFTInit(GPUInfo *g, const ObitHostUVDesc *uv, const ObitHostSkyModel *sm, int nVisPIO);
int ObitHostDFTSetMod(GPUInfo *g, const ObitHostSkyModel *sm, const float *comps, int lcomp, int ncomp);
int ObitHostDFTCalc(GPUInfo *g, float *buffer, long long nvis);
/* ObitHostDFTCalc, adding to times[0..2] the call's wall time, the time
 * inside ObitCUDASkyModelDFTCalc and the host copies' own time (copies of
 * one piece overlap the device work of the next, so times[0] can be less
 * than times[1] + times[2]). */
int ObitHostDFTCalcTimed(GPUInfo *g, float *buffer, long long nvis, double times[3]);
void ObitHostDFTShutdown(GPUInfo *g);

/* ---- gridding (ObitGPUGrid) ---- */
/* Spheroidal GCF table of ConvFunc(in,5): 7 cells x 100 points, as the
 * ObitUVGrid member convfn (src/ObitUVGrid.c:1352-1415); out[708]. */
void ObitHostConvFuncSph(float *convfn);
CUDAGridInfo *ObitHostGridCreate(int nfacet, int nGPU, const int *GPUs, const ObitHostUVDesc *uv,
                                 int nplane, int doBuff, int nVisPIO);
/* Common set-up (first call, facet 0) and per-facet set-up.  planeOfChan
 * (nchan*nif entries, may be NULL) gives each channel's plane; sigma1..3
 * may be NULL. */
int ObitHostGridSetFacet(CUDAGridInfo *c, int ifacet, const ObitHostFacet *f,
                         const ObitHostUVDesc *uv, const int *planeOfChan,
                         const float *sigma1, const float *sigma2, const float *sigma3);
int ObitHostGridBuffer(CUDAGridInfo *c, const float *buffer, long long nvis);
int ObitHostGridFlip(CUDAGridInfo *c);
/* ObitGPUGrid2Host: copy back, trim and centre; out[nplane][ny][nx/2+1][2]. */
int ObitHostGrid2Host(CUDAGridInfo *c, int ifacet, float *out);
void ObitHostGridShutdown(CUDAGridInfo *c);

/* Coarse-bin plane of every channel as ObitThreadGridSetupMF assigns it
 * (channels [bCh, eCh) of each bin, the last channel of each bin left out:
 * src/ObitThreadGrid.c:553-556) when cpuParity is 1, or every channel in
 * its bin (ObitGPUGrid.c:471-478) when 0.  bChan/eChan are 0-relative and
 * inclusive per bin; unassigned channels get nplane+100. */
void ObitHostPlanesFromBins(int nchanIF, int nbin, const int *bChan, const int *eChan,
                            int cpuParity, int *planeOfChan);

/* The geometry of one image field against the uv data that
 * ObitSkyModelLoadComps reads (ObitImageDesc / ObitUVDesc fields).  Images
 * are -SIN projections (or blank, which Obit reads as -SIN). */
typedef struct {
  double imCrval[2];     /* image reference RA, Dec (deg) */
  float imCrota;         /* image rotation (crota of the Dec axis, deg) */
  int do3D;              /* 1: 3-D image, 0: 2-D (tangent plane of the uv phase centre) */
  float xPxOff, yPxOff;  /* 2-D reference-pixel offsets from the tangent point (pixels) */
  double uvCrval[2];     /* uv phase centre RA, Dec (deg) */
  float uvCrota;         /* uv rotation (deg) */
  float imCrpix[2];      /* reference pixel (1-rel), 2-D fields */
  float imCdelt[2];      /* increments (deg), 2-D fields */
  int imNaxis[2];        /* pixels on each axis, 2-D fields */
  float xshift, yshift;  /* image shift (deg); nonzero: the centre pixel is the reference */
} ObitHostFieldGeom;

/* ObitSkyModelLoadComps for the point components of one 3-D image field
 * (src/ObitSkyModel.c:1889-1990 with ObitUVDescShiftPhase,
 * ObitSkyGeomShiftSIN and ObitUVDescShift3DMatrix, in the reference's
 * float/double arithmetic): cc rows [flux, deltaX, deltaY (deg), ...] of
 * ccStride floats -> comps rows [flux*factor, x, y, z] (Obit's 2 pi l,
 * 2 pi m, 2 pi (n-1)).  Returns ncc, or -1 for a 2-D field or bad input. */
int ObitHostLoadCompsPoint(const ObitHostFieldGeom *g, const float *cc, int ncc, int ccStride,
                           float factor, float *comps);

/* ObitTableCCUtilMergeSel (src/ObitTableCCUtil.c:1238-1370) on the rows of
 * one CC table held in memory: rows [Flux, DeltaX, DeltaY, (DeltaZ when
 * haveZ), parms[noParms]] (the table's columns), status[i] < 0 marks a
 * deselected row (status may be NULL), rows startComp..endComp (1-rel,
 * clipped to the table) are taken.  Rows at one position (DeltaX, DeltaY)
 * are merged into one entry with their fluxes summed (and their spectral
 * terms, for tabulated spectra, parms[3] in 20..29; parameterised spectra,
 * 10..19, are left unmerged), and the entries are ordered by decreasing
 * |flux| (ties in position order).  out: rows of ObitHostCCMergeWidth
 * floats [flux, X, Y, (Z), parms[4..noParms-1]]; parms: the first selected
 * row's parms[0..3] (maj, min, PA, type), zeros beyond noParms.  Returns
 * the number of merged rows, -1 on bad input. */
int ObitHostCCMergeWidth(int haveZ, int noParms);
int ObitHostCCMergeSel(const float *rows, int nrow, int haveZ, int noParms, const int *status,
                       int startComp, int endComp, float *out, float parms[4]);

/* The table-wide choices of ObitSkyModelLoadComps. */
typedef struct {
  int modType;     /* the table's model type, the highest of its fields: 0 point,
                      1 Gaussian, 10 point + spectrum, 11 Gaussian + spectrum */
  int nSpecTerm;   /* spectral terms per row (types 10, 11) */
  float factor;    /* flux factor */
  float minFlux, minDFT;  /* entries with |flux| at or below either are dropped */
  int noNeg;       /* stop at the first non-positive flux (Obit sets it only for Stokes I) */
} ObitHostLoadOpts;

/* ObitSkyModelLoadComps (src/ObitSkyModel.c:1889-2085) for one field:
 * the merged entries of ObitHostCCMergeSel (stride 3 + haveZ + nterms,
 * parms its model parameters) -> rows [flux*factor, x, y, z,
 * (g1, g2, g3 for a Gaussian table), (nSpecTerm spectral terms for types
 * 10, 11)], at most maxComp of them.  Gaussian terms are those of konst2
 * (FWHM in deg to the u^2, v^2, uv coefficients, :1722, 1939-1947); a
 * point field in a Gaussian table gets zero Gaussian terms, and in a
 * spectral table zero spectral terms.  3-D and 2-D fields (reference-pixel
 * offsets and the w term of the field centre's direction,
 * ObitUVDescShift3DMatrix's 2-D matrix).  Returns the number of rows, -1
 * for bad input or a combination of types the reference refuses. */
int ObitHostLoadComps(const ObitHostFieldGeom *g, const float *merged, int nmerged, int haveZ, int nterms,
                      const float parms[4], const ObitHostLoadOpts *o, float *comps, int maxComp);

/* GridCorrFn (src/ObitUVGrid.c:1612-1662): out[n] = the gridding
 * correction of an n-pixel axis centred on 1-relative pixel icent, for the
 * convolving-function table convfn (convWidth*convNperCell+1 entries). */
void ObitHostGridCorrFn(int n, int icent, const float *convfn, int convWidth, int convNperCell,
                        float *out);

#ifdef __cplusplus
}
#endif
#endif
