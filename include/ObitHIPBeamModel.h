/* ObitHIPBeamModel.h -- the beam-corrected GPU DFT behind Obit's own
 * entry points.
 *
 * ObitCUDABeamModelDFTInit / SetMod / Calc / Shutdown replace the
 * functions of the same names in reference include/ObitCUDABeamModel.h:
 * 49-86 (src/ObitCUDABeamModel.cu:1369-1437, 119-125).  They take the
 * GPUBeamInfo that Obit's host class src/ObitGPUBeamModel.c fills
 * (DFTInit :212-474, DFTSetMod :488-675, DFTSetBeam :686-735, DFTCalc
 * :745-836, DFTShutdown :849-896) and run this library's beam kernels
 * (ObitHIPBeam.h) on it, so that class links against this library
 * unchanged.
 *
 * The structures restate reference include/ObitCUDABeamModelInfoDef.h:
 * 61-273, include/CUDABeamInterpDef.h, include/CUDAFArrayDef.h and the
 * leading members of include/CUDAImageDescDef.h (through crval) member
 * for member: the host class allocates and fills them, this library only
 * reads them, so their layout is the interface.
 *
 * What the entry points read:
 *   h_visInfo  lenvis, nrparm, ilocu/v/w, nchan (channels x IFs), nIF,
 *              incf, incif, incs (3), nstok, kincf, kincif, chanb, chane,
 *              IFb, IFe, d_freqScale, d_freqRat;
 *   h_modelInfo type, nmodel, size, nSpec, stokType, opType, stokFact,
 *              d_model, d_specIndex;
 *   h_antInfo  numAnt, numAntType, numJChan, ilocb / iloca1 / iloca2,
 *              refSizeFact, d_BeamFreq, d_AntType, d_visJChann, d_Jones,
 *              d_BeamInterp;
 *   GPUBeamInfo nvis, lenvis, nstream, stream, d_data, h_data.
 *
 * Model rows (d_model, nmodel rows of size floats) are those
 * ObitGPUBeamModelDFTSetMod leaves for ObitSkyModelVMBeamMF's tabulated
 * spectra (src/ObitSkyModelVMBeamMF.c:1765-1776, 2112-2136;
 * src/ObitGPUBeamModel.c:574-600):
 *   GPUBeamModTypePointTSpec  [valid, field, dX, dY, x, y, z,
 *                              flux_0, SI_0, ..., flux_nSpec-1, SI_nSpec-1]
 *   GPUBeamModTypeGaussTSpec  [valid, field, dX, dY, x, y, z, g0, g1, g2,
 *                              flux_0, SI_0, ...]
 * with x, y, z the phase terms in radians per wavelength at the uv
 * reference frequency, dX, dY the offsets (deg) that place a component in
 * the beam images, g0..g2 the Gaussian's u^2, v^2, uv coefficients.  For
 * channel c of IF i (frequency table entry f = c kincf + i kincif, subband
 * s = d_specIndex[f]) a component's amplitude is
 *     flux_s (d_freqScale[f] d_freqRat[s])^SI_s   (0 where flux_s is 0)
 * as beamPointTSpecKernel forms it (src/ObitCUDABeamModel.cu:556-590).
 * The reference's Gaussian kernel (beamGaussTSpecKernel :1046-1212) takes
 * the phase from entries 1..3 and the Gaussian from 4..6 of the same row
 * layout; this library takes them from 4..6 and 7..9, where the table
 * holds them.  The other model types (Point, PointSpec, Gauss, GaussSpec)
 * are refused: their kernels read rows [amp, x, y, z, ...] that the
 * VMBeamMF loader does not produce.
 *
 * Calc applies the model to rows visRange[0]..visRange[1] (inclusive) of
 * h_data in place, in pieces of at most one d_data[] buffer's rows
 * ((nvis + 5) lenvis / nstream floats, as DFTInit allocates them) on the
 * caller's streams, and returns when the rows are back in h_data.  The
 * Jones table d_Jones ([type][component][beam channel] 2x2 complex
 * matrices, XX, XY, YX, YY) is
 *   doJones  formed from the beam images in d_BeamInterp for parAng
 *            (GetJonesKernel, :1228-1308),
 *   doUnit   set to the unit matrix (UnitJonesKernel, :1311-1334),
 *   neither  used as it stands,
 * and the model then evaluated as ObitHIPBeam.h states (FP64 phase, FP32
 * sums folded into FP64).  Channels and IFs are those of the record
 * (channel = product mod channels per IF), where the reference divides
 * the product number by the float increments incf / incif; both agree on
 * a single IF with every channel selected.  The frequency table must be
 * uniform within each IF (every entry within 1e-6 relative of the line
 * through the IF's first and last channel).
 *
 * Errors print a message naming the entry point and exit with
 * EXIT_FAILURE, as the reference's checkCudaErrors does.  Shutdown
 * releases this library's state for the GPUBeamInfo and waits for the device; it does not reset
 * the device (the reference's cudaDeviceReset frees every allocation of
 * the process).
 */
#ifndef OBIT_HIP_BEAM_MODEL_H
#define OBIT_HIP_BEAM_MODEL_H

#ifdef __cplusplus
extern "C" {
#endif

/* include/ObitCUDABeamModelInfoDef.h:61-80 */
enum gpuBeamModType {
  GPUBeamModTypePoint = 0,
  GPUBeamModTypePointSpec,
  GPUBeamModTypePointTSpec,
  GPUBeamModTypeGauss,
  GPUBeamModTypeGaussSpec,
  GPUBeamModTypeGaussTSpec
};
typedef enum gpuBeamModType GPUBeamModType;
enum gpuBeamOpType { GPUBeamOpTypeSub = 0, GPUBeamOpTypeDiv, GPUBeamOpTypeRepl };
typedef enum gpuBeamOpType GPUBeamOpType;

/* include/ObitCUDABeamModelInfoDef.h:83-142 */
typedef struct {
  int nvis, nprod, nrparm, lenvis, nchan;
  int ilocu, ilocv, ilocw, iloct, ilocb, ilocsu;
  int chanb, chane, incf, nIF, nSpec, IFb, IFe, incif;
  int nstok, stokb, stoke, incs, stype, kincif, kincf;
  float *h_freqRat, *d_freqRat;
  float *d_freqScale;
} GPUBeamVisInfo;

/* include/ObitCUDABeamModelInfoDef.h:144-179 */
typedef struct {
  int nterm, nSpec, size, doSwap;
  GPUBeamModType type;
  GPUBeamOpType opType;
  int stokType;
  int *h_specIndex, *d_specIndex;
  int nmodel;
  float *d_model;
  float stokFact[4];
} GPUBeamModelInfo;

/* include/CUDAFArrayDef.h */
typedef struct {
  float *h_array, *d_array, *d_FArray;
  float fblank;
  int ndim;
  int naxis[5];
} CUDAFArray;

/* include/CUDAImageDescDef.h, members through crval (the rest is never
 * read) */
typedef struct {
  long bitpix, naxis;
  long inaxes[7];
  char ctype[7][21];
  float cdelt[7], crpix[7], crota[7];
  float altCrpix, xshift, yshift, beamMaj, beamMin, beamPA;
  char object[41], teles[41], instrument[41], observer[41], obsdat[41], date[41], origin[41], bunit[41];
  float epoch, equinox;
  double obsra, obsdec;
  double crval[7];
} CUDAImageDesc;

/* include/CUDABeamInterpDef.h */
typedef struct {
  float *d_BeamInterp;
  int nx, ny;
  int hwidth;
  float denom[10];
  CUDAFArray *h_myArray, *d_myArray;
  CUDAImageDesc *d_myDesc;
} CUDABeamInterp;

/* include/ObitCUDABeamModelInfoDef.h:182-210; d_Jones holds the
 * reference's cudaMatx (4 x {real, imag} floats) */
typedef struct {
  int numAnt, numAntType, numJChan;
  int ilocb;
  int iloca1, iloca2, ilocsa;
  float s2p, c2p;
  float refantSize, refSizeFact;
  float *h_BeamFreq, *d_BeamFreq;
  int *h_AntType, *d_AntType;
  int *h_visJChann, *d_visJChann;
  float *d_Jones;
  CUDABeamInterp **d_BeamInterp;
} GPUBeamAntInfo;

/* include/ObitCUDABeamModelInfoDef.h:212-273 (stream and cycleDone hold
 * hipStream_t / hipEvent_t handles from ObitCUDAStreamCreate /
 * ObitCUDAEventCreate) */
typedef struct {
  int nvis, nchan, nrparm, lenvis, modelSize, nstream, prtLv, oldnVisPIO, cuda_device, bufferSize, nmodel;
  float timeRange[2];
  float delta_ParAng, curParAng;
  long visRange[2];
  int doUnit, haveUnit;
  float *d_model, *d_freq, *h_data;
  float **d_data;
  GPUBeamVisInfo *d_visInfo, *h_visInfo;
  GPUBeamModelInfo *d_modelInfo, *h_modelInfo;
  GPUBeamAntInfo *d_antInfo, *h_antInfo;
  int **stream;
  int **cycleDone;
} GPUBeamInfo;

/* Record the data layout (h_visInfo, h_antInfo, the frequency and
 * antenna tables) for this GPUBeamInfo (ObitCUDABeamModelDFTInit). */
void ObitCUDABeamModelDFTInit(GPUBeamInfo *gpuInfo);
/* Read the component rows and model description (h_modelInfo, d_model,
 * d_specIndex, d_freqRat) (ObitCUDABeamModelDFTSetMod). */
void ObitCUDABeamModelDFTSetMod(GPUBeamInfo *gpuInfo);
/* Apply the model to rows visRange[0]..visRange[1] of h_data
 * (ObitCUDABeamModelDFTCalc); parAng in radians. */
void ObitCUDABeamModelDFTCalc(GPUBeamInfo *gpuInfo, int doJones, int doUnit, float parAng, long visRange[2]);
/* Release this library's state for gpuInfo (ObitCUDABeamModelDFTShutdown). */
void ObitCUDABeamModelDFTShutdown(GPUBeamInfo *gpuInfo);

#ifdef __cplusplus
}
#endif

#endif /* OBIT_HIP_BEAM_MODEL_H */
