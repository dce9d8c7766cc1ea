/* ObitHIPBeam.h -- Jones-corrected (antenna-beam-corrected) DFT sky model
 * for linear feeds on the GPU.
 *
 * Replaces the model evaluation of Obit's beam-corrected GPU DFT: the
 * point and Gaussian kernels of reference src/ObitCUDABeamModel.cu
 * (beamPointKernel :139-310, beamGaussKernel :687-858) behind entry points
 * that follow ObitCUDABeamModelDFTInit / SetMod / Calc / Shutdown
 * (:1368-1437; include/ObitCUDABeamModel.h), with the Jones table that
 * GetJonesKernel / UnitJonesKernel (:1227-1340) leave in d_Jones either
 * handed in by the caller, set to the unit matrix, or formed on the device
 * from beam images (ObitHIPBeamDFTSetBeams / ObitHIPBeamDFTJonesFromBeams).
 *
 * For every (row, IF, channel) the model is
 *     S = stokFact (.) sum_k a_k(nu) exp(2 pi i psi_k nu/nu0) J1_k P J2_k^H
 * where psi_k = u x_k + v y_k + w z_k (x, y, z in turns per wavelength at
 * nu0), a_k(nu) the amplitude (times exp(nu^2/nu0^2 (g0 u^2 + g1 v^2 + g2 uv))
 * for a Gaussian), J1_k, J2_k the 2x2 Jones matrices of the row's two
 * antenna types at component k and the channel's beam channel, and P the
 * coherency of the model's Stokes parameter for linear feeds at parallactic
 * angle chi (c = cos 2 chi, s = sin 2 chi):
 *     I: [[1, 0], [0, 1]]    Q: [[c, -s], [-s, -c]]
 *     U: [[s, c], [c, -s]]   V: [[0, i], [-i, 0]]
 * as Obit's CPU model builds it (src/ObitSkyModelVMBeamMF.c:2686-2757,
 * calcMod: beam1 * model * beam2^H).  The GPU kernels expand the same
 * products term by term (include/CUDALinCorSum.h); their Stokes V XX term
 * takes J2's YY element where the product (and their own comment) has
 * J2's XX element, and this library follows the product.
 *
 * Correlations are XX, YY, XY, YX at offsets 0, 3, 6, 9 floats of a
 * channel (the record's Stokes axis with increment 3).  The model is
 * applied as the CPU does (src/ObitSkyModelVMBeamMF.c:3711-3800): a channel
 * whose correlations are all flagged is skipped unless replacing; each
 * correlation is changed only where its weight is positive, or always when
 * replacing; divide acts on XX and YY only (the model is conjugated into
 * the data and the weight multiplied by |S|); the cross hands are
 * subtracted or replaced only with 4 correlations.  Rows whose antenna
 * numbers fall outside the antenna-type table are left unchanged.
 *
 * Numerics are those of the split DFT (ObitHIPVis.h): FP64 phase and range
 * reduction, FP32 sincos and channel recurrence, FP32 sums folded into FP64
 * after every component tile.
 */
#ifndef OBIT_HIP_BEAM_H
#define OBIT_HIP_BEAM_H

#ifdef __cplusplus
extern "C" {
#endif

/* Operation: subtract, divide, replace (GPUBeamOpType of the reference,
 * include/ObitCUDABeamModelInfoDef.h). */
enum { ObitHIPBeamOpSub = 0, ObitHIPBeamOpDiv = 1, ObitHIPBeamOpRepl = 2 };

typedef struct {
  int lrec, nrparm;          /* floats per row, random parameters */
  int ilocu, ilocv, ilocw;   /* u, v, w (wavelengths at nu0) */
  int ilocb;                 /* BASELINE = 256 ant1 + ant2 (1-rel), or < 0 ... */
  int iloca1, iloca2;        /* ... ANTENNA1 / ANTENNA2 (1-rel) instead
                                (the reference's GetAnts, ObitCUDABeamModel.cu:82-98) */
  int nif, nchan;            /* IFs and channels of the record */
  int incf, incif;           /* channel and IF increments (floats) */
  int ncorr;                 /* 2 (XX, YY) or 4 (XX, YY, XY, YX) */
  int ncomp;                 /* components */
  int compLen;               /* 4 (point: amp, x, y, z) or 7 (Gaussian: + g0, g1, g2) */
  int ntype;                 /* antenna types (1..16) */
  int nant;                  /* entries of the antenna-type table */
  int njchan;                /* beam (Jones) channels */
  int stokType;              /* model Stokes: 1 I, 2 Q, 3 U, 4 V */
  int opType;                /* ObitHIPBeamOp* */
  float stokFact[4];         /* factor per correlation XX, XY, YX, YY
                                (the reference's modelInfo->stokFact) */
  int chanb, chane;          /* channels (0-rel, inclusive) the model is applied to,
                                and IFs; outside them the data are left as they are */
  int IFb, IFe;              /* (visInfo->chanb/chane/IFb/IFe, ObitGPUBeamModel.c:324-333).
                                chane < 0 / IFe < 0 select through the last channel / IF;
                                a zero-initialised descriptor selects channel 0 of IF 0
                                only, so set them (-1 for all) */
} ObitHIPBeamDesc;

typedef struct ObitHIPBeam ObitHIPBeam;

/* Set up for one data layout (ObitCUDABeamModelDFTInit).  Host arrays:
 *   fIF, dfIF  [nif]         nu/nu0 of channel 0 of each IF and per channel;
 *   antType    [nant]        type (0-rel) of 0-rel antenna i;
 *   visJChan   [nif*nchan]   beam channel (0-rel) of (IF, channel),
 *                            IF-major (the reference's FreqPlane map).
 * The Jones table starts as the unit matrix.  stream: hipStream_t or NULL. */
int ObitHIPBeamDFTCreate(ObitHIPBeam **out, const ObitHIPBeamDesc *desc, const double *fIF,
                         const double *dfIF, const int *antType, const int *visJChan,
                         void *stream);
/* Load the component table (ObitCUDABeamModelDFTSetMod): ncomp rows of
 * compLen floats, x, y, z in radians per wavelength at nu0 (2 pi times
 * turns), as Obit's model table holds them. */
int ObitHIPBeamDFTSetMod(ObitHIPBeam *h, const float *comps);
/* Load the Jones table (what GetJonesKernel leaves in d_Jones):
 * [ntype][ncomp][njchan] 2x2 complex matrices of 8 floats each, (XX re,
 * XX im, XY re, XY im, YX re, YX im, YY re, YY im); NULL loads the unit
 * matrix (UnitJonesKernel). */
int ObitHIPBeamDFTSetJones(ObitHIPBeam *h, const float *jones);
/* Apply the model to nrow resident rows at d_vis for parallactic angle
 * parAng (radians) (ObitCUDABeamModelDFTCalc).  Asynchronous on the
 * handle's stream. */
int ObitHIPBeamDFTCalc(ObitHIPBeam *h, float *d_vis, long long nrow, float parAng);
/* Per-channel amplitudes [ncomp][nif*nchan] (IF-major) that replace the
 * table's amplitude column, for the spectral component types: the
 * reference's beamPointSpec / beamGaussSpec kernels form
 * a exp(-sum_t c_t (-ln(nu/nu_ref))^(t+1)) and beamPointTSpec /
 * beamGaussTSpec flux_sb (nu/nu_sb)^SI_sb per subband
 * (ObitCUDABeamModel.cu:312-505, 506-686, 860-1226); they depend on the
 * component and the channel only, so they are formed once, in FP64, by
 * the caller.  A Gaussian's factor still multiplies them.  NULL returns to
 * the amplitude column. */
int ObitHIPBeamDFTSetAmps(ObitHIPBeam *h, const float *amp);
/* Copy the current Jones table to the host ([ntype][ncomp][njchan][8]). */
int ObitHIPBeamDFTGetJones(ObitHIPBeam *h, float *jones);

/* Beam images of one antenna type (the reference's eight CUDABeamInterp
 * objects per type, src/ObitGPUBeamModel.c:689-725): axes nx x ny x nchan
 * (nchan = njchan), the geometry of its image descriptor, and the
 * Lagrange half width of the interpolation (the reference uses 2). */
typedef struct {
  int nx, ny, nchan;
  int proj;          /* 0 linear, 1 -SIN, 2 -TAN, 3 -ARC (others refused) */
  int hwidth;        /* 1..4 */
  double crval[2];   /* reference position (deg) */
  float crpix[2];    /* reference pixel (1-rel) */
  float cdelt[2];    /* increments (deg) */
} ObitHIPBeamImageDesc;

/* Load the beam images: images [ntype][8][nchan][ny][nx] floats, per type
 * in the reference's order P1 re, P1 im, P2 re, P2 im, X1 re, X1 im,
 * X2 re, X2 im (RX, RXIm, LY, LYIm, RL, RLIm, LR, LRIm); beamFreq [njchan]
 * the beam planes' frequencies (Hz); refSizeFact the primary-beam factor
 * of the reference's InvBeam (antenna size x fudge x deg->rad / c);
 * fblank the blanking value of the images. */
int ObitHIPBeamDFTSetBeams(ObitHIPBeam *h, const ObitHIPBeamImageDesc *desc, const float *images,
                           const float *beamFreq, float refSizeFact, float fblank);
/* Form the Jones table from the beam images for the components at pos
 * [ncomp][2] (RA and Dec offsets, deg, as Obit's beam-model table holds
 * them) and parallactic angle parAng (radians), as GetJonesKernel does
 * (ObitCUDABeamModel.cu:1227-1308): the position (-RA, Dec) is placed in
 * each type's beam image rotated by parAng; each element is the
 * interpolated image value times 1/sqrt of the reference beam gain at the
 * offset and beam frequency; a position off the image, or a reference
 * gain or P1 beyond the reference's limits, gives the unit matrix.
 * Asynchronous on the handle's stream. */
int ObitHIPBeamDFTJonesFromBeams(ObitHIPBeam *h, const float *pos, float parAng);

/* The Set* calls and JonesFromBeams first wait for the handle's stream,
 * so tables are never replaced under a running Calc. */

/* Release the handle (ObitCUDABeamModelDFTShutdown). */
int ObitHIPBeamDFTDestroy(ObitHIPBeam *h);

#ifdef __cplusplus
}
#endif

#endif /* OBIT_HIP_BEAM_H */
