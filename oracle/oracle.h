/*
 * oracle.h -- CPU oracle for the visibility-domain hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (obit_amd/) links,
 * loads or calls this code.  It is used by tests/ (as the checker), by
 * __graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline
 * leg (as the timed CPU reference path).
 *
 * Two kinds of functions live here:
 *   1. Restatements of Obit's own CPU arithmetic (FP32 phase, cephes SSE
 *      sincos/exp, FP64 sums, FP32 gridder with its exact index arithmetic).
 *      Each function names the reference file:line it follows; paths are
 *      relative to ObitSystem/Obit/ of bill-cotton/Obit.
 *   2. FP64 oracles: the same mathematics evaluated in double precision from
 *      the same FP32 inputs, used to state absolute error of the GPU path.
 *
 * Pinning status of each function is stated in its comment ("pinned by").
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no -march: x86-64 SSE2
 * arithmetic without FMA, as Obit's stock "-O3 -msse" build).
 */
#ifndef OBIT_ORACLE_H
#define OBIT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Obit component model types (include/ObitSkyModel.h:131-163). */
enum {
  ORC_PointMod = 0, ORC_GaussMod = 1, ORC_CGaussMod = 2, ORC_USphereMod = 3,
  ORC_expDiskMod = 4, ORC_Unknown = 5,
  ORC_PointModSpec = 10, ORC_GaussModSpec = 11,
  ORC_PointModTSpec = 20, ORC_GaussModTSpec = 21
};

/* ---- FP32 vector math as a stock (HAVE_SSE) Obit build computes it ---- */
/* sincos_ps restated lane by lane (include/sse_mathfun.h:568-711). */
void orc_sincos_cephes(float x, float *s, float *c);
/* exp_ps restated lane by lane (include/sse_mathfun.h:214-289). */
float orc_exp_cephes(float x);
/* ObitSinCosVec (src/ObitSinCos.c:105-186): n==1 -> libm sinf/cosf. */
void orc_sincos_vec(int n, const float *ang, float *s, float *c);
/* ObitExpVec (src/ObitExp.c:56-134), SSE branch. */
void orc_exp_vec(int n, const float *arg, float *out);

/* ---- DFT descriptor: the ObitUVDesc / ObitSkyModel fields the path reads ---- */
typedef struct {
  int nrow;                     /* rows in buffer (numVisBuff) */
  int lrec, nrparm, ilocu, ilocv, ilocw;
  int nchan, nif, nstok;        /* inaxes of FREQ, IF, STOKES */
  int incs, incf, incif;        /* increments in floats (ObitUVDesc.c:380-392) */
  int kincf, kincif;            /* increments in freq tables */
  const float  *fscale;         /* [nchan*nif] FP32 nu/nu_ref (ObitUVDesc.c:606) */
  const double *freqArr;        /* [nchan*nif] Hz */
  int startChan, numChan, startIF, numIF, startPoln, numPoln;  /* 0-rel */
  int modType, nSpecTerm;
  int lcomp, ncomp;
  const float *comps;           /* [ncomp][lcomp] Obit comps table */
  double smRefFreq;             /* sky-model reference frequency (Hz) */
  int isMF;                     /* ObitSkyModelMF (tabulated spectra) */
  const int *specIndex;         /* [nchan*nif] coarse bin per channel (MF) */
  const double *specFreq;       /* [nSpec] coarse-bin frequencies (MF) */
  int doDivide, doReplace, doFlip;
  float stokFactor;
} orc_dft_desc;

/* ThreadSkyModelFTDFT (src/ObitSkyModel.c:2354-2857) over rows [row0,row1);
 * when isMF, ThreadSkyModelMFFTDFT (src/ObitSkyModelMF.c:1971-2440). */
void orc_dft_obit(const orc_dft_desc *d, float *buffer, int row0, int row1);
/* ObitSkyModelFTDFT decomposition (src/ObitSkyModel.c:2287-2331): contiguous
 * row ranges over nthreads (one thread below 1000 rows; MF: below 100). */
void orc_dft_obit_threaded(const orc_dft_desc *d, float *buffer, int nthreads);

/* FP64 model visibility per (row, IF, chan): out[(row*nif+if)*nchan+ch][2].
 * freqRatio[ifq] = nu/nu_uvref in FP64; specRatio[ifq] = nu/nu_ref(spectrum)
 * for parametric spectra (base class) or nu/specFreq[bin] (MF).  Channels or
 * IFs outside the selection are written as 0. */
void orc_dft_model_fp64(const orc_dft_desc *d, const float *buffer,
                        const double *freqRatio, double *out,
                        int row0, int row1, int nthreads);

/* ---- Gridding ---- */
/* ConvFunc(in,5) + sphfn (src/ObitUVGrid.c:1352-1415, 1440-1597):
 * convfn[708] as tabulated, convgfn[708] reordered for gridding. */
void orc_convfunc_sph(float *convfn, float *convgfn);
/* sphfn restated (src/ObitUVGrid.c:1440-1597) for ialf 1..5, im 4..8. */
float orc_sphfn(int ialf, int im, int iflag, float eta);

typedef struct {
  int nrow, lrec, nrparm, nchanIF;   /* nchanIF = nchan*nif */
  const float *freqArr;              /* FP32 fscale per channel */
  int convWidth, convNperCell;
  const float *convgfn;
  /* facet */
  int nx, ny;
  float uscale, vscale;
  float shift[3];
  float rot[9];
  float guardu, guardv;              /* guardband * nx, guardband * ny */
  float maxBL, minBL, bmTaper;
  int isBeam;
  int bChan, eChan;                  /* [bChan, eChan) */
  const float *sigma1, *sigma2, *sigma3;  /* MF taper per channel or NULL */
} orc_grid_desc;

/* fast_rot + fast_prep_grid (src/ObitThreadGrid.c:2045-2177, 1263-1387)
 * for one row: per channel ich in [bChan,eChan) writes
 *   iu[ich], iv[ich] (start row of the 7 rows), itu[ich], itv[ich] (GCF
 *   block index 0..100, i.e. cnvfn pointer = convgfn + convWidth*it),
 *   vis[2*ich], vis[2*ich+1] (weighted visibility).
 * Arrays are indexed by absolute channel, as in the reference. */
void orc_grid_prep_row(const orc_grid_desc *g, const float *buffer, int row,
                       int *iu, int *iv, int *itu, int *itv, float *vis);
/* Whole-facet prep for rows [row0,row1) into [row][nchanIF] arrays. */
void orc_grid_prep(const orc_grid_desc *g, const float *buffer, int row0, int row1,
                   int *iu, int *iv, int *itu, int *itv, float *vis);
/* ThreadGrid + fast_grid C path (src/ObitThreadGrid.c:960-991, 2002-2027):
 * accumulate rows [row0,row1) into grid (lrow = 2*(1+nx/2+hw), ny rows). */
void orc_grid_obit(const orc_grid_desc *g, const float *buffer, int row0, int row1,
                   float *grid);
/* Same sample indices, accumulated in FP64 (the value oracle). */
void orc_grid_fp64(const orc_grid_desc *g, const float *buffer, int row0, int row1,
                   double *grid);
/* ThreadFlip (src/ObitThreadGrid.c:998-1045), in place. */
void orc_grid_flip(float *grid, int nx, int ny, int convWidth);
void orc_grid_flip_fp64(double *grid, int nx, int ny, int convWidth);
/* ThreadMerge (src/ObitThreadGrid.c:1052-1105): add trimmed, v-swapped grid
 * into out[ny][nx/2+1][2]. */
void orc_grid_merge(const float *grid, int nx, int ny, int convWidth, float *out);
/* ObitThreadGridSetupBase/Grid/Flip/Merge decomposition for one facet:
 * nGpI private grids over row ranges, flipped then merged into out. */
void orc_grid_obit_threaded(const orc_grid_desc *g, const float *buffer,
                            int nGpI, int nthreads, float *out);

/* Floats du in (-0.5, 0.5] for which fmaf(N, du, -N/2) differs from the
 * CPU's (float)(N * ((double)du - 0.5)) (0 when the device's tap-fraction
 * form is exact). */
long orc_check_tap_fraction(float N, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
