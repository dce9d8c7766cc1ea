/*
 * ObitCUDASkyModelInfoDef.h -- descriptor structures shared between Obit's
 * C host class (ObitGPUSkyModel.c) and the DFT sky-model library.
 *
 * The field list, order and types are the ones Obit's host code fills in
 * (reference: include/ObitCUDASkyModelInfoDef.h:41-195); they are kept
 * identical so that an unmodified caller links against this library.
 * Device pointers inside the structs are produced by ObitCUDAUtilAllocGPU.
 * On this implementation the device never dereferences a struct: the
 * library reads the host copies (h_visInfo, h_modelInfo) and passes plain
 * arguments to its kernels.
 */
#ifndef OBITCUDASKYMODELINFODEF_H
#define OBITCUDASKYMODELINFODEF_H

#define STREAM_COUNT 4

/* Component model types handled by the DFT (reference :44-51). */
enum gpuModType {
  GPUModTypePoint = 0,     /* point */
  GPUModTypePointSpec,     /* point, parametrised spectrum exp(sum c_t ln^t) */
  GPUModTypePointTSpec,    /* point, tabulated coarse-bin spectrum */
  GPUModTypeGauss,         /* elliptical Gaussian */
  GPUModTypeGaussSpec,     /* Gaussian, parametrised spectrum */
  GPUModTypeGaussTSpec     /* Gaussian, tabulated spectrum */
};
typedef enum gpuModType GPUModType;

/* Operation on the data (reference :58-62). */
enum gpuOpType {
  GPUOpTypeSub = 0,        /* data -= model */
  GPUOpTypeDiv,            /* data /= model, weight *= |model| */
  GPUOpTypeRepl            /* data = model */
};
typedef enum gpuOpType GPUOpType;

/* Visibility layout (reference :66-123).  Increments are in correlations
 * (floats/3); ranges are 0-relative and inclusive. */
typedef struct {
  int nvis;          /* rows in the buffer */
  int nprod;         /* correlations per row */
  int nrparm;        /* random parameters per row */
  int lenvis;        /* floats per row */
  int nchan;         /* channels per IF */
  int ilocu, ilocv, ilocw, iloct, ilocb, ilocsu;
  int chanb, chane;  /* channel range */
  int incf;          /* channel increment (correlations) */
  int nIF, IFb, IFe; /* IF count and range */
  int incif;         /* IF increment (correlations) */
  int nstok, stokb, stoke;
  int incs;          /* Stokes increment (correlations) */
  int stype;         /* 0 IQUV, 1 RR/LL/RL/LR, 2 XX/YY/XY/YX */
  int kincif, kincf; /* increments in the frequency tables */
  float *freqRat;    /* host: nu_uvref/nu_ref per spectrum (per coarse bin) */
  float *d_freqRat;  /* device copy */
  float *freqScale;  /* device: nu/nu_uvref per (chan, IF), FP32 */
} GPUVisInfo;

/* Model description (reference :125-146). */
typedef struct {
  int nterm;          /* spectral terms */
  int size;           /* floats per component row */
  int doSwap;         /* multiply the model by sqrt(-1) */
  GPUModType type;
  GPUOpType opType;
  int *specIndex;     /* host: coarse bin per (chan, IF) */
  int *d_specIndex;   /* device copy */
  int nmodel[4];      /* components per Stokes */
  float *model[4];    /* device component tables per Stokes */
  float stokFact[4];  /* Stokes factors, see ObitCUDASkyModel.h */
} GPUModelInfo;

/* Top-level control block (reference :148-195). */
typedef struct {
  int nvis, nchan, nrparm, lenvis, modelSize, nstream, prtLv, oldnVisPIO;
  int cuda_device;
  int bufferSize;       /* floats in h_data */
  int nmodel[4];
  float *d_model[4];
  float *d_freq;        /* device FP32 frequency scaling */
  float *h_data;        /* pinned host buffer, nvis*lenvis floats */
  float **d_data;       /* per-stream device buffers (allocated by the host class) */
  GPUVisInfo *d_visInfo, *h_visInfo;
  GPUModelInfo *d_modelInfo, *h_modelInfo;
  int **stream;         /* hipStream_t per stream slot */
  int **cycleDone;      /* hipEvent_t per stream slot */
} GPUInfo;

#endif /* OBITCUDASKYMODELINFODEF_H */
