/*
 * ObitCUDAGridInfoDef.h -- descriptor structures shared between Obit's C
 * gridding host class (ObitGPUGrid.c) and the gridding library.
 *
 * Field list, order and types are those Obit's host code fills in
 * (reference: include/ObitCUDAGridInfoDef.h:42-176), kept identical so an
 * unmodified caller links against this library.  The library reads the
 * host copies (h_gridInfo, h_facetInfo) and never dereferences a struct on
 * the device.
 */
#ifndef OBITCUDAGRIDINFODEF_H
#define OBITCUDAGRIDINFODEF_H
#include <stddef.h>

#define GRID_STREAM_COUNT 4

/* Information common to all facets (reference :42-112). */
typedef struct {
  int nGPU;
  int nfacet;
  int nfacetPerGPU;
  int nplane;           /* planes per facet (coarse frequency bins) */
  int nchan;            /* channels x IFs (the host folds IFs in) */
  int nif;
  int nstok;
  int nrparm;
  int nvis;             /* rows currently in h_data_in */
  int lenvis;           /* floats per row */
  int convWidth;        /* 7 */
  int convNperCell;     /* 100 */
  float rotate;
  float guardu, guardv; /* guardband as a fraction of nx, ny */
  int oldnVisPIO;
  int nStream;
  int prtLv;
  int *GPU_device_no;
  int *h_freqPlane;     /* plane per channel; >= nplane means "skip" */
  int *d_freqPlane;
  float *h_sigma1, *h_sigma2, *h_sigma3;   /* taper per channel */
  float *d_sigma1, *d_sigma2, *d_sigma3;
  float *h_freqArr;     /* FP32 nu/nu_ref per channel */
  float *d_freqArr;
  float *h_convfn;      /* tabulated GCF, 100 points per cell */
  float *d_convfn;
  float *h_data_in;     /* pinned visibility buffer */
  float *d_data_in;
  float **h_d_data_in;
  int **stream;         /* hipStream_t slots per GPU */
  int **cycleDone;      /* hipEvent_t slots per GPU */
} GridInfo;

/* Per-facet information (reference :115-138). */
typedef struct {
  int doBeam;           /* grid (weight, 0) instead of data */
  int nx, ny;
  int GPU_num;
  float uscale, vscale, wscale;   /* wavelengths -> cells */
  float shift[3];       /* phase shift (dxc, dyc, dzc) */
  float rotUV[9];       /* 3-D uvw rotation */
  float maxBL, minBL;
  float bmTaper;
  int sizeGrid;         /* floats per plane: 2*(nx/2+1+hw)*ny + nx */
  float *h_grid;        /* pinned host copy, nplane*sizeGrid floats */
  float *d_grid;        /* device grid, nplane*sizeGrid floats */
} FacetInfo;

/* Top-level control block (reference :140-176). */
typedef struct {
  long nGPU;
  long nfacet;
  long nVisPIO;
  long *cuda_device;
  long *FacetGPU;       /* GPU index per facet */
  size_t *gpu_memory;
  void **d_base;
  void **h_base;
  GridInfo *h_gridInfo;
  GridInfo *d_gridInfo;
  FacetInfo **h_facetInfo;
  FacetInfo **d_facetInfo;
  FacetInfo ***h_d_facetInfo;
  GridInfo **h_d_gridInfo;
  float *d_data_in;
  float *h_data_in;
  float *h_grid_debug, *d_grid_debug;
} CUDAGridInfo;

#endif /* OBITCUDAGRIDINFODEF_H */
