/*
 * ObitCUDAGrid.h -- C entry points of the uv gridding library (HIP,
 * gfx950).  Each replaces the function of the same name in Obit's CUDA
 * layer; Obit's ObitGPUGrid.c calls them unchanged.
 *
 * Gridding follows Obit's CPU gridder (fast_prep_grid/fast_grid/ThreadFlip)
 * where the reference CUDA kernel disagrees with it: half-plane flip when
 * u <= 0; baseline limits on u^2+v^2; guard band tested in wavelengths;
 * GCF taps selected with the CPU's index arithmetic, so cell indices and
 * taps are identical to the CPU's; flip folds rows 1..ny/2.  Accumulation
 * is in FP32 in per-wave LDS tiles without atomics; a tile is added into
 * the grid once per work item, with plain stores when one item covers the
 * tile's whole track list and with device atomics when a long list is
 * split over several items.
 * Entry points abort with a message on a HIP error, as the reference does.
 */
#ifndef OBITCUDAGRID_H
#define OBITCUDAGRID_H
#include "ObitCUDAGridInfoDef.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces ObitCUDAGridAlloc (reference include/ObitCUDAGrid.h:43-45,
 * src/ObitCUDAGrid.cu:271-520): allocates the host-side descriptors the
 * caller fills (h_gridInfo and its arrays, h_facetInfo[], pinned vis
 * buffer when doBuff) and the per-GPU library state. */
void ObitCUDAGridAlloc(CUDAGridInfo *gpuInfo, long nfacet, long nchan, long nif,
                       long nVisPIO, long lenvis, long nplane, long iGPU, int doBuff);

/* Replaces ObitCUDAGridInit (reference :48, .cu:526-593): allocates and
 * zeroes the device grids of the facets assigned to iGPU and uploads the
 * common tables (GCF, frequencies, planes, tapers). */
void ObitCUDAGridInit(CUDAGridInfo *gpuInfo, long iGPU);

/* Replaces ObitCUDAGridGrid (reference :51, .cu:771-775): grids the
 * h_gridInfo->nvis rows of h_data_in into every facet on iGPU. */
void ObitCUDAGridGrid(CUDAGridInfo *in, long iGPU, int init);

/* Replaces ObitCUDAGridFlip (reference :54, .cu:779-806). */
void ObitCUDAGridFlip(CUDAGridInfo *in, long iGPU);

/* Replaces ObitCUDAGrid2CPU (reference :57, .cu:810-817): copies the facet
 * grid (all planes, padded layout) to h_facetInfo[ifacet]->h_grid. */
void ObitCUDAGrid2CPU(CUDAGridInfo *in, long ifacet);

/* Replaces ObitCUDAGridShutdown (reference :60, .cu:821-920): frees the
 * iGPU state, and the host descriptors on the last GPU; no device reset. */
void ObitCUDAGridShutdown(CUDAGridInfo *gpuinfo, long iGPU);

#ifdef __cplusplus
}
#endif
#endif /* OBITCUDAGRID_H */
