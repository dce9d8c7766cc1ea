/*
 * ObitCUDASkyModel.h -- C entry points of the DFT sky-model library
 * (HIP, gfx950).  Each one replaces the function of the same name in
 * Obit's CUDA layer; Obit's ObitGPUSkyModel.c calls them unchanged.
 *
 * Semantics follow Obit's CPU DFT (ThreadSkyModelFTDFT) where the
 * reference CUDA kernels disagree with it:
 *   - spectral terms use exp(+sum_t c_t ln^(t+1)(nu/nu0));
 *   - doSwap multiplies the model by sqrt(-1) for every Stokes;
 *   - a flagged correlation (weight <= 0) is left alone, except that
 *     GPUOpTypeRepl replaces it (and a channel with every Stokes flagged is
 *     still replaced);
 *   - the model applied to Stokes stokb is unscaled; after each Stokes that
 *     was written the running scale is multiplied by stokFact[1]; the
 *     component table already carries Obit's 'factor'.  A caller should
 *     therefore set stokFact[0]=1 and stokFact[1]=stokFactor.
 * All entry points abort with a message on a HIP error, as the reference's
 * checkCudaErrors does; the ObitHIP* functions in ObitHIPVis.h return a
 * status instead.
 */
#ifndef OBITCUDASKYMODEL_H
#define OBITCUDASKYMODEL_H
#include "ObitCUDASkyModelInfoDef.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces ObitCUDASkyModelDFTInit (reference include/ObitCUDASkyModel.h:44,
 * src/ObitCUDASkyModel.cu:727-731).  Creates the library's per-GPUInfo
 * state (stream, tuning) on gpuInfo->cuda_device. */
void ObitCUDASkyModelDFTInit(GPUInfo *gpuInfo);

/* Replaces ObitCUDASkyModelDFTSetMod (reference :47, .cu:739-744).  Reads
 * h_modelInfo and the device table model[0] (nmodel[0] rows of size floats)
 * and builds the resident component layout the kernels use. */
void ObitCUDASkyModelDFTSetMod(GPUInfo *gpuInfo);

/* Replaces ObitCUDASkyModelDFTCalc (reference :50, .cu:751-765).  Applies
 * the model to gpuInfo->nvis rows of h_data in place (upload, one kernel
 * launch, download). */
void ObitCUDASkyModelDFTCalc(GPUInfo *gpuInfo);

/* Replaces ObitCUDASkyModelDFTShutdown (reference :53, .cu:774-783).
 * Synchronises and frees the library's state; the device is not reset. */
void ObitCUDASkyModelDFTShutdown(GPUInfo *gpuInfo);

#ifdef __cplusplus
}
#endif
#endif /* OBITCUDASKYMODEL_H */
