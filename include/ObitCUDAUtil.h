/*
 * ObitCUDAUtil.h -- device/stream/memory helpers with the C signatures of
 * Obit's CUDA utility layer (reference include/ObitCUDAUtil.h:42-152,
 * src/ObitCUDAUtil.cu:51-402), implemented over HIP.  Streams and events
 * travel through Obit's C code as int*; here they are hipStream_t and
 * hipEvent_t.  Errors abort with a message, as the reference does.
 */
#ifndef OBITCUDAUTIL_H
#define OBITCUDAUTIL_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AIPS magic blanking value, the float whose bytes are "INDE"
 * (reference src/ObitCUDAUtil.cu:388-400). */
float CUDAMagicF(void);
/* 1 if cuda_device is a valid device number, else -1 (reference .cu:73). */
long ObitCUDACheckGPU(int cuda_device);
/* Make cuda_device current; negative selects 0 (reference .cu:52). */
void ObitCUDASetGPU(int cuda_device);
/* Reference resets the device (.cu:86); here: synchronise only, so other
 * users of the device in the same process keep their state. */
void ObitCUDAResetGPU(void);
/* Synchronise the device (reference .cu:97; the argument is unused). */
void ObitCUDADeviceSynchronize(int *event);
int *ObitCUDAStreamCreate(void);                  /* reference .cu:172 */
void ObitCUDAStreamDestroy(int *stream);          /* reference .cu:184 */
int *ObitCUDAEventCreate(void);                   /* reference .cu:194 */
void ObitCUDAEventDestroy(int *event);            /* reference .cu:206 */
void ObitCUDAEventRecord(int *event, int *stream);/* reference .cu:217 */
void ObitCUDAEventSynchronize(int *event);        /* reference .cu:227 */
/* Pinned host memory (reference .cu:238, 250). */
float *ObitCUDAUtilAllocHost(int memsize);
void ObitCUDAUtilFreeHost(float *host);
/* Device memory (reference .cu:261, 273). */
float *ObitCUDAUtilAllocGPU(int memsize);
void ObitCUDAUtilFreeGPU(float *GPU);
/* Copies; a NULL stream means synchronous (reference .cu:286-352). */
void ObitCUDAUtilHost2GPU(float *GPU, float *host, int memsize, int *stream);
void ObitCUDAUtilHostAny2GPU(void *GPU, void *host, int memsize, int *stream);
void ObitCUDAUtilGPU2Host(float *host, float *GPU, int memsize, int *stream);
void ObitCUDAUtilGPUAny2Host(void *host, void *GPU, int memsize, int *stream);
/* Total device global memory in bytes (reference .cu:356). */
size_t ObitCUDAUtilMemory(int cuda_device);
/* Print whether ptr is host or device memory (reference .cu:372). */
void ObitCUDAUtilPointerType(void *ptr, char *label);

#ifdef __cplusplus
}
#endif
#endif /* OBITCUDAUTIL_H */
