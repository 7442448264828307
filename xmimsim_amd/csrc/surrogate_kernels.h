/*
 * surrogate_kernels.h -- launch interface between the surrogate's host code
 * (surrogate_host.cpp) and its gfx950 kernels (surrogate_kernels.hip).
 */
#ifndef XMIMSIM_AMD_SURROGATE_KERNELS_H
#define XMIMSIM_AMD_SURROGATE_KERNELS_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

/* epilogue kinds of sg_launch_gemm */
#define SG_EPI_HIDDEN 0 /* z = acc + b -> cf (f32), gelu(z) -> cb (bf16) and ct (bf16, transposed) */
#define SG_EPI_OUT 1    /* acc + b -> cf (f32) */
#define SG_EPI_DX 2     /* acc * gelu'(zprev) -> cf (f32), cb (bf16), ct (bf16, transposed) */
#define SG_EPI_DW 3     /* acc -> cf (f32) */

typedef struct SgEpiArgs {
  const float *bias;
  long s_bias;        /* per-net strides (elements) */
  float *cf;
  long s_cf;
  uint16_t *cb;
  long s_cb;
  uint16_t *ct;
  long s_ct;
  const float *zprev;
  long s_zprev;
} SgEpiArgs;

#ifdef __cplusplus
extern "C" {
#endif
/* C[M x N] = A[M x K] * B[N x K]^T per net (gridDim.z = nets); M, N and K
 * multiples of 64, lda/ldb multiples of 8 */
hipError_t sg_launch_gemm(int epi, int nets, int M, int N, int K, const uint16_t *A, long lda, long s_a,
                          const uint16_t *B, long ldb, long s_b, const SgEpiArgs *a, hipStream_t st);
hipError_t sg_launch_pack(int rows, int cols, int ld, const float *X, uint16_t *Xb, uint16_t *XbT, hipStream_t st);
/* inv_count scales the gradient; a non-NULL count (device) overrides it with 1 / *count */
hipError_t sg_launch_loss(int rows, int N, int nets, const float *Y, const float *T, const float *Mk, float inv_count,
                          const unsigned int *count, float *G, uint16_t *Gb, uint16_t *GT, float *partial, int blocks,
                          hipStream_t st);
/* column sums (bias gradient); part: sg_colsum_scratch(nets, rows, N) floats of scratch */
size_t sg_colsum_scratch(int nets, int rows, int N);
hipError_t sg_launch_colsum(int nets, int rows, int N, const float *G, long s_g, float *part, float *db, long s_db,
                            hipStream_t st);
/* mini-batch rows idx[batch] of the resident data into the packed inputs, targets, mask; count += mask entries */
hipError_t sg_launch_gather(int batch, int in_f, int ld, int out_f, const int *idx, const float *Xall, const float *Tall,
                            const float *Mall, uint16_t *Xb, uint16_t *XbT, float *T, float *Mk, unsigned int *count,
                            hipStream_t st);
hipError_t sg_launch_refresh(size_t n, int N, int K, const float *p, uint16_t *pb, uint16_t *pt, hipStream_t st);
hipError_t sg_launch_adam(size_t n, int N, int K, float *p, const float *g, float *m, float *v, float lr, float b1,
                          float b2, float eps, float c1, float c2, uint16_t *pb, uint16_t *pt, hipStream_t st);
#ifdef __cplusplus
}
#endif

#endif
