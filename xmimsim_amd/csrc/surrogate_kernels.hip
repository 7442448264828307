/*
 * surrogate_kernels.hip -- gfx950 kernels of the net-line surrogate: a
 * multilayer perceptron (log composition/thickness/density/energy -> log net
 * line intensities) trained on spectra of the GPU Monte Carlo.
 *
 * Every product is one kernel shape, C[M x N] = A[M x K] * B[N x K]^T with both
 * operands K-contiguous BF16 and FP32 accumulation on the matrix cores
 * (v_mfma_f32_16x16x32_bf16).  The forward pass, the input gradient and the
 * weight gradient are all of this form because each activation and gradient is
 * stored both row-major and transposed by the epilogue that produces it.
 *
 * Tiles: 64 x 64 per workgroup of four wave64s, K in steps of 32 through
 * double-buffered LDS (k_sg_gemm_nt).  The ensemble members run
 * side by side as gridDim.z.  Sizes must be multiples of 64 (M, N and
 * K); the host checks this.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "surrogate_kernels.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define SG_TK 32
#define SG_LDS_STRIDE (SG_TK + 8) /* bf16 per LDS row: 80 B, staggers the banks */

/* bf16 <-> f32 with round-to-nearest-even, written out so that the host and the
 * device round identically (NaN is not produced by these kernels) */
__device__ __forceinline__ uint16_t sg_f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float sg_bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float sg_gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float sg_gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

struct SgEpi {
  const float *bias;    /* [N] (+ net stride) */
  long s_bias;
  float *cf;            /* f32 result [M x N] */
  long s_cf;
  uint16_t *cb;         /* bf16 result [M x N] */
  long s_cb;
  uint16_t *ct;         /* bf16 transposed result [N x M] */
  long s_ct;
  const float *zprev;   /* SG_EPI_DX: pre-activation [M x N] */
  long s_zprev;
};

/* one tile of C = A * B^T: BM x BN per workgroup of four wave64s (2 x 2
 * waves of BM/2 x BN/2, each a grid of 16 x 16 MFMA blocks), K in steps of
 * SG_TK through two LDS buffers: the next step's operands are loaded from
 * global memory into registers while the matrix cores work on the current
 * ones, and stored into the other buffer afterwards, so one barrier per step
 * suffices.  Epilogue per EPI; the transposed bf16 result is written four
 * rows (8 bytes) at a time. */
template <int EPI, int BM, int BN>
__global__ void __launch_bounds__(256) k_sg_gemm_nt(int M, int N, int K, const uint16_t *A, long lda, long s_a,
                                                     const uint16_t *B, long ldb, long s_b, SgEpi e) {
  constexpr int WM = BM / 2, WN = BN / 2, FI = WM / 16, FJ = WN / 16;
  constexpr int LA = BM * SG_TK / 8 / 256, LB = BN * SG_TK / 8 / 256; /* 16-byte loads per thread */
  __shared__ __attribute__((aligned(16))) uint16_t As[2][BM * SG_LDS_STRIDE];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN * SG_LDS_STRIDE];
  const int net = blockIdx.z;
  A += (size_t)net * s_a;
  B += (size_t)net * s_b;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * WM, wc = (wave & 1) * WN;
  f32x4 acc[FI][FJ];
#pragma unroll
  for (int i = 0; i < FI; i++)
#pragma unroll
    for (int j = 0; j < FJ; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  /* global -> LDS: load q of a thread moves 8 bf16 of row (tid + 256 q) / RL
   * at k ((tid + 256 q) % RL) * 8, RL = SG_TK / 8 loads per row */
  constexpr int RL = SG_TK / 8;
  uint4 ra[LA], rb[LB];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < LA; q++) {
      const int t = tid + 256 * q;
      ra[q] = *(const uint4 *)(A + (size_t)(m0 + t / RL) * lda + k0 + (t % RL) * 8);
    }
#pragma unroll
    for (int q = 0; q < LB; q++) {
      const int t = tid + 256 * q;
      rb[q] = *(const uint4 *)(B + (size_t)(n0 + t / RL) * ldb + k0 + (t % RL) * 8);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < LA; q++) {
      const int t = tid + 256 * q;
      *(uint4 *)(&As[buf][(t / RL) * SG_LDS_STRIDE + (t % RL) * 8]) = ra[q];
    }
#pragma unroll
    for (int q = 0; q < LB; q++) {
      const int t = tid + 256 * q;
      *(uint4 *)(&Bs[buf][(t / RL) * SG_LDS_STRIDE + (t % RL) * 8]) = rb[q];
    }
  };
  const int nk = K / SG_TK;
  gload(0);
  lstore(0);
  __syncthreads();
  /* operands: lane l holds row (l % 16) and the 8 consecutive k of group l / 16;
   * A and B use the same k assignment, so the contraction is complete */
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  for (int kt = 0; kt < nk; kt++) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * SG_TK);
#pragma unroll
    for (int kk = 0; kk < SG_TK; kk += 32) {
      bf16x8 a[FI], b[FJ];
#pragma unroll
      for (int i = 0; i < FI; i++) a[i] = *(const bf16x8 *)(&As[cur][(wr + i * 16 + fr) * SG_LDS_STRIDE + kk + fk]);
#pragma unroll
      for (int j = 0; j < FJ; j++) b[j] = *(const bf16x8 *)(&Bs[cur][(wc + j * 16 + fr) * SG_LDS_STRIDE + kk + fk]);
#pragma unroll
      for (int i = 0; i < FI; i++)
#pragma unroll
        for (int j = 0; j < FJ; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) lstore(cur ^ 1);
    __syncthreads();
  }
  /* result layout of a 16 x 16 block: lane l holds column l % 16, rows 4 (l / 16) + v */
#pragma unroll
  for (int i = 0; i < FI; i++)
#pragma unroll
    for (int j = 0; j < FJ; j++) {
      const int row0 = m0 + wr + i * 16 + (lane >> 4) * 4;
      const int col = n0 + wc + j * 16 + (lane & 15);
      uint16_t hv[4];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = row0 + v;
        const size_t o = (size_t)row * N + col;
        const float x = acc[i][j][v];
        if (EPI == SG_EPI_HIDDEN) {
          const float z = x + e.bias[(size_t)net * e.s_bias + col];
          e.cf[(size_t)net * e.s_cf + o] = z;
          hv[v] = sg_f2bf(sg_gelu(z));
          e.cb[(size_t)net * e.s_cb + o] = hv[v];
        } else if (EPI == SG_EPI_OUT) {
          e.cf[(size_t)net * e.s_cf + o] = x + e.bias[(size_t)net * e.s_bias + col];
        } else if (EPI == SG_EPI_DX) {
          const float g = x * sg_gelu_grad(e.zprev[(size_t)net * e.s_zprev + o]);
          e.cf[(size_t)net * e.s_cf + o] = g;
          hv[v] = sg_f2bf(g);
          e.cb[(size_t)net * e.s_cb + o] = hv[v];
        } else {
          e.cf[(size_t)net * e.s_cf + o] = x;
        }
      }
      if (EPI == SG_EPI_HIDDEN || EPI == SG_EPI_DX) {
        const uint2 pk = make_uint2((uint32_t)hv[0] | ((uint32_t)hv[1] << 16), (uint32_t)hv[2] | ((uint32_t)hv[3] << 16));
        *(uint2 *)(&e.ct[(size_t)net * e.s_ct + (size_t)col * M + row0]) = pk;
      }
    }
}

/* f32 [rows x cols] -> bf16 row-major [rows x ld] (zero padded) and its
 * transpose [ld x rows]; one thread per padded element */
__global__ void k_sg_pack(int rows, int cols, int ld, const float *X, uint16_t *Xb, uint16_t *XbT) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)rows * ld) return;
  const int r = (int)(idx / ld), c = (int)(idx % ld);
  const uint16_t h = c < cols ? sg_f2bf(X[(size_t)r * cols + c]) : (uint16_t)0;
  Xb[idx] = h;
  if (XbT) XbT[(size_t)c * rows + r] = h;
}

/* masked mean-square loss of the output layer and its gradient:
 * dY = 2 (Y - T) mask / count, stored f32, bf16 and transposed bf16;
 * loss partials per (net, row block) */
__global__ void __launch_bounds__(256) k_sg_loss(int rows, int N, int nets, const float *Y, const float *T,
                                                  const float *Mk, float inv_count, const unsigned int *count, float *G,
                                                  uint16_t *Gb, uint16_t *GT, float *partial) {
  const int net = blockIdx.y;
  /* with a device count (mask entries of the batch), the same rounding as the host's (float)(1.0 / count) */
  if (count) inv_count = *count ? (float)(1.0 / (double)*count) : 0.0f;
  const size_t base = (size_t)net * rows * N;
  __shared__ uint16_t tile[64][66];
  float s = 0.0f;
  /* 64 x 64 tiles (rows and N are multiples of 64), this block's in turn:
   * row-major results straight out, the transpose through LDS */
  const int tc = N / 64, ntiles = (rows / 64) * tc;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int r0 = (t / tc) * 64, c0 = (t % tc) * 64;
    for (int k = threadIdx.x; k < 64 * 64; k += 256) {
      const int rr = k >> 6, cc = k & 63;
      const size_t idx = (size_t)(r0 + rr) * N + c0 + cc;
      const float d = (Y[base + idx] - T[idx]) * Mk[idx];
      s += d * d;
      const float g = 2.0f * d * inv_count;
      G[base + idx] = g;
      const uint16_t gb = sg_f2bf(g);
      Gb[base + idx] = gb;
      tile[rr][cc] = gb;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 64 * 64; k += 256) {
      const int cc = k >> 6, rr = k & 63;
      GT[base + (size_t)(c0 + cc) * rows + r0 + rr] = tile[rr][cc];
    }
    __syncthreads();
  }
  /* fixed-order block sum */
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)net * gridDim.x + blockIdx.x] = red[0];
}

/* column sums of G [rows x N] per net, the bias gradient, in a fixed order:
 * workgroup (column tile of 64, net, chunk of SG_CS_ROWS rows) sums its
 * chunk with 4 threads per column and an LDS step into part[net][chunk][N];
 * k_sg_colsum2 adds the chunks in order */
#define SG_CS_ROWS 256
__global__ void __launch_bounds__(256) k_sg_colsum(int rows, int N, const float *G, long s_g, float *part) {
  __shared__ float red[4][64];
  const int net = blockIdx.y, chunk = blockIdx.z;
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), t = threadIdx.x >> 6;
  const float *g = G + (size_t)net * s_g;
  float s = 0.0f;
  const int r1 = min(rows, (chunk + 1) * SG_CS_ROWS);
  for (int r = chunk * SG_CS_ROWS + t; r < r1; r += 4) s += g[(size_t)r * N + c];
  red[t][threadIdx.x & 63] = s;
  __syncthreads();
  if (t == 0)
    part[((size_t)net * gridDim.z + chunk) * N + c] = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                                                      (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

__global__ void k_sg_colsum2(int N, int chunks, const float *part, float *db, long s_db) {
  const int net = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.0f;
  for (int k = 0; k < chunks; k++) s += part[((size_t)net * chunks + k) * N + c];
  db[(size_t)net * s_db + c] = s;
}

/* the rows idx[0..batch) of the resident inputs packed to bf16, row-major
 * (zero padded to ld) and transposed, through a 64 x 64 LDS tile so that
 * both stores are coalesced */
__global__ void __launch_bounds__(256) k_sg_gather_x(int batch, int in_f, int ld, const int *idx, const float *Xall,
                                                      uint16_t *Xb, uint16_t *XbT) {
  __shared__ uint16_t tile[64][66];
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  for (int k = threadIdx.x; k < 64 * 64; k += 256) {
    const int rr = k >> 6, cc = k & 63, r = r0 + rr, c = c0 + cc;
    uint16_t h = 0;
    if (r < batch && c < in_f) h = sg_f2bf(Xall[(size_t)idx[r] * in_f + c]);
    if (r < batch && c < ld) Xb[(size_t)r * ld + c] = h;
    tile[rr][cc] = h;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 64 * 64; k += 256) {
    const int cc = k >> 6, rr = k & 63, r = r0 + rr, c = c0 + cc;
    if (r < batch && c < ld) XbT[(size_t)c * batch + r] = tile[rr][cc];
  }
}

/* targets and mask of rows idx[0..batch) in f32, and the count of mask
 * entries of the batch (wave-aggregated) */
__global__ void k_sg_gather(int batch, int out_f, const int *idx, const float *Tall, const float *Mall, float *T,
                            float *Mk, unsigned int *count) {
  const size_t nt = (size_t)batch * out_f;
  unsigned int on = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nt; i += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(i / out_f), c = (int)(i % out_f);
    const size_t src = (size_t)idx[r] * out_f + c;
    T[i] = Tall[src];
    const float m = Mall[src];
    Mk[i] = m;
    on += m != 0.0f;
  }
  /* one atomic per workgroup: thousands on one address would serialise */
  __shared__ unsigned int red[4];
  unsigned int w = on;
  for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off);
  if (__lane_id() == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(count, red[0] + red[1] + red[2] + red[3]);
}

/* Adam (Kingma & Ba 2015) on f32 master parameters; refreshes the bf16 copy
 * (row-major [N x K]) and its transpose [K x N] when given */
__global__ void k_sg_adam(size_t n, int N, int K, float *p, const float *g, float *m, float *v, float lr, float b1,
                          float b2, float eps, float c1, float c2, uint16_t *pb, uint16_t *pt) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (pb) {
    /* n = nets * N * K; element (net, r, c) */
    const size_t per = (size_t)N * K;
    const size_t net = i / per, rem = i % per;
    const int r = (int)(rem / K), c = (int)(rem % K);
    const uint16_t h = sg_f2bf(pi);
    pb[i] = h;
    pt[net * per + (size_t)c * N + r] = h;
  }
}

/* bf16 copy (row-major [N x K] per net) and its transpose from f32 parameters */
__global__ void k_sg_refresh(size_t n, int N, int K, const float *p, uint16_t *pb, uint16_t *pt) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t per = (size_t)N * K;
  const size_t net = i / per, rem = i % per;
  const int r = (int)(rem / K), c = (int)(rem % K);
  const uint16_t h = sg_f2bf(p[i]);
  pb[i] = h;
  pt[net * per + (size_t)c * N + r] = h;
}

/* ------------------------------------------------------------------------ */

extern "C" {

hipError_t sg_launch_refresh(size_t n, int N, int K, const float *p, uint16_t *pb, uint16_t *pt, hipStream_t st) {
  hipLaunchKernelGGL(k_sg_refresh, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, N, K, p, pb, pt);
  return hipGetLastError();
}

hipError_t sg_launch_gemm(int epi, int nets, int M, int N, int K, const uint16_t *A, long lda, long s_a,
                          const uint16_t *B, long ldb, long s_b, const SgEpiArgs *a, hipStream_t st) {
  SgEpi e;
  e.bias = a->bias;
  e.s_bias = a->s_bias;
  e.cf = a->cf;
  e.s_cf = a->s_cf;
  e.cb = a->cb;
  e.s_cb = a->s_cb;
  e.ct = a->ct;
  e.s_ct = a->s_ct;
  e.zprev = a->zprev;
  e.s_zprev = a->s_zprev;
  /* 64 x 64 tiles and K steps of 32: at these sizes measured faster than
   * 128 x 128, 128 x 64 and 64 x 128 tiles and than K steps of 64 (whose
   * 144-byte LDS rows conflict) */
#define SG_GEMM(E)                                                                                         \
  hipLaunchKernelGGL((k_sg_gemm_nt<E, 64, 64>), dim3(N / 64, M / 64, nets), dim3(256), 0, st, M, N, K, A, lda, \
                     s_a, B, ldb, s_b, e)
  switch (epi) {
    case SG_EPI_HIDDEN: SG_GEMM(SG_EPI_HIDDEN); break;
    case SG_EPI_OUT: SG_GEMM(SG_EPI_OUT); break;
    case SG_EPI_DX: SG_GEMM(SG_EPI_DX); break;
    default: SG_GEMM(SG_EPI_DW); break;
  }
#undef SG_GEMM
  return hipGetLastError();
}

hipError_t sg_launch_pack(int rows, int cols, int ld, const float *X, uint16_t *Xb, uint16_t *XbT, hipStream_t st) {
  size_t n = (size_t)rows * ld;
  hipLaunchKernelGGL(k_sg_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rows, cols, ld, X, Xb, XbT);
  return hipGetLastError();
}

hipError_t sg_launch_loss(int rows, int N, int nets, const float *Y, const float *T, const float *Mk, float inv_count,
                          const unsigned int *count, float *G, uint16_t *Gb, uint16_t *GT, float *partial, int blocks,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_sg_loss, dim3(blocks, nets), dim3(256), 0, st, rows, N, nets, Y, T, Mk, inv_count, count, G, Gb,
                     GT, partial);
  return hipGetLastError();
}

size_t sg_colsum_scratch(int nets, int rows, int N) {
  return (size_t)nets * ((rows + SG_CS_ROWS - 1) / SG_CS_ROWS) * N;
}

hipError_t sg_launch_colsum(int nets, int rows, int N, const float *G, long s_g, float *part, float *db, long s_db,
                            hipStream_t st) {
  const int chunks = (rows + SG_CS_ROWS - 1) / SG_CS_ROWS;
  hipLaunchKernelGGL(k_sg_colsum, dim3(N / 64, nets, chunks), dim3(256), 0, st, rows, N, G, s_g, part);
  hipLaunchKernelGGL(k_sg_colsum2, dim3((N + 255) / 256, nets), dim3(256), 0, st, N, chunks, part, db, s_db);
  return hipGetLastError();
}

hipError_t sg_launch_gather(int batch, int in_f, int ld, int out_f, const int *idx, const float *Xall, const float *Tall,
                            const float *Mall, uint16_t *Xb, uint16_t *XbT, float *T, float *Mk, unsigned int *count,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_sg_gather_x, dim3((batch + 63) / 64, (ld + 63) / 64), dim3(256), 0, st, batch, in_f, ld, idx,
                     Xall, Xb, XbT);
  hipLaunchKernelGGL(k_sg_gather, dim3(512), dim3(256), 0, st, batch, out_f, idx, Tall, Mall, T, Mk, count);
  return hipGetLastError();
}

hipError_t sg_launch_adam(size_t n, int N, int K, float *p, const float *g, float *m, float *v, float lr, float b1,
                          float b2, float eps, float c1, float c2, uint16_t *pb, uint16_t *pt, hipStream_t st) {
  hipLaunchKernelGGL(k_sg_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, N, K, p, g, m, v, lr, b1, b2,
                     eps, c1, c2, pb, pt);
  return hipGetLastError();
}

} /* extern "C" */
