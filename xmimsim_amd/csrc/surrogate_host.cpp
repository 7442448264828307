/*
 * surrogate_host.cpp -- C ABI of the net-line surrogate: an ensemble of
 * multilayer perceptrons (GELU hidden layers, linear output) trained with
 * Adam on a masked mean-square loss, all on the GPU with the kernels of
 * surrogate_kernels.hip (BF16 operands on the matrix cores, FP32 master
 * weights, FP32 accumulation).
 *
 * No counterpart in the reference: the surrogate screens compositions for the
 * inverse quantification loop, and the Monte Carlo engine of this library is
 * both its training-data source and its judge.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <vector>

#include "../../include/xmimsim_amd.h"
#include "surrogate_kernels.h"

extern "C" void xmi_set_error(const char *fmt, ...);

namespace {

struct Buf {
  void *p = nullptr;
  size_t n = 0;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e == hipSuccess) {
      n = bytes;
      e = hipMemset(p, 0, bytes ? bytes : 16);
    }
    return e;
  }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
};

struct Layer {
  int K, N;     /* padded input and output widths */
  Buf W, Wb, WT, b, gW, gb, mW, vW, mb, vb;
  Buf Z, H, HT; /* hidden layers: pre-activation f32, activation bf16 and its transpose */
  Buf G, Gb, GT; /* gradient of the loss with respect to this layer's pre-activation */
};

}  // namespace

struct xmi_surrogate {
  int nets = 0, n_layers = 0, max_batch = 0, in_features = 0, out_features = 0, steps = 0;
  std::vector<Layer> L;
  Buf Xf, X0b, X0T, T, Mk, Y, partial, colsum;
  /* resident training data (xmi_surrogate_set_data) and the per-call
   * buffers of xmi_surrogate_train_resident */
  Buf Xall, Tall, Mall, idx, counts, losses;
  long n_data = 0;
  hipStream_t st = nullptr;
};

#define SG_TRY(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      xmi_set_error("surrogate: %s failed: %s", #x, hipGetErrorString(e_));              \
      return 0;                                                                          \
    }                                                                                    \
  } while (0)

static const int SG_LOSS_BLOCKS = 120;

static int round_up(int x, int m) { return (x + m - 1) / m * m; }

extern "C" {

xmi_surrogate *xmi_surrogate_create(int n_nets, int n_layers, const int *dims, int max_batch, uint64_t seed) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    xmi_set_error("surrogate: no HIP device");
    return nullptr;
  }
  if (n_nets < 1 || n_layers < 1 || max_batch < 64 || max_batch % 64 != 0) {
    xmi_set_error("surrogate: need n_nets >= 1, n_layers >= 1 and a batch that is a multiple of 64");
    return nullptr;
  }
  for (int l = 1; l <= n_layers; l++)
    if (dims[l] % 64 != 0) {
      xmi_set_error("surrogate: layer widths must be multiples of 64 (got %d)", dims[l]);
      return nullptr;
    }
  xmi_surrogate *s = new xmi_surrogate();
  s->nets = n_nets;
  s->n_layers = n_layers;
  s->max_batch = max_batch;
  s->in_features = dims[0];
  s->out_features = dims[n_layers];
  s->L.resize(n_layers);
  std::mt19937_64 rng(seed);
  const size_t B = (size_t)max_batch;
  bool ok = hipStreamCreate(&s->st) == hipSuccess;
  for (int l = 0; l < n_layers && ok; l++) {
    Layer &y = s->L[l];
    y.K = l == 0 ? round_up(dims[0], 64) : dims[l];
    y.N = dims[l + 1];
    const size_t nw = (size_t)n_nets * y.N * y.K, nb = (size_t)n_nets * y.N;
    ok = y.W.alloc(nw * 4) == hipSuccess && y.Wb.alloc(nw * 2) == hipSuccess && y.WT.alloc(nw * 2) == hipSuccess &&
         y.b.alloc(nb * 4) == hipSuccess && y.gW.alloc(nw * 4) == hipSuccess && y.gb.alloc(nb * 4) == hipSuccess &&
         y.mW.alloc(nw * 4) == hipSuccess && y.vW.alloc(nw * 4) == hipSuccess && y.mb.alloc(nb * 4) == hipSuccess &&
         y.vb.alloc(nb * 4) == hipSuccess;
    const size_t na = (size_t)n_nets * B * y.N;
    if (ok && l + 1 < n_layers)
      ok = y.Z.alloc(na * 4) == hipSuccess && y.H.alloc(na * 2) == hipSuccess && y.HT.alloc(na * 2) == hipSuccess;
    if (ok) ok = y.G.alloc(na * 4) == hipSuccess && y.Gb.alloc(na * 2) == hipSuccess && y.GT.alloc(na * 2) == hipSuccess;
    if (!ok) break;
    /* He-normal weights (zero on the padded input columns), zero biases */
    std::vector<float> w(nw, 0.0f);
    std::normal_distribution<float> nd(0.0f, sqrtf(2.0f / (float)(l == 0 ? dims[0] : y.K)));
    const int kreal = l == 0 ? dims[0] : y.K;
    for (int e = 0; e < n_nets; e++)
      for (int r = 0; r < y.N; r++)
        for (int c = 0; c < kreal; c++) w[((size_t)e * y.N + r) * y.K + c] = nd(rng);
    ok = hipMemcpy(y.W.p, w.data(), nw * 4, hipMemcpyHostToDevice) == hipSuccess;
  }
  if (ok) {
    const size_t K0 = (size_t)s->L[0].K, No = (size_t)s->out_features;
    ok = s->Xf.alloc(B * dims[0] * 4) == hipSuccess && s->X0b.alloc(B * K0 * 2) == hipSuccess &&
         s->X0T.alloc(B * K0 * 2) == hipSuccess && s->T.alloc(B * No * 4) == hipSuccess &&
         s->Mk.alloc(B * No * 4) == hipSuccess && s->Y.alloc((size_t)n_nets * B * No * 4) == hipSuccess &&
         s->partial.alloc((size_t)n_nets * SG_LOSS_BLOCKS * 4) == hipSuccess;
    size_t cs = 0;
    for (int l = 0; l < n_layers; l++) {
      const size_t c = sg_colsum_scratch(n_nets, max_batch, s->L[l].N);
      cs = c > cs ? c : cs;
    }
    if (ok) ok = s->colsum.alloc(cs * 4) == hipSuccess;
  }
  /* bf16 copies of the initial weights */
  for (int l = 0; l < n_layers && ok; l++) {
    Layer &y = s->L[l];
    ok = sg_launch_refresh((size_t)n_nets * y.N * y.K, y.N, y.K, y.W.as<float>(), y.Wb.as<uint16_t>(),
                           y.WT.as<uint16_t>(), s->st) == hipSuccess;
  }
  if (ok) ok = hipStreamSynchronize(s->st) == hipSuccess;
  if (!ok) {
    xmi_set_error("surrogate: device allocation or initialisation failed");
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
    return nullptr;
  }
  return s;
}

void xmi_surrogate_free(xmi_surrogate *s) {
  if (!s) return;
  if (s->st) {
    (void)hipStreamSynchronize(s->st);
    (void)hipStreamDestroy(s->st);
  }
  delete s;
}

long xmi_surrogate_param_count(const xmi_surrogate *s) {
  long n = 0;
  for (const Layer &y : s->L) n += (long)y.N * y.K + y.N;
  return n;
}

int xmi_surrogate_padded_width(const xmi_surrogate *s, int layer) {
  return (layer >= 0 && layer < s->n_layers) ? s->L[layer].K : 0;
}

}  // extern "C"

/* forward pass of every net on `batch` rows already packed in X0b/X0T */
static int forward(xmi_surrogate *s, int batch) {
  const size_t B = (size_t)batch;
  for (int l = 0; l < s->n_layers; l++) {
    Layer &y = s->L[l];
    const uint16_t *A;
    long lda, sa;
    if (l == 0) {
      A = s->X0b.as<uint16_t>();
      lda = y.K;
      sa = 0; /* every net reads the same inputs */
    } else {
      A = s->L[l - 1].H.as<uint16_t>();
      lda = y.K;
      sa = (long)(B * y.K);
    }
    SgEpiArgs e;
    memset(&e, 0, sizeof(e));
    e.bias = y.b.as<float>();
    e.s_bias = y.N;
    if (l + 1 < s->n_layers) {
      e.cf = y.Z.as<float>();
      e.s_cf = (long)(B * y.N);
      e.cb = y.H.as<uint16_t>();
      e.s_cb = (long)(B * y.N);
      e.ct = y.HT.as<uint16_t>();
      e.s_ct = (long)(B * y.N);
      SG_TRY(sg_launch_gemm(SG_EPI_HIDDEN, s->nets, batch, y.N, y.K, A, lda, sa, y.Wb.as<uint16_t>(), y.K,
                            (long)y.N * y.K, &e, s->st));
    } else {
      e.cf = s->Y.as<float>();
      e.s_cf = (long)(B * y.N);
      SG_TRY(sg_launch_gemm(SG_EPI_OUT, s->nets, batch, y.N, y.K, A, lda, sa, y.Wb.as<uint16_t>(), y.K,
                            (long)y.N * y.K, &e, s->st));
    }
  }
  return 1;
}

/* loss partials (into partial[nets][SG_LOSS_BLOCKS]) and the gradients of all
 * parameters into gW/gb, for targets and mask already in T/Mk; the gradient
 * scale is inv, or 1 / *count when count (device) is given.  Launches only. */
static int backward_launch(xmi_surrogate *s, int batch, float inv, const unsigned int *count, float *partial) {
  const size_t B = (size_t)batch, No = (size_t)s->out_features;
  Layer &out = s->L[s->n_layers - 1];
  SG_TRY(sg_launch_loss(batch, (int)No, s->nets, s->Y.as<float>(), s->T.as<float>(), s->Mk.as<float>(), inv, count,
                        out.G.as<float>(), out.Gb.as<uint16_t>(), out.GT.as<uint16_t>(), partial, SG_LOSS_BLOCKS,
                        s->st));
  for (int l = s->n_layers - 1; l >= 0; l--) {
    Layer &y = s->L[l];
    /* dW = G^T X: [N x B] . [K x B]^T */
    const uint16_t *XT = l == 0 ? s->X0T.as<uint16_t>() : s->L[l - 1].HT.as<uint16_t>();
    const long sxt = l == 0 ? 0 : (long)(B * y.K);
    SgEpiArgs e;
    memset(&e, 0, sizeof(e));
    e.cf = y.gW.as<float>();
    e.s_cf = (long)y.N * y.K;
    SG_TRY(sg_launch_gemm(SG_EPI_DW, s->nets, y.N, y.K, batch, y.GT.as<uint16_t>(), batch, (long)(B * y.N), XT, batch,
                          sxt, &e, s->st));
    SG_TRY(sg_launch_colsum(s->nets, batch, y.N, y.G.as<float>(), (long)(B * y.N), s->colsum.as<float>(),
                            y.gb.as<float>(), y.N, s->st));
    if (l > 0) {
      /* gradient at the previous pre-activation: (G W) * gelu'(Z_prev): [B x N] . [K x N]^T */
      Layer &p = s->L[l - 1];
      SgEpiArgs d;
      memset(&d, 0, sizeof(d));
      d.cf = p.G.as<float>();
      d.s_cf = (long)(B * p.N);
      d.cb = p.Gb.as<uint16_t>();
      d.s_cb = (long)(B * p.N);
      d.ct = p.GT.as<uint16_t>();
      d.s_ct = (long)(B * p.N);
      d.zprev = p.Z.as<float>();
      d.s_zprev = (long)(B * p.N);
      SG_TRY(sg_launch_gemm(SG_EPI_DX, s->nets, batch, y.K, y.N, y.Gb.as<uint16_t>(), y.N, (long)(B * y.N),
                            y.WT.as<uint16_t>(), y.N, (long)y.N * y.K, &d, s->st));
    }
  }
  return 1;
}

/* loss, gradients of all parameters into gW/gb; returns per-net loss in *loss */
static int backward(xmi_surrogate *s, int batch, const float *T, const float *mask, float *loss) {
  const size_t B = (size_t)batch, No = (size_t)s->out_features;
  SG_TRY(hipMemcpyAsync(s->T.p, T, B * No * 4, hipMemcpyHostToDevice, s->st));
  SG_TRY(hipMemcpyAsync(s->Mk.p, mask, B * No * 4, hipMemcpyHostToDevice, s->st));
  double count = 0.0;
  for (size_t k = 0; k < B * No; k++) count += mask[k] != 0.0f ? 1.0 : 0.0;
  const float inv = count > 0.0 ? (float)(1.0 / count) : 0.0f;
  if (!backward_launch(s, batch, inv, nullptr, s->partial.as<float>())) return 0;
  std::vector<float> part((size_t)s->nets * SG_LOSS_BLOCKS);
  SG_TRY(hipMemcpyAsync(part.data(), s->partial.p, part.size() * 4, hipMemcpyDeviceToHost, s->st));
  SG_TRY(hipStreamSynchronize(s->st));
  for (int e = 0; e < s->nets; e++) {
    double t = 0.0;
    for (int k = 0; k < SG_LOSS_BLOCKS; k++) t += part[(size_t)e * SG_LOSS_BLOCKS + k];
    if (loss) loss[e] = (float)(t * inv);
  }
  return 1;
}

static int pack_inputs(xmi_surrogate *s, const float *X, int batch) {
  SG_TRY(hipMemcpyAsync(s->Xf.p, X, (size_t)batch * s->in_features * 4, hipMemcpyHostToDevice, s->st));
  SG_TRY(sg_launch_pack(batch, s->in_features, s->L[0].K, s->Xf.as<float>(), s->X0b.as<uint16_t>(),
                        s->X0T.as<uint16_t>(), s->st));
  return 1;
}

static int check_batch(const xmi_surrogate *s, int batch) {
  if (batch < 64 || batch % 64 != 0 || batch > s->max_batch) {
    xmi_set_error("surrogate: batch %d must be a multiple of 64 and at most %d", batch, s->max_batch);
    return 0;
  }
  return 1;
}

/* the Adam update of every parameter after a backward pass */
static int adam_launch(xmi_surrogate *s, float lr) {
  s->steps++;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)s->steps), c2 = 1.0f - powf(b2, (float)s->steps);
  for (Layer &y : s->L) {
    const size_t nw = (size_t)s->nets * y.N * y.K, nb = (size_t)s->nets * y.N;
    SG_TRY(sg_launch_adam(nw, y.N, y.K, y.W.as<float>(), y.gW.as<float>(), y.mW.as<float>(), y.vW.as<float>(), lr, b1,
                          b2, eps, c1, c2, y.Wb.as<uint16_t>(), y.WT.as<uint16_t>(), s->st));
    SG_TRY(sg_launch_adam(nb, 1, 1, y.b.as<float>(), y.gb.as<float>(), y.mb.as<float>(), y.vb.as<float>(), lr, b1, b2,
                          eps, c1, c2, nullptr, nullptr, s->st));
  }
  return 1;
}

extern "C" {

/* one Adam step of every net on the same batch; loss[n_nets] receives the
 * masked mean-square loss before the step */
int xmi_surrogate_train_step(xmi_surrogate *s, const float *X, const float *T, const float *mask, int batch, float lr,
                             float *loss) {
  if (!s || !check_batch(s, batch)) return 0;
  if (!pack_inputs(s, X, batch) || !forward(s, batch) || !backward(s, batch, T, mask, loss)) return 0;
  if (!adam_launch(s, lr)) return 0;
  SG_TRY(hipStreamSynchronize(s->st));
  return 1;
}

int xmi_surrogate_set_data(xmi_surrogate *s, const float *X, const float *T, const float *mask, long n) {
  if (!s || !X || !T || !mask || n < 1) {
    xmi_set_error("surrogate: set_data needs data and n >= 1");
    return 0;
  }
  const size_t ni = (size_t)n * s->in_features * 4, no = (size_t)n * s->out_features * 4;
  s->n_data = 0;
  SG_TRY(s->Xall.alloc(ni));
  SG_TRY(s->Tall.alloc(no));
  SG_TRY(s->Mall.alloc(no));
  SG_TRY(hipMemcpy(s->Xall.p, X, ni, hipMemcpyHostToDevice));
  SG_TRY(hipMemcpy(s->Tall.p, T, no, hipMemcpyHostToDevice));
  SG_TRY(hipMemcpy(s->Mall.p, mask, no, hipMemcpyHostToDevice));
  s->n_data = n;
  return 1;
}

int xmi_surrogate_train_resident(xmi_surrogate *s, const int32_t *idx, int steps, int batch, const float *lr,
                                 float *loss) {
  if (!s || !idx || !lr || !loss || steps < 1 || !check_batch(s, batch)) return 0;
  if (s->n_data < 1) {
    xmi_set_error("surrogate: no resident training data (xmi_surrogate_set_data)");
    return 0;
  }
  for (size_t k = 0; k < (size_t)steps * batch; k++)
    if (idx[k] < 0 || idx[k] >= s->n_data) {
      xmi_set_error("surrogate: row index %d outside the %ld resident rows", idx[k], s->n_data);
      return 0;
    }
  const size_t ni = (size_t)steps * batch * 4, nl = (size_t)steps * s->nets * SG_LOSS_BLOCKS;
  if (s->idx.n < ni) SG_TRY(s->idx.alloc(ni));
  if (s->counts.n < (size_t)steps * 4) SG_TRY(s->counts.alloc((size_t)steps * 4));
  if (s->losses.n < nl * 4) SG_TRY(s->losses.alloc(nl * 4));
  SG_TRY(hipMemcpyAsync(s->idx.p, idx, ni, hipMemcpyHostToDevice, s->st));
  SG_TRY(hipMemsetAsync(s->counts.p, 0, (size_t)steps * 4, s->st));
  const int ld = s->L[0].K;
  for (int k = 0; k < steps; k++) {
    unsigned int *cnt = s->counts.as<unsigned int>() + k;
    SG_TRY(sg_launch_gather(batch, s->in_features, ld, s->out_features, s->idx.as<int>() + (size_t)k * batch,
                            s->Xall.as<float>(), s->Tall.as<float>(), s->Mall.as<float>(), s->X0b.as<uint16_t>(),
                            s->X0T.as<uint16_t>(), s->T.as<float>(), s->Mk.as<float>(), cnt, s->st));
    if (!forward(s, batch) ||
        !backward_launch(s, batch, 0.0f, cnt, s->losses.as<float>() + (size_t)k * s->nets * SG_LOSS_BLOCKS) ||
        !adam_launch(s, lr[k]))
      return 0;
  }
  std::vector<float> part(nl);
  std::vector<unsigned int> cnt(steps);
  SG_TRY(hipMemcpyAsync(part.data(), s->losses.p, nl * 4, hipMemcpyDeviceToHost, s->st));
  SG_TRY(hipMemcpyAsync(cnt.data(), s->counts.p, (size_t)steps * 4, hipMemcpyDeviceToHost, s->st));
  SG_TRY(hipStreamSynchronize(s->st));
  for (int k = 0; k < steps; k++) {
    const float inv = cnt[k] ? (float)(1.0 / (double)cnt[k]) : 0.0f;
    for (int e = 0; e < s->nets; e++) {
      double t = 0.0;
      for (int b = 0; b < SG_LOSS_BLOCKS; b++) t += part[((size_t)k * s->nets + e) * SG_LOSS_BLOCKS + b];
      loss[(size_t)k * s->nets + e] = (float)(t * inv);
    }
  }
  return 1;
}

/* loss and gradients without an update; grad[n_nets][param_count] in the
 * order of xmi_surrogate_get_params */
int xmi_surrogate_gradients(xmi_surrogate *s, const float *X, const float *T, const float *mask, int batch,
                            float *loss, float *grad) {
  if (!s || !check_batch(s, batch)) return 0;
  if (!pack_inputs(s, X, batch) || !forward(s, batch) || !backward(s, batch, T, mask, loss)) return 0;
  const long per = xmi_surrogate_param_count(s);
  for (int e = 0; e < s->nets; e++) {
    float *g = grad + (size_t)e * per;
    for (Layer &y : s->L) {
      SG_TRY(hipMemcpy(g, y.gW.as<float>() + (size_t)e * y.N * y.K, (size_t)y.N * y.K * 4, hipMemcpyDeviceToHost));
      g += (size_t)y.N * y.K;
      SG_TRY(hipMemcpy(g, y.gb.as<float>() + (size_t)e * y.N, (size_t)y.N * 4, hipMemcpyDeviceToHost));
      g += y.N;
    }
  }
  return 1;
}

/* outputs of every net: Y[n_nets][batch][out_features] */
int xmi_surrogate_predict(xmi_surrogate *s, const float *X, int batch, float *Y) {
  if (!s || !check_batch(s, batch)) return 0;
  if (!pack_inputs(s, X, batch) || !forward(s, batch)) return 0;
  const size_t No = (size_t)s->out_features, B = (size_t)batch;
  for (int e = 0; e < s->nets; e++)
    SG_TRY(hipMemcpyAsync(Y + (size_t)e * B * No, s->Y.as<float>() + (size_t)e * B * No, B * No * 4,
                          hipMemcpyDeviceToHost, s->st));
  SG_TRY(hipStreamSynchronize(s->st));
  return 1;
}

/* f32 master parameters, per net: for each layer W [N][K padded] then b [N] */
int xmi_surrogate_get_params(const xmi_surrogate *s, float *out) {
  const long per = xmi_surrogate_param_count(s);
  for (int e = 0; e < s->nets; e++) {
    float *o = out + (size_t)e * per;
    for (const Layer &y : s->L) {
      SG_TRY(hipMemcpy(o, y.W.as<float>() + (size_t)e * y.N * y.K, (size_t)y.N * y.K * 4, hipMemcpyDeviceToHost));
      o += (size_t)y.N * y.K;
      SG_TRY(hipMemcpy(o, y.b.as<float>() + (size_t)e * y.N, (size_t)y.N * 4, hipMemcpyDeviceToHost));
      o += y.N;
    }
  }
  return 1;
}

int xmi_surrogate_set_params(xmi_surrogate *s, const float *in) {
  const long per = xmi_surrogate_param_count(s);
  for (int e = 0; e < s->nets; e++) {
    const float *p = in + (size_t)e * per;
    for (Layer &y : s->L) {
      SG_TRY(hipMemcpy(y.W.as<float>() + (size_t)e * y.N * y.K, p, (size_t)y.N * y.K * 4, hipMemcpyHostToDevice));
      p += (size_t)y.N * y.K;
      SG_TRY(hipMemcpy(y.b.as<float>() + (size_t)e * y.N, p, (size_t)y.N * 4, hipMemcpyHostToDevice));
      p += y.N;
    }
  }
  /* refresh the bf16 copies */
  for (Layer &y : s->L)
    SG_TRY(sg_launch_refresh((size_t)s->nets * y.N * y.K, y.N, y.K, y.W.as<float>(), y.Wb.as<uint16_t>(),
                             y.WT.as<uint16_t>(), s->st));
  SG_TRY(hipStreamSynchronize(s->st));
  return 1;
}

}  // extern "C"
