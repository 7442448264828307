/*
 * detector.hip -- detector response of the simulated spectra on the GPU:
 * the reference's xmi_detector_convolute_all (src/xmi_detector_f.F90:219-580).
 *
 * Per interaction order: window/absorber transmission and crystal absorption
 * per channel (src/xmi_detector_f.F90:404-426), escape peaks from given escape
 * ratios (:582-809), pile-up (:56-217), the Gaussian response with the
 * Si(Li)/SDD low-energy tail (:446-530) and Poisson counting noise (:862-905);
 * the net-line histories get the detector absorption at their line energies
 * (:275-335).
 *
 * Kernels: k_det_norm, one thread per source channel, sums its response row
 * (the normalisation); k_det_conv, one thread per target channel, sums the
 * contributions of the source channels in increasing order, the order in
 * which the reference accumulates them.  k_det_pileup runs independent pulse
 * trains, each on its own Philox stream; k_det_poisson draws one Poisson
 * variate per channel (inversion for small means, Hormann's PTRS rejection
 * otherwise).
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/xmimsim_amd.h"
#include "philox.h"
#include "xmi_internal.h"
#include "xrl_model.h"

extern "C" void xmi_set_error(const char *fmt, ...);

#define DET_PHX_SUB 5u /* Philox substream of the detector response */

struct DetResp {
  double zero, gain, a, b, c;
  int nch, type;
};

/* response of channel I to a unit line in channel I0 (before normalisation) */
__device__ __forceinline__ double det_r(const DetResp &d, int I0, int I) {
  const double E0 = d.zero + d.gain * I0;
  const double FWHM = sqrt(d.a + d.b * E0);
  const double B0 = d.c * FWHM;
  const double A0 = 1.0 / (B0 * 1.7724538509055160273);
  const double E = d.zero + d.gain * I;
  const double X = (E - E0) / B0;
  const double G = exp(-X * X);
  if (E0 > 50.0) return A0 * G;
  const double A3 = 2.73E-3 * exp(-0.21 * E0) + 1.E-4;
  const double A4 = 0.000188 * exp(-0.00296 * pow(E0, 0.763)) + 1.355E-5 * exp(0.968 * pow(E0, 0.498));
  const double ALFA = 1.179 * exp(8.6E-4 * pow(E0, 1.877)) - 7.793 * exp(-3.81 * pow(E0, -0.0716));
  const double F = erfc(X);
  const double tail = d.type == XMI_DETECTOR_CONVOLUTION_PROFILE_SI_SDD ? 0.63 : 2.7;
  return A0 * G + 1.0 * (tail * A3 + 15.0 * A4 * exp(ALFA * (E - E0))) * F;
}

__global__ void k_det_norm(DetResp d, double *norm) {
  const int I0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (I0 >= d.nch) return;
  const double E0 = d.zero + d.gain * I0;
  double s = 0.0;
  if (E0 >= 1.0) {
    const int last = I0 + 100 < d.nch - 1 ? I0 + 100 : d.nch - 1;
    for (int I = 0; I <= last; I++) s += det_r(d, I0, I);
  }
  norm[I0] = s;
}

/* conv[I] = sum over I0 >= I - 100 of R(I0, I) * src[I0] / norm[I0] */
__global__ void k_det_conv(DetResp d, const double *src, const double *norm, double *conv) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= d.nch) return;
  double s = 0.0;
  for (int I0 = I - 100 > 0 ? I - 100 : 0; I0 < d.nch; I0++) {
    if (d.zero + d.gain * I0 < 1.0 || src[I0] == 0.0) continue;
    s += det_r(d, I0, I) * src[I0] / norm[I0];
  }
  conv[I] = s;
}

/* Pile-up, src/xmi_detector_f.F90:56-217: trains of pulses whose successive
 * gaps (exponential, mean live_time / N) stay below the pulse width sum into
 * one count.  Trains are independent, so each thread runs its own until it
 * has drawn n_each pulses (and finishes the train it is in).  As in the
 * reference, a pulse's energy is taken at (channel + 1) * gain + zero and the
 * sum lands in channel (sum - zero) / gain - 1.  cdf: cumulative channel
 * contents normalised to 1. */
__global__ void k_det_pileup(int nch, const double *cdf, double mean_gap, double width, double zero, double gain,
                             long n_each, long n_total, uint64_t seed, unsigned long long *hist,
                             unsigned int *overflow) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long first = t * n_each;
  if (first >= n_total) return;
  const long mine = (first + n_each <= n_total) ? n_each : n_total - first;
  phx_stream s;
  phx_stream_init(&s, seed, 0u, 0u, DET_PHX_SUB, (uint64_t)t);
  long done = 0;
  int np = 0, first_ch = 0;
  double esum = 0.0;
  while (done < mine || np > 0) {
    const double u = phx_uniform(&s);
    int lo = 0, hi = nch - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cdf[mid] > u) hi = mid;
      else lo = mid + 1;
    }
    const double gap = -mean_gap * log1p(-phx_uniform(&s));
    np++;
    done++;
    if (np == 1) first_ch = lo;
    esum += (double)(lo + 1) * gain + zero;
    if (np > 100) { /* the reference stops with an error; here the train is dropped and counted */
      atomicAdd(overflow, 1u);
      np = 0;
      esum = 0.0;
      continue;
    }
    if (gap > width) {
      if (np == 1) {
        atomicAdd(&hist[first_ch], 1ull);
      } else {
        const long ps = (long)((esum - zero) / gain);
        if (ps > 0 && ps <= nch) atomicAdd(&hist[ps - 1], 1ull);
      }
      np = 0;
      esum = 0.0;
    }
  }
}

/* Poisson variate of mean lam (lam > 1): inversion below 10, otherwise the
 * transformed rejection with squeeze of W. Hormann, "The transformed
 * rejection method for generating Poisson random variables", Insurance:
 * Mathematics and Economics 12 (1993) 39-45. */
__device__ double det_poisson(double lam, phx_stream *s) {
  if (lam < 10.0) {
    double p = exp(-lam), F = p, u = phx_uniform(s);
    double k = 0.0;
    while (u > F && k < 1000.0) {
      k += 1.0;
      p *= lam / k;
      F += p;
    }
    return k;
  }
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (int it = 0; it < 10000; it++) {
    const double U = phx_uniform(s) - 0.5;
    const double V = phx_uniform(s);
    const double us = 0.5 - fabs(U);
    const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
  }
  return floor(lam + 0.5);
}

__global__ void k_det_poisson(int nch, double *ch, uint64_t seed, uint32_t order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nch) return;
  const double v = ch[i];
  if (v > 4294967295.0 || !(v > 1.0)) return; /* src/xmi_detector_f.F90:891-897 */
  phx_stream s;
  phx_stream_init(&s, seed, 1u, order, DET_PHX_SUB, (uint64_t)i);
  ch[i] = det_poisson(v, &s);
}

/* ------------------------------------------------------------------------ */
/* host */

static double layer_mu(const xmi_layerF *L, double E) {
  double rv = 0.0;
  for (int i = 0; i < L->n_elements; i++) rv += xrl_cs_total(xrl_atom_get(L->Z[i]), E) * L->weight[i];
  return rv;
}

/* detector window/absorber transmission times crystal absorption at E */
static double det_efficiency(const xmi_inputF *in, double E) {
  double c = 1.0;
  for (int l = 0; l < in->n_det_layers; l++)
    c *= exp(-1.0 * in->det_layers[l].density * in->det_layers[l].thickness * layer_mu(&in->det_layers[l], E));
  for (int l = 0; l < in->n_crystal_layers; l++)
    c = -1.0 * c * expm1(-1.0 * in->crystal_layers[l].density * in->crystal_layers[l].thickness *
                         layer_mu(&in->crystal_layers[l], E));
  return c;
}

/* 0-based index p with a[p] < x <= a[p+1] (x == a[0] gives 0), else -1:
 * the findpos of src/xmi_aux_f.F90 used by the escape-peak code */
static int det_findpos(const double *a, int n, double x) {
  if (fabs(x - a[0]) < 1e-10) return 0;
  for (int i = 0; i < n; i++)
    if (x <= a[i]) return i - 1;
  return -1;
}

/* escape peaks with given escape ratios, src/xmi_detector_f.F90:582-809.
 * fluo_escape_ratios holds the Fortran array (element, line 1..109, input
 * energy) in column-major order; compton_escape_ratios (input, output). */
static void det_escape(std::vector<double> &ch, const xmi_inputF *in, const xmi_escape_ratios *er) {
  const int nch = (int)ch.size();
  const double zero = in->detector.zero, gain = in->detector.gain;
  const double channel_1e = 0.5 * gain + zero;
  const int ne = er->n_elements, nfe = er->n_fluo_input_energies;
  const int nce = er->n_compton_input_energies, nco = er->n_compton_output_energies;
  const double *fe = er->fluo_escape_input_energies;
  const double *ci = er->compton_escape_input_energies, *co = er->compton_escape_output_energies;
  const double compton_out_diff = co[1] - co[0];
  /* lines of the K, L1, L2 and L3 vacancies that the reference loops over */
  static const int shell_first[4] = {1, 32, 60, 86}, shell_last[4] = {29, 58, 85, 109};
  for (int i = 0; i < nch; i++) {
    const double channel_e = ((double)i + 0.5) * gain + zero;
    double sum_ratio = 0.0;
    for (int j = 0; j < ne; j++) {
      const xrl_atom *a = xrl_atom_get(er->Z[j]);
      for (int sh = 0; sh < 4; sh++) {
        if (!(channel_e > a->edge[sh])) continue;
        for (int line = shell_first[sh]; line <= shell_last[sh]; line++) {
          const double line_e = a->line_energy[line];
          if (!((channel_e - line_e) >= channel_1e && channel_e >= fe[0] && channel_e < fe[nfe - 1])) continue;
          const int pos = det_findpos(fe, nfe, channel_e);
          if (pos < 0) continue;
          const double r0 = er->fluo_escape_ratios[((size_t)pos * 109 + (line - 1)) * ne + j];
          const double r1 = er->fluo_escape_ratios[((size_t)(pos + 1) * 109 + (line - 1)) * ne + j];
          /* interpolate_simple, src/xmi_aux_f.F90:1267-1276 */
          const double ratio = r0 + (r1 - r0) * (channel_e - fe[pos]) / (fe[pos + 1] - fe[pos]);
          sum_ratio += ratio;
          const int escape_i = (int)((channel_e - line_e - zero) / gain);
          if (escape_i >= 0 && escape_i < nch) ch[escape_i] += ratio * ch[i];
        }
      }
    }
    for (int j = 0; j < i; j++) {
      const double channel_c = ((double)j + 0.5) * gain + zero;
      const double compton_diff = channel_e - channel_c;
      if (!(channel_e >= ci[0] && channel_e < ci[nce - 1] && compton_diff >= co[0] && compton_diff < co[nco - 1]))
        continue;
      const int p1 = det_findpos(ci, nce, channel_e), p2 = det_findpos(co, nco, compton_diff);
      if (p1 < 0 || p2 < 0) continue;
      const double denom = (ci[p1 + 1] - ci[p1]) * (co[p2 + 1] - co[p2]);
      const double c1 = (ci[p1 + 1] - channel_e) * (co[p2 + 1] - compton_diff) / denom;
      const double c2 = (channel_e - ci[p1]) * (co[p2 + 1] - compton_diff) / denom;
      const double c3 = (ci[p1 + 1] - channel_e) * (compton_diff - co[p2]) / denom;
      const double c4 = (channel_e - ci[p1]) * (compton_diff - co[p2]) / denom;
      const double *g = er->compton_escape_ratios;
      const double v = c1 * g[(size_t)p2 * nce + p1] + c2 * g[(size_t)p2 * nce + p1 + 1] +
                       c3 * g[(size_t)(p2 + 1) * nce + p1] + c4 * g[(size_t)(p2 + 1) * nce + p1 + 1];
      const double ratio = v * gain / compton_out_diff;
      ch[j] += ratio * ch[i];
      sum_ratio += ratio;
    }
    ch[i] *= (1.0 - sum_ratio);
  }
}

#define DET_TRY(x)                                                          \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      xmi_set_error("detector response: %s: %s", #x, hipGetErrorString(e_)); \
      return 0;                                                             \
    }                                                                       \
  } while (0)

extern "C" uint64_t xmi_engine_current_seed(const xmi_main_options *options);

/* clears n words (a kernel rather than a memset, so that a caller may
 * capture the launch sequence in a graph) */
__global__ void k_det_clear(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

/* device buffers of one convolution, nch entries each (over: one word) */
struct DetBufs {
  double *src, *norm, *conv, *cdf;
  unsigned long long *hist, *over;
};

/* the response of every order into channels_conv[n] (nch doubles each,
 * owned by the caller), src/xmi_detector_f.F90:219-580; 1 on success */
static int det_convolute(const xmi_inputF *in, double **channels_noconv, double **channels_conv,
                         double *brute_history, double *var_red_history, const xmi_main_options *options,
                         const xmi_escape_ratios *escape_ratios, int n_interactions_all, int zero_interaction,
                         const DetBufs &b, hipStream_t st) {
  const int nch = in->detector.nchannels;
  const int n_int = in->general.n_interactions_trajectory;
  DetResp d;
  d.zero = in->detector.zero;
  d.gain = in->detector.gain;
  d.a = in->detector.noise * in->detector.noise;
  /* the reference writes (2.3548)**2 * 3.85 with default-real literals */
  d.b = (double)((2.3548f * 2.3548f) * 3.85f) * in->detector.fano / 1000.0;
  d.c = sqrt(2.0) / (2.0 * sqrt(2.0 * log(2.0)));
  d.nch = nch;
  d.type = (int)in->detector.detector_type;
  const uint64_t seed = xmi_engine_current_seed(options);
  std::vector<double> eff(nch);
  for (int i = 0; i < nch; i++) eff[i] = det_efficiency(in, i * d.gain + d.zero);
  const int nb = (nch + 255) / 256;
  hipLaunchKernelGGL(k_det_norm, dim3(nb), dim3(256), 0, st, d, b.norm);
  DET_TRY(hipGetLastError());
  for (int n = zero_interaction ? 0 : 1; n <= n_interactions_all; n++) {
    std::vector<double> temp(channels_noconv[n], channels_noconv[n] + nch);
    for (int i = 0; i < nch; i++) temp[i] *= eff[i];
    if (options->use_escape_peaks && escape_ratios) det_escape(temp, in, escape_ratios);
    if (options->use_sum_peaks) {
      double Nt = 0.0;
      for (int i = 0; i < nch; i++) Nt += temp[i];
      const long Nt_long = (long)Nt;
      if (Nt_long > 0) {
        std::vector<double> cdf(nch);
        double acc = 0.0;
        for (int i = 0; i < nch; i++) {
          acc += temp[i];
          cdf[i] = acc / Nt;
        }
        cdf[nch - 1] = 1.0;
        DET_TRY(hipMemcpyAsync(b.cdf, cdf.data(), nch * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_det_clear, dim3(nb), dim3(256), 0, st, b.hist, nch);
        hipLaunchKernelGGL(k_det_clear, dim3(1), dim3(64), 0, st, b.over, 1);
        const long threads = 65536;
        const long each = (Nt_long + threads - 1) / threads;
        hipLaunchKernelGGL(k_det_pileup, dim3((unsigned)(threads / 256)), dim3(256), 0, st, nch, b.cdf,
                           in->detector.live_time / Nt, in->detector.pulse_width, d.zero, d.gain, each, Nt_long,
                           seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(n + 1)), b.hist, (unsigned int *)b.over);
        DET_TRY(hipGetLastError());
        std::vector<unsigned long long> h(nch);
        unsigned int over = 0;
        DET_TRY(hipMemcpyAsync(h.data(), b.hist, nch * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        DET_TRY(hipMemcpyAsync(&over, b.over, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        DET_TRY(hipStreamSynchronize(st));
        if (over) { /* src/xmi_detector_f.F90:152-156 stops the program here */
          xmi_set_error("detector response: pulse train maximum (100 pulses) reached %u times; adjust the pulse "
                        "width or disable the pile-up generation", over);
          return 0;
        }
        for (int i = 0; i < nch; i++) temp[i] = (double)h[i];
      }
    }
    DET_TRY(hipMemcpyAsync(b.src, temp.data(), nch * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_det_conv, dim3(nb), dim3(256), 0, st, d, b.src, b.norm, b.conv);
    DET_TRY(hipGetLastError());
    if (options->use_poisson) {
      hipLaunchKernelGGL(k_det_poisson, dim3(nb), dim3(256), 0, st, nch, b.conv, seed, (uint32_t)n);
      DET_TRY(hipGetLastError());
    }
    DET_TRY(hipMemcpyAsync(channels_conv[n], b.conv, nch * sizeof(double), hipMemcpyDeviceToHost, st));
    DET_TRY(hipStreamSynchronize(st));
  }
  if (!zero_interaction) memset(channels_conv[0], 0, nch * sizeof(double));
  /* net-line histories: detector absorption at the line energies */
  for (double *hist : {var_red_history, brute_history}) {
    if (!hist) continue;
    for (int Z = 1; Z <= 100 && Z <= XMI_MAXZ; Z++) {
      const xrl_atom *a = xrl_atom_get(Z);
      if (!a) continue;
      for (int line = 1; line <= 383; line++) {
        const double E = a->line_energy[line];
        if (!(E > 0.0)) continue;
        double f = -1.0;
        for (int i = 0; i < n_int; i++) {
          double &c = hist[((size_t)(Z - 1) * 385 + (line - 1)) * n_int + i];
          if (c > 0.0) {
            if (f < 0.0) f = det_efficiency(in, E);
            c *= f;
          }
        }
      }
    }
  }
  return 1;
}

extern "C" void xmi_detector_convolute_all(xmi_inputFPtr inputFPtr, double **channels_noconv, double **channels_conv,
                                           double *brute_history, double *var_red_history, xmi_main_options *options,
                                           xmi_escape_ratios *escape_ratios, int n_interactions_all,
                                           int zero_interaction) {
  const xmi_inputF *in = (const xmi_inputF *)inputFPtr;
  const size_t nch = (size_t)in->detector.nchannels;
  for (int n = 0; n <= n_interactions_all; n++) channels_conv[n] = nullptr;
  DetBufs b;
  void *block = nullptr;
  hipStream_t st = nullptr;
  int ok = 0;
  hipError_t e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipMalloc(&block, nch * 5 * sizeof(double) + 256);
  if (e != hipSuccess) {
    xmi_set_error("detector response: device allocation: %s", hipGetErrorString(e));
  } else {
    b.src = (double *)block;
    b.norm = b.src + nch;
    b.conv = b.norm + nch;
    b.cdf = b.conv + nch;
    b.hist = (unsigned long long *)(b.cdf + nch);
    b.over = b.hist + nch;
    for (int n = 0; n <= n_interactions_all; n++) channels_conv[n] = (double *)malloc(nch * sizeof(double));
    ok = det_convolute(in, channels_noconv, channels_conv, brute_history, var_red_history, options, escape_ratios,
                       n_interactions_all, zero_interaction, b, st);
  }
  if (!ok) /* no partial result: every order's pointer is NULL after a failure */
    for (int n = 0; n <= n_interactions_all; n++) {
      free(channels_conv[n]);
      channels_conv[n] = nullptr;
    }
  if (block) (void)hipFree(block);
  if (st) (void)hipStreamDestroy(st);
}

/* ---- the allocation-free variant ------------------------------------------
 * The buffers live in this file-scope array; each device holds its own copy,
 * found once per device through hipGetSymbolAddress.  Calls that overlap on
 * one device share the array and race. */
#define DET_NOALLOC_MAX_CHANNELS 65536
static __device__ __align__(256) unsigned char scratch[(size_t)DET_NOALLOC_MAX_CHANNELS * 5 * sizeof(double) + 256];

#define DET_MAX_DEVICES 64
static void *g_scratch_ptr[DET_MAX_DEVICES];
static std::mutex g_scratch_mutex;

extern "C" int xmi_detector_convolute_all_noalloc(xmi_inputFPtr inputFPtr, double **channels_noconv,
                                                  double **channels_conv, double *brute_history,
                                                  double *var_red_history, xmi_main_options *options,
                                                  xmi_escape_ratios *escape_ratios, int n_interactions_all,
                                                  int zero_interaction) {
  const xmi_inputF *in = (const xmi_inputF *)inputFPtr;
  if (!in || !channels_noconv || !channels_conv || !options) {
    xmi_set_error("xmi_detector_convolute_all_noalloc: NULL argument");
    return 0;
  }
  const int nch = in->detector.nchannels;
  if (nch < 1 || nch > DET_NOALLOC_MAX_CHANNELS) {
    xmi_set_error("xmi_detector_convolute_all_noalloc: %d channels; the allocation-free response holds at most %d",
                  nch, DET_NOALLOC_MAX_CHANNELS);
    return 0;
  }
  for (int n = 0; n <= n_interactions_all; n++)
    if (!channels_conv[n]) {
      xmi_set_error("xmi_detector_convolute_all_noalloc: channels_conv[%d] is NULL (the caller owns the outputs)", n);
      return 0;
    }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= DET_MAX_DEVICES) {
    xmi_set_error("xmi_detector_convolute_all_noalloc: no usable device (%s)", hipGetErrorString(e));
    return 0;
  }
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        xmi_set_error("xmi_detector_convolute_all_noalloc: scratch address: %s", hipGetErrorString(e));
        return 0;
      }
      g_scratch_ptr[dev] = p;
    }
    base = g_scratch_ptr[dev];
  }
  const size_t n = (size_t)nch;
  DetBufs b;
  b.src = (double *)base;
  b.norm = b.src + n;
  b.conv = b.norm + n;
  b.cdf = b.conv + n;
  b.hist = (unsigned long long *)(b.cdf + n);
  b.over = b.hist + n;
  return det_convolute(in, channels_noconv, channels_conv, brute_history, var_red_history, options, escape_ratios,
                       n_interactions_all, zero_interaction, b, (hipStream_t)0);
}

/* the name under which the reference loads a custom detector response
 * (bin/xmimsim.c:516) */
extern "C" void xmi_detector_convolute_all_custom(xmi_inputFPtr inputFPtr, double **channels_noconv,
                                                  double **channels_conv, double *brute_history,
                                                  double *var_red_history, xmi_main_options *options,
                                                  xmi_escape_ratios *escape_ratios, int n_interactions_all,
                                                  int zero_interaction) {
  xmi_detector_convolute_all(inputFPtr, channels_noconv, channels_conv, brute_history, var_red_history, options,
                             escape_ratios, n_interactions_all, zero_interaction);
}
