/*
 * engine_host.cpp -- host driver of the photon-history engine: the C ABI
 * entry points xmi_main_msim and xmi_solid_angle_calculation_cl.
 *
 * xmi_main_msim (src/xmi_main.F90:66-954) becomes:
 *   1. device-resident copies of the loaded tables, cached per table handle;
 *   2. a per-call plan: layers, per-layer attenuation at every fluorescence
 *      line (precalc_mu_cs, src/xmi_main.F90:227-237), the active-line lists
 *      of the estimator, the photon sources (discrete lines, continuous
 *      intervals) with their weights (:253-317, :319-378, :580-601);
 *   3. batches of histories through the event kernels;
 *   4. the 128-bit fixed-point tallies read back once, summed exactly, made
 *      cumulative over interaction order and scaled by live_time (:923-950).
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <chrono>
#include <mutex>
#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "engine_kernels.h"
#include "xmi_internal.h"

long xmi_hits_per_single = 5000;

static uint64_t g_seed = 0;
static int g_seed_set = 0;
static int g_profile = 0;

extern "C" void xmi_msim_set_profiling(int on) { g_profile = on; }

static xmi_msim_launch g_launch = {0, 0, 0, 0};

extern "C" int xmi_msim_set_launch(const xmi_msim_launch *l) {
  if (!l) {
    xmi_set_error("xmi_msim_set_launch: NULL argument");
    return 0;
  }
  if (l->batch_log2 && (l->batch_log2 < 12 || l->batch_log2 > 21)) {
    xmi_set_error("xmi_msim_set_launch: batch_log2 %d outside 12..21", l->batch_log2);
    return 0;
  }
  if (l->lines_kp < 0 || l->lines_kp > 4 || l->wg_per_cu < 0 || l->wg_per_cu > 32 || l->use_graph < 0 ||
      l->use_graph > 1) {
    xmi_set_error("xmi_msim_set_launch: lines_kp %d, wg_per_cu %d or use_graph %d out of range", l->lines_kp,
                  l->wg_per_cu, l->use_graph);
    return 0;
  }
  g_launch = *l;
  return 1;
}

extern "C" void xmi_msim_get_launch(xmi_msim_launch *l) {
  if (l) *l = g_launch;
}
static thread_local xmi_msim_stats g_stats;
static thread_local int64_t g_layer_events[ENG_MAX_LAYERS];

extern "C" int xmi_msim_get_layer_events(int64_t *out, int n) {
  int m = n < ENG_MAX_LAYERS ? n : ENG_MAX_LAYERS;
  for (int l = 0; l < m; l++) out[l] = g_layer_events[l];
  return m;
}

extern "C" void xmi_set_seed(uint64_t seed) {
  g_seed = seed;
  g_seed_set = 1;
}

/* the seed the engine uses under these options (1 with default seeds, as the
 * reference's use_default_seeds; else the xmi_set_seed or time-based one) */
extern "C" uint64_t xmi_engine_current_seed(const xmi_main_options *options) {
  if (options && options->use_default_seeds) return 1;
  if (!g_seed_set) {
    g_seed = (uint64_t)time(NULL) * 6364136223846793005ull + (uint64_t)clock();
    g_seed_set = 1;
  }
  return g_seed;
}

extern "C" void xmi_msim_get_stats(xmi_msim_stats *out) {
  if (out) *out = g_stats;
}

#define HIP_TRY(x)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) {                                                                 \
      xmi_set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
      goto fail;                                                                            \
    }                                                                                       \
  } while (0)

/* ------------------------------------------------------------------------ */
/* device allocations */

/* held while a stream captures a graph and around every device allocation
 * and release of the engines: HIP refuses those (they synchronise the
 * device) while another thread's stream is capturing */
static std::mutex g_capture_lock;

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t bytes) {
    if (bytes <= n && p) return hipSuccess;
    std::lock_guard<std::mutex> g(g_capture_lock);
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e == hipSuccess) n = bytes;
    return e;
  }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
};

/* the energies bounding each cell of the provider's ln(E) grid, with the
 * margins of the cell records' edge tests: lo[c] below cell c, hi[c] above it */
struct CellBounds {
  double lo[XRL_NCS], hi[XRL_NCS];
  CellBounds() {
    const double h = (XRL_CS_LNE1 - XRL_CS_LNE0) / (double)(XRL_NCS - 1);
    for (int c = 0; c < XRL_NCS; c++) {
      lo[c] = exp(XRL_CS_LNE0 + (double)c * h) * (1.0 - 1e-9);
      hi[c] = exp(XRL_CS_LNE0 + (double)(c + 1) * h) * (1.0 + 1e-9);
    }
  }
};
static const CellBounds g_cells;

/* page-locked staging memory of one engine's host <-> device copies, so
 * that no copy takes the runtime's path for pageable memory */
struct Staging {
  void *p = nullptr;
  size_t n = 0;
  ~Staging() {
    if (p) (void)hipHostFree(p);
  }
  hipError_t ensure(size_t bytes) {
    if (bytes <= n && p) return hipSuccess;
    std::lock_guard<std::mutex> g(g_capture_lock);
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipHostMalloc(&p, bytes ? bytes : 16, hipHostMallocDefault);
    if (e == hipSuccess) n = bytes;
    return e;
  }
};

static const size_t STAGE_CHUNK = (size_t)16 << 20;

/* copies through the staging memory on the caller's stream, complete on
 * return (not on the legacy stream, which would wait for streams that may
 * be capturing) */
static hipError_t copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t st, Staging &sg) {
  hipError_t e = sg.ensure(bytes < STAGE_CHUNK ? bytes : STAGE_CHUNK);
  for (size_t off = 0; e == hipSuccess && off < bytes; off += STAGE_CHUNK) {
    const size_t c = bytes - off < STAGE_CHUNK ? bytes - off : STAGE_CHUNK;
    memcpy(sg.p, (const char *)src + off, c);
    e = hipMemcpyAsync((char *)dst + off, sg.p, c, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  return e;
}

static hipError_t copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t st, Staging &sg) {
  hipError_t e = sg.ensure(bytes < STAGE_CHUNK ? bytes : STAGE_CHUNK);
  for (size_t off = 0; e == hipSuccess && off < bytes; off += STAGE_CHUNK) {
    const size_t c = bytes - off < STAGE_CHUNK ? bytes - off : STAGE_CHUNK;
    e = hipMemcpyAsync(sg.p, (const char *)src + off, c, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) memcpy((char *)dst + off, sg.p, c);
  }
  return e;
}

/* many uploads through one staging fill: every array copied into its own
 * 256-byte aligned part of the staging memory, one asynchronous copy each,
 * and one synchronisation at the end */
struct UploadSet {
  struct Item {
    DevBuf *b;
    const void *src;
    size_t bytes;
  };
  std::vector<Item> items;
  template <typename T>
  void add(DevBuf &b, const T *src, size_t count) {
    items.push_back({&b, src, sizeof(T) * count});
  }
  hipError_t run(hipStream_t st, Staging &sg) {
    size_t total = 0;
    for (const Item &it : items) total += (it.bytes + 255) & ~(size_t)255;
    hipError_t e = sg.ensure(total);
    size_t off = 0;
    for (const Item &it : items) {
      if (e == hipSuccess) e = it.b->ensure(it.bytes);
      if (e != hipSuccess) break;
      if (it.bytes) {
        memcpy((char *)sg.p + off, it.src, it.bytes);
        e = hipMemcpyAsync(it.b->p, (char *)sg.p + off, it.bytes, hipMemcpyHostToDevice, st);
        off += (it.bytes + 255) & ~(size_t)255;
      }
    }
    /* the staging memory is reused after this */
    hipError_t e2 = hipStreamSynchronize(st);
    return e != hipSuccess ? e : e2;
  }
};

template <typename T>
static hipError_t upload(DevBuf &b, const T *src, size_t count, hipStream_t st, Staging &sg) {
  hipError_t e = b.ensure(sizeof(T) * count);
  if (e != hipSuccess || !count) return e;
  return copy_h2d(b.p, src, sizeof(T) * count, st, sg);
}

/* tables of one xmi_hdf5F handle on the device */
struct DeviceTables {
  const xmi_hdf5F *owner = nullptr;
  int nZ = 0;
  int Zs[XMI_MAXZ + 1];
  DevBuf atoms, ray, comp, tabE, tabR, fycorr, ipE, ipr, iprc, ipoff, ipacc, cp_icdf, cp_rn, precalc_xrf, lineE;
  DevBuf cp_cdfg, cp_gscale;
  DevBuf phi, phiT, phiR;
  /* the advanced Compton model's subshell profiles, uploaded on first use */
  bool adv_ready = false;
  bool has_precalc = true; /* every element has its cascade cross sections */
  DevBuf cp_shoff, cp_shidx, cp_pcdf, cp_qinv, cp_Qs, cp_inv;
  int n_e = 0, n_r = 0, cp_len = 0, phi_nt = 0, phi_nr = 0;
  double tab_E0 = 0, tab_Einv = 0, tab_R0 = 0, tab_Rinv = 0, phi_T0 = 0, phi_Tinv = 0, phi_R0 = 0, phi_Rinv = 0;
  double ip_acc_E0 = 0, ip_acc_inv = 0;
};

/* device copies of the tables of each xmi_hdf5F handle, shared by all engines */
static std::mutex g_tables_lock;
static std::vector<DeviceTables *> g_tables;

#define ENG_MAX_ORDERS_PROFILED 16

/* one engine: a stream and its working buffers (the default engine serves
 * xmi_main_msim; xmi_main_msim_batch runs several side by side) */
struct Engine {
  std::mutex lock;
  bool ready = false;
  hipStream_t stream = nullptr;
  Staging stage;
  hipEvent_t prof_ev[ENG_ORDER_EVENTS * ENG_MAX_ORDERS_PROFILED] = {};
  /* solid-angle grid cache */
  const void *sa_key = nullptr;
  uint64_t sa_hash = 0;
  DevBuf sa_grid, sa_r, sa_t;
  int sa_nr = 0, sa_nt = 0;
  /* per-call buffers (grow only) */
  DevBuf el_Z, el_w, el_u, precalc_mu, act_n, act_line, act_shell, act_factor, act_E, act_ch, act_murho;
  DevBuf fl_info, fl_wf, fl_mr, lbin_g;
  DevBuf lay_rc, lay_edge, lay_pb, lay_kstart, lay_cell, el_cell;
  DevBuf src_desc, src_line_mus, exc_layers, exc_atoms;
  DevBuf state, vr, tallies, counters, queue, lay_list, lay_woff, lay_total, lay_segbase, cone_acc, esc_tally;
  DevBuf queue2, qcount2, olist;
  DevBuf batch_queue;
};

static Engine g_engine;

static int engine_init(Engine &E) {
  if (E.ready) return 1;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    xmi_set_error("no HIP device available");
    return 0;
  }
  if (hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking) != hipSuccess) {
    xmi_set_error("hipStreamCreate failed");
    return 0;
  }
  E.ready = true;
  return 1;
}
static int engine_init() { return engine_init(g_engine); }

static double uniform_inv(const double *a, int n) { return (double)(n - 1) / (a[n - 1] - a[0]); }

static DeviceTables *get_tables(const xmi_hdf5F *h, hipStream_t st, Staging &sg) {
  std::lock_guard<std::mutex> g(g_tables_lock);
  for (DeviceTables *t : g_tables)
    if (t->owner == h) return t;
  DeviceTables *t = new DeviceTables();
  t->owner = h;
  t->nZ = h->n_Z;
  const int nZ = h->n_Z;
  std::vector<xrl_atom> atoms(nZ);
  for (int i = 0; i < nZ; i++) {
    t->Zs[i] = h->Zs[i].Z;
    const xrl_atom *a = xrl_atom_get(h->Zs[i].Z);
    if (!a) {
      delete t;
      return nullptr;
    }
    atoms[i] = *a;
  }
  const xmi_hdf5_ZF *z0 = &h->Zs[0];
  t->n_e = z0->n_e;
  t->n_r = z0->n_r;
  t->cp_len = z0->cp_len;
  size_t plane = (size_t)t->n_e * t->n_r;
  std::vector<double> ray(plane * nZ), comp(plane * nZ), fy(9 * (size_t)nZ), cp((size_t)t->cp_len * nZ);
  std::vector<double> cdfg((size_t)XMI_CP_GRID * nZ), gscale(nZ);
  std::vector<double> ipE, ipr, iprc;
  std::vector<int> ipoff(nZ + 1), ipacc((size_t)nZ * ENG_IP_ACC);
  std::vector<double> pxrf((size_t)nZ * 9 * nZ * XMI_NLINES_VR, 0.0), lineE((size_t)nZ * XMI_NLINES_VR, 0.0);
  t->ip_acc_E0 = z0->ip_energies[0];
  double Emax = z0->ip_energies[z0->ip_len - 1];
  t->ip_acc_inv = (double)ENG_IP_ACC / (Emax - t->ip_acc_E0);
  for (int i = 0; i < nZ; i++) {
    const xmi_hdf5_ZF *z = &h->Zs[i];
    if (z->n_e != t->n_e || z->n_r != t->n_r || z->cp_len != t->cp_len) {
      xmi_set_error("inconsistent table dimensions for Z=%d", z->Z);
      delete t;
      return nullptr;
    }
    memcpy(&ray[plane * i], z->RayleighTheta_ICDF, sizeof(double) * plane);
    memcpy(&comp[plane * i], z->ComptonTheta_ICDF, sizeof(double) * plane);
    memcpy(&fy[9 * (size_t)i], z->FluorYieldsCorr, sizeof(double) * 9);
    memcpy(&cp[(size_t)t->cp_len * i], z->cp_total_icdf, sizeof(double) * t->cp_len);
    xmi_profile_cdf_grid(z->cp_random_numbers, z->cp_total_icdf, z->cp_len, XMI_CP_GRID,
                         &cdfg[(size_t)XMI_CP_GRID * i]);
    gscale[i] = (double)(XMI_CP_GRID - 1) / z->cp_total_icdf[z->cp_len - 2];
    ipoff[i] = (int)ipE.size();
    ipE.insert(ipE.end(), z->ip_energies, z->ip_energies + z->ip_len);
    ipr.insert(ipr.end(), z->ip_rayl, z->ip_rayl + z->ip_len);
    iprc.insert(iprc.end(), z->ip_rayl_compt, z->ip_rayl_compt + z->ip_len);
    /* accelerator: largest p with e[p] < cell start */
    int p = 0;
    for (int c = 0; c < ENG_IP_ACC; c++) {
      double E = t->ip_acc_E0 + c / t->ip_acc_inv;
      while (p + 2 < z->ip_len && z->ip_energies[p + 1] < E) p++;
      ipacc[(size_t)i * ENG_IP_ACC + c] = p;
    }
    /* a table file read without variance reduction holds no cascade cross
     * sections (src/xmi_data_f.F90 reads them for that mode only) */
    if (z->precalc_xrf_cs)
      memcpy(&pxrf[(size_t)i * 9 * nZ * XMI_NLINES_VR], z->precalc_xrf_cs, sizeof(double) * 9 * nZ * XMI_NLINES_VR);
    else
      t->has_precalc = false;
    for (int l = 1; l <= XMI_NLINES_VR; l++) lineE[(size_t)i * XMI_NLINES_VR + l - 1] = atoms[i].line_energy[l];
  }
  ipoff[nZ] = (int)ipE.size();
  t->tab_E0 = z0->Energies[0];
  t->tab_Einv = uniform_inv(z0->Energies, t->n_e);
  t->tab_R0 = z0->RandomNumbers[0];
  t->tab_Rinv = uniform_inv(z0->RandomNumbers, t->n_r);
  t->phi_nt = h->phi_n_theta;
  t->phi_nr = h->phi_n_r;
  t->phi_T0 = h->Thetas[0];
  t->phi_Tinv = uniform_inv(h->Thetas, h->phi_n_theta);
  t->phi_R0 = h->RandomNumbers[0];
  t->phi_Rinv = uniform_inv(h->RandomNumbers, h->phi_n_r);
  hipError_t e = hipSuccess;
  e = upload(t->atoms, atoms.data(), atoms.size(), st, sg);
  if (e == hipSuccess) e = upload(t->ray, ray.data(), ray.size(), st, sg);
  if (e == hipSuccess) e = upload(t->comp, comp.data(), comp.size(), st, sg);
  if (e == hipSuccess) e = upload(t->tabE, z0->Energies, (size_t)t->n_e, st, sg);
  if (e == hipSuccess) e = upload(t->tabR, z0->RandomNumbers, (size_t)t->n_r, st, sg);
  if (e == hipSuccess) e = upload(t->fycorr, fy.data(), fy.size(), st, sg);
  if (e == hipSuccess) e = upload(t->ipE, ipE.data(), ipE.size(), st, sg);
  if (e == hipSuccess) e = upload(t->ipr, ipr.data(), ipr.size(), st, sg);
  if (e == hipSuccess) e = upload(t->iprc, iprc.data(), iprc.size(), st, sg);
  if (e == hipSuccess) e = upload(t->ipoff, ipoff.data(), ipoff.size(), st, sg);
  if (e == hipSuccess) e = upload(t->ipacc, ipacc.data(), ipacc.size(), st, sg);
  if (e == hipSuccess) e = upload(t->cp_icdf, cp.data(), cp.size(), st, sg);
  if (e == hipSuccess) e = upload(t->cp_rn, z0->cp_random_numbers, (size_t)t->cp_len, st, sg);
  if (e == hipSuccess) e = upload(t->cp_cdfg, cdfg.data(), cdfg.size(), st, sg);
  if (e == hipSuccess) e = upload(t->cp_gscale, gscale.data(), gscale.size(), st, sg);
  if (e == hipSuccess) e = upload(t->precalc_xrf, pxrf.data(), pxrf.size(), st, sg);
  if (e == hipSuccess) e = upload(t->lineE, lineE.data(), lineE.size(), st, sg);
  if (e == hipSuccess) e = upload(t->phi, h->Phi_ICDF, (size_t)h->phi_n_theta * h->phi_n_r, st, sg);
  if (e == hipSuccess) e = upload(t->phiT, h->Thetas, (size_t)h->phi_n_theta, st, sg);
  if (e == hipSuccess) e = upload(t->phiR, h->RandomNumbers, (size_t)h->phi_n_r, st, sg);
  if (e != hipSuccess) {
    xmi_set_error("table upload failed: %s", hipGetErrorString(e));
    delete t;
    return nullptr;
  }
  g_tables.push_back(t);
  return t;
}

/* the partial Compton profiles of every element, flattened over the
 * occupied subshells (advanced Compton model only) */
static int ensure_adv_tables(DeviceTables *t, const xmi_hdf5F *h, hipStream_t st, Staging &sg) {
  std::lock_guard<std::mutex> g(g_tables_lock);
  if (t->adv_ready) return 1;
  const int nZ = h->n_Z, N = t->cp_len;
  std::vector<int> off(nZ + 1), idx;
  std::vector<double> pcdf, qinv;
  for (int i = 0; i < nZ; i++) {
    const xmi_hdf5_ZF *z = &h->Zs[i];
    off[i] = (int)idx.size();
    if (!z->cp_partial_cdf || !z->cp_Qs_inv || !z->cp_Qs || !z->cp_partial_cdf_inv) {
      xmi_set_error("advanced Compton model: no partial Compton profiles for Z=%d", z->Z);
      return 0;
    }
    for (int k = 0; k < z->cp_nshells; k++) idx.push_back(z->cp_shell_indices[k]);
    pcdf.insert(pcdf.end(), z->cp_partial_cdf, z->cp_partial_cdf + (size_t)z->cp_nshells * N);
    qinv.insert(qinv.end(), z->cp_Qs_inv, z->cp_Qs_inv + (size_t)z->cp_nshells * N);
  }
  off[nZ] = (int)idx.size();
  hipError_t e = upload(t->cp_shoff, off.data(), off.size(), st, sg);
  if (e == hipSuccess) e = upload(t->cp_shidx, idx.data(), idx.size(), st, sg);
  if (e == hipSuccess) e = upload(t->cp_pcdf, pcdf.data(), pcdf.size(), st, sg);
  if (e == hipSuccess) e = upload(t->cp_qinv, qinv.data(), qinv.size(), st, sg);
  if (e == hipSuccess) e = upload(t->cp_Qs, h->Zs[0].cp_Qs, (size_t)N, st, sg);
  if (e == hipSuccess) e = upload(t->cp_inv, h->Zs[0].cp_partial_cdf_inv, (size_t)N, st, sg);
  if (e != hipSuccess) {
    xmi_set_error("advanced Compton table upload failed: %s", hipGetErrorString(e));
    return 0;
  }
  t->adv_ready = true;
  return 1;
}

/* released through xmi_hdf5F's destructor hook */
extern "C" void xmi_engine_forget_tables(const void *h) {
  std::lock_guard<std::mutex> g(g_tables_lock);
  for (size_t k = 0; k < g_tables.size(); k++)
    if (g_tables[k]->owner == h) {
      delete g_tables[k];
      g_tables.erase(g_tables.begin() + k);
      return;
    }
}

/* hash of every cell and both axes of a grid: a grid edited in place, or a new
 * grid at a reused address, is uploaded again unless its contents are the same
 * (four independent multiply-xorshift chains over the 64-bit words) */
static uint64_t grid_hash(const xmi_solid_angle *sa) {
  uint64_t h[4] = {0x9E3779B97F4A7C15ull, 0xC2B2AE3D27D4EB4Full, 0x165667B19E3779F9ull, 0x27D4EB2F165667C5ull};
  auto mix = [&](const double *p, size_t n) {
    size_t i = 0;
    for (; i + 4 <= n; i += 4)
      for (int k = 0; k < 4; k++) {
        uint64_t w;
        memcpy(&w, p + i + k, 8);
        h[k] = (h[k] ^ w) * 0xff51afd7ed558ccdull;
        h[k] ^= h[k] >> 29;
      }
    for (; i < n; i++) {
      uint64_t w;
      memcpy(&w, p + i, 8);
      h[0] = (h[0] ^ w) * 0xff51afd7ed558ccdull;
      h[0] ^= h[0] >> 29;
    }
  };
  size_t tot = (size_t)sa->grid_dims_r_n * sa->grid_dims_theta_n;
  mix(sa->grid_dims_r_vals, (size_t)sa->grid_dims_r_n);
  mix(sa->grid_dims_theta_vals, (size_t)sa->grid_dims_theta_n);
  mix(sa->solid_angles, tot);
  return (h[0] ^ (h[1] * 3) ^ (h[2] * 5) ^ (h[3] * 7)) + tot;
}

static int ensure_grid(Engine &E, const xmi_solid_angle *sa) {
  uint64_t hsh = grid_hash(sa);
  if (E.sa_key == sa && E.sa_hash == hsh) return 1;
  size_t tot = (size_t)sa->grid_dims_r_n * sa->grid_dims_theta_n;
  if (upload(E.sa_grid, sa->solid_angles, tot, E.stream, E.stage) != hipSuccess ||
      upload(E.sa_r, sa->grid_dims_r_vals, (size_t)sa->grid_dims_r_n, E.stream, E.stage) != hipSuccess ||
      upload(E.sa_t, sa->grid_dims_theta_vals, (size_t)sa->grid_dims_theta_n, E.stream, E.stage) != hipSuccess) {
    xmi_set_error("solid-angle grid upload failed");
    return 0;
  }
  E.sa_key = sa;
  E.sa_hash = hsh;
  E.sa_nr = (int)sa->grid_dims_r_n;
  E.sa_nt = (int)sa->grid_dims_theta_n;
  g_stats.grid_uploads++;
  return 1;
}

/* ------------------------------------------------------------------------ */

static double layer_mu_host(const xmi_layerF *L, double E) {
  double rv = 0.0;
  for (int i = 0; i < L->n_elements; i++) rv += xrl_cs_total(xrl_atom_get(L->Z[i]), E) * L->weight[i];
  return rv;
}

static void fill_geometry(EngineParams &P, const xmi_inputF *in) {
  for (int k = 0; k < 3; k++) {
    P.n_sample[k] = in->geometry.n_sample_orientation[k];
    P.p_det[k] = in->geometry.p_detector_window[k];
    P.n_det[k] = in->geometry.n_detector_orientation[k];
    for (int j = 0; j < 3; j++) {
      P.det_new[k][j] = in->n_detector_orientation_new[k][j];
      P.det_inv[k][j] = in->n_detector_orientation_inverse[k][j];
    }
  }
  P.det_radius = in->detector_radius;
  P.coll_radius = in->collimator_radius;
  P.coll_height = in->geometry.collimator_height;
  P.coll_present = in->collimator_present;
  for (int k = 0; k < 3; k++) P.det_vertex[k] = in->vertex[k];
  {
    const double c = cos(in->half_apex);
    P.det_cos2 = c * c;
  }
  P.slit_x = in->geometry.slit_size_x;
  P.slit_y = in->geometry.slit_size_y;
  P.d_source_slit = in->geometry.d_source_slit;
}

static int cascade_of(const xmi_main_options *o) {
  if (!o->use_cascade_auger && !o->use_cascade_radiative) return XRL_CASCADE_NONE;
  if (o->use_cascade_auger && !o->use_cascade_radiative) return XRL_CASCADE_NONRADIATIVE;
  if (!o->use_cascade_auger && o->use_cascade_radiative) return XRL_CASCADE_RADIATIVE;
  return XRL_CASCADE_FULL;
}

typedef unsigned __int128 u128;

static inline u128 fx_get(const unsigned long long *b) { return ((u128)b[1] << 64) | (u128)b[0]; }
static inline double fx_to_double(u128 v, int shift) { return ldexp((double)v, -shift); }

/* an escape-ratio run (src/xmi_main.F90:5473-5801): photons of nE input
 * energies, n_photons each, through the detector crystal; results in the
 * reference's Fortran orders */
struct EscapeRun {
  int nE = 0, nOut = 0;
  long n_photons = 0;
  double E0 = 0.0, dE = 0.0, out0 = 0.0, dOut = 0.0;
  std::vector<double> fluo;    /* (element, line 1..109, input energy), element fastest */
  std::vector<double> compton; /* (input energy, output energy), input energy fastest */
};

static int msim_run(Engine &E, xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr, int n_mpi_hosts,
                    double **channelsPtr, xmi_main_options *options, double **brute_historyPtr,
                    double **var_red_historyPtr, xmi_solid_angle *solid_angles, EscapeRun *esc = nullptr,
                    const uint64_t *seed = nullptr) {
  auto t_wall0 = std::chrono::steady_clock::now();
  memset(&g_stats, 0, sizeof(g_stats));
  xmi_inputF *in = (xmi_inputF *)inputFPtr;
  xmi_hdf5F *h = (xmi_hdf5F *)hdf5FPtr;
  if (!in || !h || !options || (!esc && (!channelsPtr || !brute_historyPtr || !var_red_historyPtr))) {
    xmi_set_error("xmi_main_msim: NULL argument");
    return 0;
  }
  if (!in->initialised) {
    xmi_set_error("xmi_main_msim: xmi_init_input was not called");
    return 0;
  }
  /* without variance reduction: the brute-force path, no solid angles,
   * detector hits tallied (src/xmi_main.F90:1274-1407,445-535) */
  const bool brute = !esc && !options->use_variance_reduction;
  /* the Philox counter holds the interaction order in 8 bits (philox.h):
   * order n and n + 256 would share their random streams */
  if (in->general.n_interactions_trajectory < 1 || in->general.n_interactions_trajectory > ENG_MAX_INTERACTIONS) {
    xmi_set_error("xmi_main_msim: n_interactions_trajectory must be in 1..%d, got %ld", ENG_MAX_INTERACTIONS,
                  (long)in->general.n_interactions_trajectory);
    return 0;
  }
  if (!esc && !brute && (!solid_angles || !solid_angles->solid_angles)) {
    xmi_set_error("xmi_main_msim: variance reduction needs a solid-angle grid");
    return 0;
  }
  if (in->n_layers > ENG_MAX_LAYERS) {
    xmi_set_error("xmi_main_msim: at most %d layers", ENG_MAX_LAYERS);
    return 0;
  }
  if (n_mpi_hosts < 1) n_mpi_hosts = 1;
  if (!engine_init(E)) return 0;
  DeviceTables *T = get_tables(h, E.stream, E.stage);
  if (!T) return 0;
  if (options->use_advanced_compton && !ensure_adv_tables(T, h, E.stream, E.stage)) return 0;
  if (!esc && !brute && !T->has_precalc) {
    xmi_set_error("xmi_main_msim: the tables were read without the cross sections of the variance reduction "
                  "(xmi_init_from_hdf5 with use_variance_reduction=0)");
    return 0;
  }
  if (!esc && !brute && !ensure_grid(E, solid_angles)) return 0;
  hipStream_t st = E.stream;
  const int nZ = T->nZ;
  const int n_int = in->general.n_interactions_trajectory;
  const int nch = in->detector.nchannels;
  const int line_last = options->use_M_lines ? XRL_M5P5 : XRL_L3Q1;
  const int NL = in->n_layers;
  double *ch_out = nullptr, *vr_out = nullptr, *bf_out = nullptr;
  hipEvent_t evs[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int ok = 0;

  EngineParams P;
  memset(&P, 0, sizeof(P));
  fill_geometry(P, in);
  P.n_layers = NL;
  P.n_int = n_int;
  P.nch = nch;
  P.gain = in->detector.gain;
  P.zero = in->detector.zero;
  P.line_last = line_last;
  P.cascade = cascade_of(options);
  P.use_M_lines = options->use_M_lines;
  P.hits_per_single = xmi_hits_per_single;
  P.seed = seed ? *seed : xmi_engine_current_seed(options);
  P.sample = 0;
  P.nZ = nZ;
  P.brute = brute ? 1 : 0;
  P.casc = brute ? ((options->use_cascade_auger ? 1 : 0) | (options->use_cascade_radiative ? 2 : 0)) : 0;

  /* layers */
  std::vector<int> elZ, elU;
  std::vector<double> elW;
  for (int l = 0; l < NL; l++) {
    const xmi_layerF *L = &in->layers[l];
    P.layer_density[l] = L->density;
    P.layer_zb[l] = L->Z_coord_begin;
    P.layer_ze[l] = L->Z_coord_end;
    P.layer_nel[l] = L->n_elements;
    P.layer_off[l] = (int)elZ.size();
    for (int k = 0; k < L->n_elements; k++) {
      int u = h->uniqZ[L->Z[k]] - 1;
      if (u < 0) {
        xmi_set_error("xmi_main_msim: Z=%d has no tables (xmi_update_input_from_hdf5 not called?)", L->Z[k]);
        return 0;
      }
      elZ.push_back(L->Z[k]);
      elW.push_back(L->weight[k]);
      elU.push_back(u);
    }
  }
  /* per-layer attenuation tables (d_layer_mu_tab) */
  std::vector<double> lay_rc((size_t)NL * XRL_NCS, 0.0), lay_edge, lay_pb;
  std::vector<int> lay_kstart((size_t)NL * XRL_NCS, 0);
  for (int l = 0; l < NL; l++) {
    const xmi_layerF *L = &in->layers[l];
    std::vector<std::pair<double, int>> edges; /* (edge, element index in layer) */
    for (int k = 0; k < L->n_elements; k++) {
      const xrl_atom *a = xrl_atom_get(L->Z[k]);
      for (int c = 0; c < XRL_NCS; c++) lay_rc[(size_t)l * XRL_NCS + c] += L->weight[k] * a->cs_rc[c];
      for (int sh = 0; sh < 9; sh++)
        if (a->edge[sh] > 0.0) edges.push_back({a->edge[sh], k});
    }
    std::sort(edges.begin(), edges.end(), [](const std::pair<double, int> &x, const std::pair<double, int> &y) {
      return x.first > y.first;
    });
    P.lay_eoff[l] = (int)lay_edge.size();
    std::vector<int> above(L->n_elements, 0);
    for (size_t K = 0; K <= edges.size(); K++) {
      /* K edges above E: each element's own count among them */
      double pb = 0.0;
      for (int k = 0; k < L->n_elements; k++) pb += L->weight[k] * xrl_atom_get(L->Z[k])->photo_Bsum[above[k]];
      lay_pb.push_back(pb);
      if (K < edges.size()) {
        lay_edge.push_back(edges[K].first);
        above[edges[K].second]++;
      }
    }
    for (int c = 0; c < XRL_NCS; c++) {
      const double e_hi = g_cells.hi[c];
      int K = 0;
      while (K < (int)edges.size() && edges[K].first > e_hi) K++;
      lay_kstart[(size_t)l * XRL_NCS + c] = K;
    }
  }
  P.lay_eoff[NL] = (int)lay_edge.size();
  /* the same tables per cell, for one record load per lookup */
  std::vector<LayCell> lay_cell((size_t)NL * XRL_NCS);
  {
    for (int l = 0; l < NL; l++) {
      const int eo = P.lay_eoff[l], ne = P.lay_eoff[l + 1] - eo;
      for (int c = 0; c < XRL_NCS; c++) {
        LayCell &r = lay_cell[(size_t)l * XRL_NCS + c];
        const double *T = &lay_rc[(size_t)l * XRL_NCS];
        r.T0 = T[c];
        r.dT = c + 1 < XRL_NCS ? T[c + 1] - T[c] : 0.0;
        r.K = lay_kstart[(size_t)l * XRL_NCS + c];
        /* every energy of the cell lies above the next edge (with a margin
         * far beyond the rounding of the cell index); the end cells also
         * take the energies beyond the grid, clamped, and keep the walk */
        const double e_lo = g_cells.lo[c];
        r.edge_free = c > 0 && c < XRL_NCS - 2 && (r.K >= ne || lay_edge[eo + r.K] < e_lo);
        r.pb = lay_pb[eo + l + r.K];
      }
    }
  }
  /* per element of every layer, the same records of its own total cross
   * section (the atom selection of the step: one record load per element
   * instead of the atom's edges, photo sum and table, in a chain) */
  std::vector<LayCell> el_cell(elZ.size() * (size_t)XRL_NCS);
  {
    for (size_t j = 0; j < elZ.size(); j++) {
      const xrl_atom *a = xrl_atom_get(elZ[j]);
      std::vector<double> ed;
      for (int sh = 0; sh < 9; sh++)
        if (a->edge[sh] > 0.0) ed.push_back(a->edge[sh]);
      std::sort(ed.begin(), ed.end(), [](double x, double y) { return x > y; });
      const int ne = (int)ed.size();
      for (int c = 0; c < XRL_NCS; c++) {
        LayCell &r = el_cell[j * XRL_NCS + c];
        r.T0 = a->cs_rc[c];
        r.dT = c + 1 < XRL_NCS ? a->cs_rc[c + 1] - a->cs_rc[c] : 0.0;
        const double e_hi = g_cells.hi[c], e_lo = g_cells.lo[c];
        int K = 0;
        while (K < ne && ed[K] > e_hi) K++;
        r.K = K;
        r.edge_free = c > 0 && c < XRL_NCS - 2 && (K >= ne || ed[K] < e_lo);
        r.pb = a->photo_Bsum[K];
      }
    }
  }
  /* the same lookup on the host (as d_layer_mu_tab on the device) */
  auto layer_mu_tab = [&](int l, double E) -> double {
    const xrl_ectx c = xrl_energy(E);
    if (!(c.E > 0.0)) return 0.0;
    const double *T = &lay_rc[(size_t)l * XRL_NCS];
    const int eo = P.lay_eoff[l], ne = P.lay_eoff[l + 1] - eo;
    int K = lay_kstart[(size_t)l * XRL_NCS + c.ci];
    while (K < ne && lay_edge[eo + K] > c.E) K++;
    return c.pe * lay_pb[eo + l + K] + (T[c.ci] + c.cf * (T[c.ci + 1] - T[c.ci]));
  };

  /* attenuation at the fluorescence lines, and the estimator's active lines,
   * for the elements of this input (a table handle may hold more) */
  std::vector<char> present(nZ, 0);
  for (int u : elU) present[u] = 1;
  std::vector<double> pmu((size_t)NL * nZ * XMI_NLINES_VR, 0.0);
  for (int l = 0; l < NL; l++)
    for (int u = 0; u < nZ; u++) {
      if (!present[u]) continue;
      const xrl_atom *a = xrl_atom_get(T->Zs[u]);
      for (int line = 1; line <= XMI_NLINES_VR; line++) {
        double E = a->line_energy[line];
        pmu[((size_t)l * nZ + u) * XMI_NLINES_VR + line - 1] = E > 0.0 ? layer_mu_tab(l, E) : 0.0;
      }
    }
  std::vector<int> actn(nZ, 0), actl((size_t)nZ * ENG_MAX_ACT, 0), acts((size_t)nZ * ENG_MAX_ACT, 0),
      actc((size_t)nZ * ENG_MAX_ACT, -1);
  std::vector<double> actf((size_t)nZ * ENG_MAX_ACT, 0.0), actE((size_t)nZ * ENG_MAX_ACT, 0.0),
      actm((size_t)NL * nZ * ENG_MAX_ACT, 0.0);
  for (int u = 0; u < nZ; u++) {
    if (!present[u]) continue;
    const xrl_atom *a = xrl_atom_get(T->Zs[u]);
    int k = 0;
    for (int line = 1; line <= line_last; line++) {
      double E = a->line_energy[line];
      if (E < XMI_ENERGY_THRESHOLD) continue;
      int s = xrl_line_shell(line);
      double f = a->fluor_yield[s] * a->rad_rate[line];
      if (f == 0.0) continue;
      if (k >= ENG_MAX_ACT) {
        xmi_set_error("xmi_main_msim: more than %d active lines for Z=%d", ENG_MAX_ACT, T->Zs[u]);
        return 0;
      }
      size_t o = (size_t)u * ENG_MAX_ACT + k;
      actl[o] = line;
      acts[o] = s;
      actf[o] = f;
      actE[o] = E;
      int ch = (int)((E - in->detector.zero) / in->detector.gain);
      actc[o] = (ch >= 0 && ch <= nch - 1) ? ch : -1;
      for (int l = 0; l < NL; l++)
        actm[((size_t)l * nZ + u) * ENG_MAX_ACT + k] =
            pmu[((size_t)l * nZ + u) * XMI_NLINES_VR + line - 1] * in->layers[l].density;
      k++;
    }
    actn[u] = k;
  }
  /* compact line bins (estimator LDS) and the per-layer flat (element, line)
   * entries of the lane-per-line phase */
  std::vector<int> act_base(nZ + 1, 0), lbin_g;
  for (int u = 0; u < nZ; u++) act_base[u + 1] = act_base[u] + actn[u];
  for (int u = 0; u < nZ; u++)
    for (int q = 0; q < actn[u]; q++) lbin_g.push_back(u * ENG_MAX_ACT + q);
  P.n_lbins = act_base[nZ];
  std::vector<int4> fl_info;
  std::vector<double> fl_wf;
  std::vector<double> fl_mr;
  const int fl_stride = NL <= 1 ? 1 : NL <= 2 ? 2 : NL <= 4 ? 4 : NL <= 8 ? 8 : 16;
  for (int l = 0; l < NL; l++) {
    P.fl_off[l] = (int)fl_info.size();
    for (int k = 0; k < P.layer_nel[l]; k++) {
      int u = elU[(size_t)P.layer_off[l] + k];
      for (int q = 0; q < actn[u]; q++) {
        size_t o = (size_t)u * ENG_MAX_ACT + q;
        fl_info.push_back(make_int4(u, q, acts[o], act_base[u] + q));
        fl_wf.push_back(elW[(size_t)P.layer_off[l] + k] * actf[o]);
        /* padded to the kernels' layer bound, zeros past the stack */
        for (int l2 = 0; l2 < fl_stride; l2++)
          fl_mr.push_back(l2 < NL ? actm[((size_t)l2 * nZ + u) * ENG_MAX_ACT + q] : 0.0);
      }
    }
  }
  P.fl_off[NL] = (int)fl_info.size();
  /* one zero entry past the end: the walk reads entry t + 1 ahead of use */
  fl_info.push_back(make_int4(0, 0, 0, 0));
  fl_wf.push_back(0.0);
  for (int l2 = 0; l2 < fl_stride; l2++) fl_mr.push_back(0.0);

  /* sources: continuous intervals, then discrete lines */
  std::vector<SourceDesc> src;
  std::vector<double> src_mus;
  double total_weight = 0.0;
  uint64_t hid = 0;
  const long nint_ph = in->general.n_photons_interval, nline_ph = in->general.n_photons_line;
  for (int i = 0; i + 1 < in->n_continuous; i++) {
    const xmi_energy_continuous *e0 = &in->continuous[i], *e1 = &in->continuous[i + 1];
    SourceDesc d;
    memset(&d, 0, sizeof(d));
    d.kind = 0;
    d.x1 = e0->energy;
    d.x2 = e1->energy;
    d.y1 = e0->vertical_intensity + e0->horizontal_intensity;
    d.y2 = e1->vertical_intensity + e1->horizontal_intensity;
    d.h1 = e0->horizontal_intensity;
    d.h2 = e1->horizontal_intensity;
    double tot = (d.y1 + d.y2) * (d.x2 - d.x1) / 2.0;
    d.valid = tot != 0.0;
    if (d.valid && d.x1 >= d.x2) {
      xmi_set_error("xmi_main_msim: continuous energies must increase");
      return 0;
    }
    d.weight = tot / (double)nint_ph;
    d.count = nint_ph / n_mpi_hosts;
    d.hid0 = hid;
    d.sx = e0->sigma_x;
    d.sxp = e0->sigma_xp;
    d.sy = e0->sigma_y;
    d.syp = e0->sigma_yp;
    hid += (uint64_t)nint_ph;
    if (d.valid) total_weight += tot;
    src.push_back(d);
    for (int l = 0; l < NL; l++) src_mus.push_back(0.0);
  }
  for (int i = 0; i < in->n_discrete; i++) {
    const xmi_energy_discrete *e = &in->discrete[i];
    SourceDesc d;
    memset(&d, 0, sizeof(d));
    d.kind = 1;
    d.valid = 1;
    d.dist = (int)e->distribution_type;
    d.energy = e->energy;
    d.scale = e->scale_parameter;
    double tot = e->vertical_intensity + e->horizontal_intensity;
    d.count = nline_ph / n_mpi_hosts;
    d.hor_ver_ratio = e->horizontal_intensity * (double)d.count / tot;
    double exc = 1.0;
    for (int k = 0; k < in->n_exc_layers; k++) {
      const xmi_layerF *L = &in->exc_layers[k];
      exc *= exp(-1.0 * L->density * L->thickness * layer_mu_host(L, e->energy));
    }
    d.weight = tot * exc / (double)nline_ph;
    d.hid0 = hid;
    d.sx = e->sigma_x;
    d.sxp = e->sigma_xp;
    d.sy = e->sigma_y;
    d.syp = e->sigma_yp;
    hid += (uint64_t)nline_ph;
    total_weight += tot;
    src.push_back(d);
    for (int l = 0; l < NL; l++) src_mus.push_back(layer_mu_host(&in->layers[l], e->energy));
  }
  if (esc) {
    /* escape-ratio run: one line per input energy, weight 1, random
     * polarisation azimuth (src/xmi_main.F90:5640-5690) */
    src.clear();
    src_mus.clear();
    for (int k = 0; k < esc->nE; k++) {
      SourceDesc d;
      memset(&d, 0, sizeof(d));
      d.kind = 1;
      d.valid = 1;
      d.dist = 0;
      d.pol = 1;
      d.energy = esc->E0 + k * esc->dE;
      d.count = esc->n_photons;
      d.weight = 1.0;
      d.hid0 = (uint64_t)k * (uint64_t)esc->n_photons;
      src.push_back(d);
      for (int l = 0; l < NL; l++) src_mus.push_back(layer_mu_host(&in->layers[l], d.energy));
    }
    total_weight = (double)esc->n_photons; /* the largest tally: one energy's photons */
  }
  /* history ids are laid out with the full per-source budget; batches only run
   * the first `count` of each source */
  std::vector<double> exc_flat;
  for (int k = 0; k < in->n_exc_layers; k++) {
    const xmi_layerF *L = &in->exc_layers[k];
    exc_flat.push_back(L->density);
    exc_flat.push_back(L->thickness);
    exc_flat.push_back((double)L->n_elements);
    for (int j = 0; j < L->n_elements; j++) exc_flat.push_back((double)L->Z[j]);
    for (int j = 0; j < L->n_elements; j++) exc_flat.push_back(L->weight[j]);
  }
  std::vector<xrl_atom> exc_atoms;
  if (in->n_exc_layers > 0) {
    exc_atoms.resize(XMI_MAXZ + 1);
    for (int k = 0; k < in->n_exc_layers; k++)
      for (int j = 0; j < in->exc_layers[k].n_elements; j++) exc_atoms[in->exc_layers[k].Z[j]] = *xrl_atom_get(in->exc_layers[k].Z[j]);
  }
  /* fixed-point scale: the total source intensity maps to 2^100, leaving a
   * factor 2^27 of headroom below 2^127 and a resolution of 2^-100 of it */
  int shift = 100;
  if (total_weight > 0.0) shift = 100 - (int)ceil(log2(total_weight));
  if (shift > 900) shift = 900;
  if (shift < -900) shift = -900;
  P.tally_shift = shift;
  P.lines_kp = g_launch.lines_kp;
  P.wg_per_cu = g_launch.wg_per_cu;

  {
    Engine &G = E;
    /* histories: each source's first `count` ids, all sources back to back in
     * one run order, so that every batch is full whatever the source sizes */
    int64_t total_hist = 0;
    for (auto &d : src) {
      d.run0 = (uint64_t)total_hist;
      total_hist += (d.valid && d.count > 0) ? d.count : 0;
    }
    const int cap_max = 1 << (g_launch.batch_log2 ? g_launch.batch_log2 : 21);
    int cap = (int)(total_hist < cap_max ? (total_hist > 0 ? total_hist : 1) : cap_max);
    /* a brute-force run with cascades keeps a second slot per primary for
     * the offspring photon of its cascade */
    const int slots = (brute && P.casc) ? 2 * cap : cap;
    P.cap = slots;
    P.off_half = (brute && P.casc) ? cap : 0;
    size_t per = (size_t)slots;
    /* the kernels' layer-count bound (estimator records hold nl_bound paths) */
    const int nl_bound = NL <= 1 ? 1 : NL <= 2 ? 2 : NL <= 4 ? 4 : NL <= 8 ? 8 : 16;
    size_t state_bytes = per * (15 * sizeof(double) + (size_t)NL * sizeof(double) + 5 * sizeof(int) + sizeof(uint64_t));
    size_t vr_bytes = per * ((6 + (size_t)nl_bound) * sizeof(double) + 2 * sizeof(int));
    size_t n_ch2 = (size_t)n_int * nch * 2, n_line2 = (size_t)n_int * nZ * ENG_MAX_ACT * 2, n_rc2 = (size_t)n_int * nZ * 4;
    /* brute force: hits by interactions so far, and by the last interaction */
    size_t n_bch2 = brute ? (size_t)(n_int + 1) * nch * 2 : 0, n_bh2 = brute ? (size_t)n_int * nZ * 385 * 2 : 0;
    size_t tally_words = n_ch2 + n_line2 + n_rc2 + n_bch2 + n_bh2;
    HIP_TRY(G.state.ensure(state_bytes));
    HIP_TRY(G.vr.ensure(vr_bytes));
    HIP_TRY(G.tallies.ensure(tally_words * sizeof(unsigned long long)));
    int dev = 0, n_cu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    /* sub-queue size: k_step runs ceil(cap / ENG_BLOCK) workgroups, and each
     * sub-queue takes the photons of every ENG_SPREAD-th of their waves */
    const int step_waves = ((slots + ENG_BLOCK - 1) / ENG_BLOCK) * (ENG_BLOCK / 64);
    P.qreg = ((step_waves + ENG_SPREAD - 1) / ENG_SPREAD) * 64;
    /* the brute-force grid of bblocks workgroups strides over at most `slots`
     * live photons: each of its waves takes at most ceil(slots / (64 W))
     * trips of 64 entries, and a sub-queue serves ceil(W / ENG_SPREAD) of
     * its W waves */
    const int bblocks = brute ? eng_brute_blocks(n_cu) : 0;
    if (brute) {
      const long W = (long)bblocks * (ENG_BLOCK / 64);
      const long trips = ((long)slots + 64 * W - 1) / (64 * W);
      const long need = ((W + ENG_SPREAD - 1) / ENG_SPREAD) * 64 * trips;
      if (need > P.qreg) P.qreg = (int)need;
    }
    HIP_TRY(G.counters.ensure((size_t)ENG_NCOUNTERS * ENG_SPREAD * sizeof(unsigned long long)));
    HIP_TRY(G.queue.ensure((size_t)ENG_NQUEUES * ENG_SPREAD * P.qreg * sizeof(int)));
    /* brute force: a second queue set (the interactions of order n - 1 while
     * order n's step fills the first), and two offspring lists */
    if (brute) {
      HIP_TRY(G.queue2.ensure((size_t)ENG_NQUEUES * ENG_SPREAD * P.qreg * sizeof(int)));
      HIP_TRY(G.qcount2.ensure((size_t)ENG_NQUEUES * ENG_SPREAD * sizeof(unsigned long long) + 2 * sizeof(unsigned int)));
      HIP_TRY(G.olist.ensure((size_t)2 * (P.off_half ? P.off_half : 1) * sizeof(int)));
    }
    /* layer lists: a row of cap slots per layer; per-wave offsets of the
     * k_step launch (ceil(cap / ENG_BLOCK) workgroups of ENG_BLOCK / 64 waves) */
    P.n_waves = step_waves;
    if (!brute && step_waves > 32 * 1024) {
      xmi_set_error("xmi_main_msim: %d waves per batch exceed the layer scan's 32768", step_waves);
      return 0;
    }
    HIP_TRY(G.lay_list.ensure((size_t)NL * cap * sizeof(int)));
    HIP_TRY(G.lay_woff.ensure((size_t)NL * step_waves * sizeof(int)));
    HIP_TRY(G.lay_total.ensure((size_t)ENG_MAX_LAYERS * sizeof(int)));
    HIP_TRY(G.lay_segbase.ensure((size_t)ENG_MAX_LAYERS * 32 * sizeof(int)));
    /* off-grid cone accumulators: zero here, and cleared again by the
     * kernel after each photon's estimate */
    {
      const size_t cone_bytes = (size_t)ENG_SPREAD * P.qreg * sizeof(unsigned long long);
      const bool fresh = G.cone_acc.n < cone_bytes || !G.cone_acc.p;
      HIP_TRY(G.cone_acc.ensure(cone_bytes));
      if (fresh) HIP_TRY(hipMemsetAsync(G.cone_acc.p, 0, cone_bytes, st));
    }
    {
      UploadSet up;
      up.add(G.el_Z, elZ.data(), elZ.size());
      up.add(G.el_w, elW.data(), elW.size());
      up.add(G.el_u, elU.data(), elU.size());
      up.add(G.precalc_mu, pmu.data(), pmu.size());
      up.add(G.act_n, actn.data(), actn.size());
      up.add(G.act_line, actl.data(), actl.size());
      up.add(G.act_shell, acts.data(), acts.size());
      up.add(G.act_factor, actf.data(), actf.size());
      up.add(G.act_E, actE.data(), actE.size());
      up.add(G.act_ch, actc.data(), actc.size());
      up.add(G.act_murho, actm.data(), actm.size());
      up.add(G.fl_info, fl_info.data(), fl_info.size());
      up.add(G.fl_wf, fl_wf.data(), fl_wf.size());
      up.add(G.fl_mr, fl_mr.data(), fl_mr.size());
      up.add(G.lbin_g, lbin_g.data(), lbin_g.size());
      up.add(G.lay_rc, lay_rc.data(), lay_rc.size());
      up.add(G.lay_edge, lay_edge.data(), lay_edge.size());
      up.add(G.lay_pb, lay_pb.data(), lay_pb.size());
      up.add(G.lay_kstart, lay_kstart.data(), lay_kstart.size());
      up.add(G.lay_cell, lay_cell.data(), lay_cell.size());
      up.add(G.el_cell, el_cell.data(), el_cell.size());
      up.add(G.src_desc, src.data(), src.size());
      up.add(G.src_line_mus, src_mus.data(), src_mus.size());
      up.add(G.exc_layers, exc_flat.data(), exc_flat.size());
      up.add(G.exc_atoms, exc_atoms.data(), exc_atoms.size());
      HIP_TRY(up.run(st, G.stage));
    }
    HIP_TRY(hipMemsetAsync(G.tallies.p, 0, tally_words * sizeof(unsigned long long), st));
    HIP_TRY(hipMemsetAsync(G.counters.p, 0, (size_t)ENG_NCOUNTERS * ENG_SPREAD * sizeof(unsigned long long), st));

    P.el_Z = G.el_Z.as<int>();
    P.el_w = G.el_w.as<double>();
    P.el_u = G.el_u.as<int>();
    P.atoms = T->atoms.as<xrl_atom>();
    P.ray_icdf = T->ray.as<double>();
    P.comp_icdf = T->comp.as<double>();
    P.tab_E = T->tabE.as<double>();
    P.tab_R = T->tabR.as<double>();
    P.n_e = T->n_e;
    P.n_r = T->n_r;
    P.tab_E0 = T->tab_E0;
    P.tab_Einv = T->tab_Einv;
    P.tab_R0 = T->tab_R0;
    P.tab_Rinv = T->tab_Rinv;
    P.fycorr = T->fycorr.as<double>();
    P.ip_E = T->ipE.as<double>();
    P.ip_r = T->ipr.as<double>();
    P.ip_rc = T->iprc.as<double>();
    P.ip_off = T->ipoff.as<int>();
    P.ip_acc = T->ipacc.as<int>();
    P.ip_acc_E0 = T->ip_acc_E0;
    P.ip_acc_inv = T->ip_acc_inv;
    P.cp_icdf = T->cp_icdf.as<double>();
    P.cp_rn = T->cp_rn.as<double>();
    P.cp_len = T->cp_len;
    P.cp_cdfg = T->cp_cdfg.as<double>();
    P.cp_gscale = T->cp_gscale.as<double>();
    P.adv_compton = options->use_advanced_compton ? 1 : 0;
    if (P.adv_compton) {
      P.cp_sh_off = T->cp_shoff.as<int>();
      P.cp_sh_idx = T->cp_shidx.as<int>();
      P.cp_pcdf = T->cp_pcdf.as<double>();
      P.cp_qinv = T->cp_qinv.as<double>();
      P.cp_Qs = T->cp_Qs.as<double>();
      P.cp_inv = T->cp_inv.as<double>();
    }
    P.precalc_xrf = T->precalc_xrf.as<double>();
    P.line_E = T->lineE.as<double>();
    P.precalc_mu = G.precalc_mu.as<double>();
    P.act_n = G.act_n.as<int>();
    P.act_line = G.act_line.as<int>();
    P.act_shell = G.act_shell.as<int>();
    P.act_factor = G.act_factor.as<double>();
    P.act_E = G.act_E.as<double>();
    P.act_ch = G.act_ch.as<int>();
    P.act_murho = G.act_murho.as<double>();
    P.fl_info = G.fl_info.as<int4>();
    P.fl_wf = G.fl_wf.as<double>();
    P.fl_mr = G.fl_mr.as<double>();
    P.lbin_g = G.lbin_g.as<int>();
    P.lay_rc = G.lay_rc.as<double>();
    P.lay_edge = G.lay_edge.as<double>();
    P.lay_pb = G.lay_pb.as<double>();
    P.lay_kstart = G.lay_kstart.as<int>();
    P.lay_cell = G.lay_cell.as<LayCell>();
    P.el_cell = G.el_cell.as<LayCell>();
    P.phi_icdf = T->phi.as<double>();
    P.phi_T = T->phiT.as<double>();
    P.phi_R = T->phiR.as<double>();
    P.phi_nt = T->phi_nt;
    P.phi_nr = T->phi_nr;
    P.phi_T0 = T->phi_T0;
    P.phi_Tinv = T->phi_Tinv;
    P.phi_R0 = T->phi_R0;
    P.phi_Rinv = T->phi_Rinv;
    if (!esc && !brute) {
      P.sa_grid = G.sa_grid.as<double>();
      P.sa_r = G.sa_r.as<double>();
      P.sa_t = G.sa_t.as<double>();
      P.sa_nr = G.sa_nr;
      P.sa_nt = G.sa_nt;
      P.sa_r0 = solid_angles->grid_dims_r_vals[0];
      P.sa_rinv = uniform_inv(solid_angles->grid_dims_r_vals, (int)solid_angles->grid_dims_r_n);
      P.sa_t0 = solid_angles->grid_dims_theta_vals[0];
      P.sa_tinv = uniform_inv(solid_angles->grid_dims_theta_vals, (int)solid_angles->grid_dims_theta_n);
    }
    /* escape tallies: fluorescence, Compton, normalisation */
    EscParams Q;
    memset(&Q, 0, sizeof(Q));
    size_t esc_words = 0;
    if (esc) {
      const size_t nf = (size_t)esc->nE * ENG_ESC_LINES * nZ, nc = (size_t)esc->nOut * esc->nE;
      esc_words = 2 * (nf + nc + (size_t)esc->nE);
      HIP_TRY(G.esc_tally.ensure(esc_words * sizeof(unsigned long long)));
      HIP_TRY(hipMemsetAsync(G.esc_tally.p, 0, esc_words * sizeof(unsigned long long), st));
      Q.fluo = G.esc_tally.as<unsigned long long>();
      Q.compton = Q.fluo + 2 * nf;
      Q.interacted = Q.compton + 2 * nc;
      Q.n_photons = esc->n_photons;
      Q.nE = esc->nE;
      Q.nOut = esc->nOut;
      Q.out0 = esc->out0;
      Q.dOut = esc->dOut;
    }
    unsigned long long *tw = G.tallies.as<unsigned long long>();
    P.t_ch = tw;
    P.t_line = tw + n_ch2;
    P.t_rc = tw + n_ch2 + n_line2;
    P.t_bch = brute ? tw + n_ch2 + n_line2 + n_rc2 : nullptr;
    P.t_bhist = brute ? P.t_bch + n_bch2 : nullptr;
    P.counters = G.counters.as<unsigned long long>();
    P.queue = G.queue.as<int>();
    P.qcount = P.counters;
    P.pqueue = nullptr;
    P.pqcount = nullptr;
    P.olist_in = nullptr;
    P.ocount_in = nullptr;
    P.olist_out = nullptr;
    P.ocount_out = nullptr;
    P.lay_list = G.lay_list.as<int>();
    P.lay_woff = G.lay_woff.as<int>();
    P.lay_total = G.lay_total.as<int>();
    P.lay_segbase = G.lay_segbase.as<int>();
    P.cone_acc = G.cone_acc.as<unsigned long long>();

    PhotonSoA S;
    char *sp = G.state.as<char>();
    double **dfields[] = {&S.x, &S.y, &S.z, &S.dx, &S.dy, &S.dz, &S.ex, &S.ey, &S.ez, &S.theta, &S.phi, &S.energy, &S.weight};
    for (double **f : dfields) {
      *f = (double *)sp;
      sp += per * sizeof(double);
    }
    sp += 2 * per * sizeof(double);
    S.mus = (double *)sp;
    sp += per * NL * sizeof(double);
    S.hid = (uint64_t *)sp;
    sp += per * sizeof(uint64_t);
    S.layer = (int *)sp;
    sp += per * sizeof(int);
    S.elem = (int *)sp;
    sp += per * sizeof(int);
    S.prev_line = (int *)sp;
    sp += per * sizeof(int);
    S.prev_u = (int *)sp;
    sp += per * sizeof(int);
    S.cflags = (int *)sp;

    VrSoA V;
    char *vp = G.vr.as<char>();
    double **vfields[] = {&V.theta, &V.phi, &V.omega, &V.pesc, &V.sar, &V.sat};
    for (double **f : vfields) {
      *f = (double *)vp;
      vp += per * sizeof(double);
    }
    V.dist = (double *)vp;
    vp += per * nl_bound * sizeof(double);
    V.info = (int *)vp;
    vp += per * sizeof(int);
    V.blk = (unsigned int *)vp;

    SourceTable ST;
    ST.d = G.src_desc.as<SourceDesc>();
    ST.n = (int)src.size();
    ST.line_mus = G.src_line_mus.as<double>();
    ST.exc_layers = G.exc_layers.as<double>();
    ST.n_exc = in->n_exc_layers;
    ST.exc_atoms = G.exc_atoms.as<xrl_atom>();


    HIP_TRY(hipEventCreate(&evs[0]));
    HIP_TRY(hipEventCreate(&evs[1]));
    if (g_profile)
      for (int k = 0; k < ENG_ORDER_EVENTS * ENG_MAX_ORDERS_PROFILED; k++)
        if (!G.prof_ev[k]) HIP_TRY(hipEventCreate(&G.prof_ev[k]));
    HIP_TRY(hipEventRecord(evs[0], st));
    /* batches of `cap` consecutive positions of the run order */
    {
      const int64_t nb = (total_hist + cap - 1) / cap;
      g_stats.batches = nb;
      /* the launches of one batch after its source kernel (count: the
       * batch's primaries, or the capacity in a captured graph, whose
       * source kernel marks the slots beyond the batch empty) */
      int *qs[2] = {G.queue.as<int>(), G.queue2.as<int>()};
      unsigned long long *qc[2] = {G.counters.as<unsigned long long>(), G.qcount2.as<unsigned long long>()};
      unsigned int *oc = brute ? (unsigned int *)(qc[1] + ENG_NQUEUES * ENG_SPREAD) : nullptr;
      int *ol[2] = {G.olist.as<int>(), brute ? G.olist.as<int>() + (P.off_half ? P.off_half : 1) : nullptr};
      auto orders = [&](int count) -> hipError_t {
        hipError_t e = hipSuccess;
        if (esc) return eng_launch_escape(&P, &S, &V, &Q, count, st);
        if (brute) {
          /* order n_int + 1: the step after the last interaction, in which a
           * photon can still reach the detector.  Order n fills queue set
           * n & 1 and offspring list n & 1, and reads those of order n - 1 */
          const int bprof = g_profile && 3 * (n_int + 1) <= ENG_ORDER_EVENTS * ENG_MAX_ORDERS_PROFILED;
          for (int n = 1; n <= n_int + 1 && e == hipSuccess; n++) {
            EngineParams Pn = P;
            const int c = n & 1, p = c ^ 1;
            Pn.queue = qs[c];
            Pn.qcount = qc[c];
            Pn.pqueue = qs[p];
            Pn.pqcount = qc[p];
            Pn.olist_in = ol[p];
            Pn.ocount_in = oc + p;
            Pn.olist_out = ol[c];
            Pn.ocount_out = oc + c;
            e = eng_launch_order_brute(&Pn, &S, n, count, bblocks, st, bprof ? &G.prof_ev[3 * (n - 1)] : nullptr);
          }
          if (bprof && e == hipSuccess) {
            if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
            for (int n = 0; n <= n_int; n++) {
              float t0 = 0.f, t1 = 0.f;
              if ((e = hipEventElapsedTime(&t0, G.prof_ev[3 * n], G.prof_ev[3 * n + 1])) != hipSuccess) return e;
              if ((e = hipEventElapsedTime(&t1, G.prof_ev[3 * n + 1], G.prof_ev[3 * n + 2])) != hipSuccess) return e;
              g_stats.step_kernel_ms += t0;
              g_stats.interaction_kernel_ms += t1;
            }
            g_stats.launches += n_int + 1; /* step launches */
          }
          return e;
        }
        const int prof = g_profile && n_int <= ENG_MAX_ORDERS_PROFILED;
        for (int n = 1; n <= n_int && e == hipSuccess; n++)
          e = eng_launch_order(&P, &S, &V, n, count, n_cu, st, prof ? &G.prof_ev[ENG_ORDER_EVENTS * (n - 1)] : nullptr);
        if (prof && e == hipSuccess) {
          if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
          for (int n = 0; n < n_int; n++) {
            float t[ENG_ORDER_EVENTS - 1];
            const hipEvent_t *ev = &G.prof_ev[ENG_ORDER_EVENTS * n];
            for (int k = 0; k + 1 < ENG_ORDER_EVENTS; k++)
              if ((e = hipEventElapsedTime(&t[k], ev[k], ev[k + 1])) != hipSuccess) return e;
            g_stats.step_kernel_ms += t[0];
            g_stats.vr_geom_ms += t[1];
            g_stats.vr_cone_ms += t[2];
            g_stats.vr_scatter_ms += t[3];
            g_stats.vr_lines_ms += t[4];
            g_stats.vr_kernel_ms += t[1] + t[2] + t[3] + t[4];
            g_stats.interaction_kernel_ms += t[5];
          }
          g_stats.launches += n_int; /* launches of each kernel class */
        }
        return e;
      };
      if (g_launch.use_graph && !esc && !g_profile) {
        /* one captured graph per run: the pop of the next batch from the
         * device batch queue, the source kernel over `cap` slots and the
         * orders; replayed once per batch */
        HIP_TRY(G.batch_queue.ensure(sizeof(BatchQueue)));
        BatchQueue *bq = G.batch_queue.as<BatchQueue>();
        HIP_TRY(eng_launch_batch_init(bq, (unsigned long long)total_hist, st));
        hipError_t ce, ee, ie = hipSuccess;
        {
          std::lock_guard<std::mutex> g(g_capture_lock);
          ce = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
          if (ce == hipSuccess) {
            ce = eng_launch_source_queued(&P, &S, &ST, bq, cap, st);
            if (ce == hipSuccess) ce = orders(cap);
            ee = hipStreamEndCapture(st, &graph);
            if (ce == hipSuccess && ee == hipSuccess) ie = hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0);
          } else {
            ee = hipSuccess;
          }
        }
        HIP_TRY(ce);
        HIP_TRY(ee);
        HIP_TRY(ie);
        for (int64_t b = 0; b < nb; b++) HIP_TRY(hipGraphLaunch(graph_exec, st));
        g_stats.graph_replays = nb;
      } else {
        for (int64_t done = 0; done < total_hist; done += cap) {
          const int count = (int)((total_hist - done) < cap ? (total_hist - done) : cap);
          HIP_TRY(eng_launch_source(&P, &S, &ST, (uint64_t)done, count, st));
          HIP_TRY(orders(count));
        }
      }
    }
    HIP_TRY(hipEventRecord(evs[1], st));
    std::vector<unsigned long long> th(tally_words);
    std::vector<unsigned long long> cnt_sub((size_t)ENG_NCOUNTERS * ENG_SPREAD);
    unsigned long long cnt[ENG_NCOUNTERS];
    HIP_TRY(copy_d2h(th.data(), G.tallies.p, tally_words * sizeof(unsigned long long), st, G.stage));
    HIP_TRY(copy_d2h(cnt_sub.data(), G.counters.p, cnt_sub.size() * sizeof(unsigned long long), st, G.stage));
    for (int c = 0; c < ENG_NCOUNTERS; c++) {
      cnt[c] = 0ull;
      for (int j = 0; j < ENG_SPREAD; j++) cnt[c] += cnt_sub[(size_t)c * ENG_SPREAD + j];
    }
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, evs[0], evs[1]));
    g_stats.transport_ms = ms;
    g_stats.histories = cnt[ENG_CNT_HISTORIES];
    g_stats.vr_events = (int64_t)cnt[ENG_CNT_VR];
    g_stats.solid_angle_fallbacks = cnt[ENG_CNT_FALLBACKS];
    g_stats.photon_steps = (int64_t)cnt[ENG_CNT_STEPS];
    g_stats.detector_hits = (int64_t)cnt[ENG_CNT_HITS];
    for (int l = 0; l < ENG_MAX_LAYERS; l++) g_layer_events[l] = cnt[ENG_CNT_LAYER0 + l];
    if (cnt[ENG_CNT_ERRORS]) {
      xmi_set_error("xmi_main_msim: %llu photons hit an invalid state on the device (source %llu, estimator Doppler %llu, "
                    "transport Doppler %llu)",
                    cnt[ENG_CNT_ERRORS], cnt[ENG_CNT_ERR_SOURCE], cnt[ENG_CNT_ERR_VR_DOPPLER], cnt[ENG_CNT_ERR_DOPPLER]);
      goto fail;
    }

    if (esc) {
      /* ratios: escaped weight over the weight of the photons that
       * interacted, per input energy (src/xmi_main.F90:5773-5776) */
      std::vector<unsigned long long> et(esc_words);
      HIP_TRY(copy_d2h(et.data(), G.esc_tally.p, esc_words * sizeof(unsigned long long), st, G.stage));
      const size_t nf = (size_t)esc->nE * ENG_ESC_LINES * nZ, nc = (size_t)esc->nOut * esc->nE;
      const unsigned long long *tf = et.data(), *tc = tf + 2 * nf, *ti = tc + 2 * nc;
      esc->fluo.assign(nf, 0.0);
      esc->compton.assign(nc, 0.0);
      for (int k = 0; k < esc->nE; k++) {
        const double norm = fx_to_double(fx_get(&ti[(size_t)k * 2]), shift);
        if (!(norm > 0.0)) continue;
        for (int line = 0; line < ENG_ESC_LINES; line++)
          for (int u = 0; u < nZ; u++)
            esc->fluo[(size_t)u + (size_t)nZ * (line + (size_t)ENG_ESC_LINES * k)] =
                fx_to_double(fx_get(&tf[(((size_t)k * ENG_ESC_LINES + line) * nZ + u) * 2]), shift) / norm;
        for (int o = 0; o < esc->nOut; o++)
          esc->compton[(size_t)k + (size_t)esc->nE * o] =
              fx_to_double(fx_get(&tc[((size_t)o * esc->nE + k) * 2]), shift) / norm;
      }
      ok = 1;
      goto fail;
    }
    if (brute) {
      /* channels(k, :) holds the photons that reached the detector after at
       * most k interactions; the history rows by the order of the last */
      const unsigned long long *tbc = th.data() + n_ch2 + n_line2 + n_rc2, *tbh = tbc + n_bch2;
      const double live = in->detector.live_time;
      const size_t nhist = (size_t)100 * XMI_HIST_ROWS * n_int;
      ch_out = (double *)calloc((size_t)(n_int + 1) * nch, sizeof(double));
      bf_out = (double *)calloc(nhist, sizeof(double));
      if (!ch_out || !bf_out) {
        xmi_set_error("out of host memory");
        goto fail;
      }
      for (int c = 0; c < nch; c++) {
        u128 acc = 0;
        for (int k = 0; k <= n_int; k++) {
          acc += fx_get(&tbc[((size_t)k * nch + c) * 2]);
          ch_out[(size_t)k * nch + c] = fx_to_double(acc, shift) * live;
        }
      }
      for (int u = 0; u < nZ; u++) {
        const int Z = T->Zs[u];
        for (int n = 0; n < n_int; n++)
          for (int row = 0; row < XMI_HIST_ROWS; row++) {
            const u128 v = fx_get(&tbh[(((size_t)n * nZ + u) * 385 + row) * 2]);
            if (v) bf_out[((size_t)(Z - 1) * XMI_HIST_ROWS + row) * n_int + n] = fx_to_double(v, shift) * live * n_mpi_hosts;
          }
      }
      ok = 1;
      goto fail;
    }
    /* exact sums: lines into their channels, prefix over orders */
    const unsigned long long *tch = th.data(), *tln = th.data() + n_ch2, *trc = th.data() + n_ch2 + n_line2;
    std::vector<u128> order_ch((size_t)n_int * nch);
    for (int n = 0; n < n_int; n++) {
      for (int c = 0; c < nch; c++) order_ch[(size_t)n * nch + c] = fx_get(&tch[((size_t)n * nch + c) * 2]);
      for (int u = 0; u < nZ; u++)
        for (int k = 0; k < actn[u]; k++) {
          int c = actc[(size_t)u * ENG_MAX_ACT + k];
          if (c < 0) continue;
          order_ch[(size_t)n * nch + c] += fx_get(&tln[(((size_t)n * nZ + u) * ENG_MAX_ACT + k) * 2]);
        }
    }
    const double live = in->detector.live_time;
    size_t nhist = (size_t)100 * XMI_HIST_ROWS * n_int;
    ch_out = (double *)calloc((size_t)(n_int + 1) * nch, sizeof(double));
    vr_out = (double *)calloc(nhist, sizeof(double));
    bf_out = (double *)calloc(nhist, sizeof(double));
    if (!ch_out || !vr_out || !bf_out) {
      xmi_set_error("out of host memory");
      goto fail;
    }
    for (int c = 0; c < nch; c++) {
      u128 acc = 0;
      for (int k = 1; k <= n_int; k++) {
        acc += order_ch[(size_t)(k - 1) * nch + c];
        ch_out[(size_t)k * nch + c] = fx_to_double(acc, shift) * live;
      }
    }
    for (int u = 0; u < nZ; u++) {
      int Z = T->Zs[u];
      for (int n = 0; n < n_int; n++) {
        for (int k = 0; k < actn[u]; k++) {
          int line = actl[(size_t)u * ENG_MAX_ACT + k];
          vr_out[((size_t)(Z - 1) * XMI_HIST_ROWS + (line - 1)) * n_int + n] =
              fx_to_double(fx_get(&tln[(((size_t)n * nZ + u) * ENG_MAX_ACT + k) * 2]), shift) * live * n_mpi_hosts;
        }
        for (int rc = 0; rc < 2; rc++)
          vr_out[((size_t)(Z - 1) * XMI_HIST_ROWS + (383 + rc)) * n_int + n] =
              fx_to_double(fx_get(&trc[(((size_t)n * nZ + u) * 2 + rc) * 2]), shift) * live * n_mpi_hosts;
      }
    }
    ok = 1;
  }
fail:
  if (graph_exec || graph) {
    /* replays may still run when a later step failed */
    (void)hipStreamSynchronize(E.stream);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  for (int k = 0; k < 2; k++)
    if (evs[k]) (void)hipEventDestroy(evs[k]);
  if (!ok) {
    free(ch_out);
    free(vr_out);
    free(bf_out);
    return 0;
  }
  if (esc) {
    g_stats.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall0).count();
    return 1;
  }
  *channelsPtr = ch_out;
  *var_red_historyPtr = vr_out;
  *brute_historyPtr = bf_out;
  g_stats.total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall0).count();
  return 1;
}

/* ------------------------------------------------------------------------ */
/* solid-angle grid (xmi_solid_angle_inputs_f, src/xmi_solid_angle_f.F90:62-301,
 * then the per-cell estimate of :432-710 on the GPU) */

/* ------------------------------------------------------------------------ */
/* escape ratios of the detector crystal (src/xmi_detector.c:91-141 builds
 * the input, src/xmi_main.F90:5473-5801 runs it): the crystal as a sample
 * at normal incidence of a pencil beam, one forced interaction per photon */

extern "C" xmi_escape_ratios_options xmi_get_default_escape_ratios_options(void) {
  xmi_escape_ratios_options o;
  o.n_input_energies = 1990;
  o.n_compton_output_energies = 1999;
  o.n_photons = 500000;
  o.input_energy_min = 1.0;
  o.input_energy_delta = 0.1;
  o.compton_output_energy_min = 0.1;
  o.compton_output_energy_delta = 0.1;
  return o;
}

extern "C" void xmi_free_escape_ratios(xmi_escape_ratios *r) {
  if (!r) return;
  free(r->Z);
  free(r->fluo_escape_ratios);
  free(r->fluo_escape_input_energies); /* shared with compton_escape_input_energies */
  free(r->compton_escape_ratios);
  free(r->compton_escape_output_energies);
  free(r->xmi_input_string);
  free(r);
}

extern "C" void xmi_escape_ratios_calculation(xmi_input *input, xmi_escape_ratios **escape_ratios, char *input_string,
                                              char *hdf5_file, xmi_main_options *options,
                                              xmi_escape_ratios_options ero) {
  if (escape_ratios) *escape_ratios = nullptr;
  if (!input || !escape_ratios || !options || !input->general || !input->geometry || !input->detector ||
      input->detector->n_crystal_layers < 1 || ero.n_input_energies < 1 || ero.n_compton_output_energies < 1 ||
      ero.n_photons < 1) {
    xmi_set_error("xmi_escape_ratios_calculation: invalid argument");
    return;
  }
  /* the escape input: crystal layers as the composition, reference layer 1,
   * source 1 cm from the sample behind a 1 um slit, sample normal along z */
  xmi_general gen = *input->general;
  gen.n_interactions_trajectory = 1;
  gen.n_photons_line = 100000;
  xmi_composition comp;
  comp.n_layers = input->detector->n_crystal_layers;
  comp.layers = input->detector->crystal_layers;
  comp.reference_layer = 1;
  xmi_geometry geo = *input->geometry;
  geo.d_sample_source = 1.0;
  geo.d_source_slit = 1.0;
  geo.slit_size_x = 0.0001;
  geo.slit_size_y = 0.0001;
  geo.n_sample_orientation[0] = 0.0;
  geo.n_sample_orientation[1] = 0.0;
  geo.n_sample_orientation[2] = 1.0;
  xmi_input esc_in = *input;
  esc_in.general = &gen;
  esc_in.composition = &comp;
  esc_in.geometry = &geo;
  xmi_main_options opts = *options;
  opts.use_M_lines = 0;
  opts.use_cascade_auger = 0;
  opts.use_cascade_radiative = 0;
  opts.use_variance_reduction = 0;
  opts.use_sum_peaks = 0;
  opts.escape_ratios_mode = 1;
  opts.use_advanced_compton = 0;

  xmi_inputFPtr inF = nullptr;
  xmi_hdf5FPtr h5 = nullptr;
  EscapeRun run;
  run.nE = (int)ero.n_input_energies;
  run.nOut = (int)ero.n_compton_output_energies;
  run.n_photons = ero.n_photons;
  run.E0 = ero.input_energy_min;
  run.dE = ero.input_energy_delta;
  run.out0 = ero.compton_output_energy_min;
  run.dOut = ero.compton_output_energy_delta;
  int ok = xmi_input_C2F(&esc_in, &inF) && xmi_init_input(&inF) &&
           (hdf5_file ? xmi_init_from_hdf5(hdf5_file, inF, &h5, &opts) : xmi_init_from_provider(inF, &h5, &opts)) &&
           xmi_update_input_from_hdf5(inF, h5);
  if (ok) {
    std::lock_guard<std::mutex> guard(g_engine.lock);
    ok = msim_run(g_engine, inF, h5, 1, nullptr, &opts, nullptr, nullptr, nullptr, &run);
  }
  xmi_escape_ratios *r = nullptr;
  if (ok) {
    const xmi_hdf5F *h = (const xmi_hdf5F *)h5;
    const int nZ = h->n_Z;
    r = (xmi_escape_ratios *)calloc(1, sizeof(xmi_escape_ratios));
    r->n_elements = nZ;
    r->n_fluo_input_energies = run.nE;
    r->n_compton_input_energies = run.nE;
    r->n_compton_output_energies = run.nOut;
    r->Z = (int *)malloc(sizeof(int) * nZ);
    for (int u = 0; u < nZ; u++) r->Z[u] = h->Zs[u].Z;
    r->fluo_escape_ratios = (double *)malloc(sizeof(double) * run.fluo.size());
    memcpy(r->fluo_escape_ratios, run.fluo.data(), sizeof(double) * run.fluo.size());
    r->compton_escape_ratios = (double *)malloc(sizeof(double) * run.compton.size());
    memcpy(r->compton_escape_ratios, run.compton.data(), sizeof(double) * run.compton.size());
    r->fluo_escape_input_energies = (double *)malloc(sizeof(double) * run.nE);
    for (int k = 0; k < run.nE; k++) r->fluo_escape_input_energies[k] = run.E0 + k * run.dE;
    r->compton_escape_input_energies = r->fluo_escape_input_energies;
    r->compton_escape_output_energies = (double *)malloc(sizeof(double) * run.nOut);
    for (int k = 0; k < run.nOut; k++) r->compton_escape_output_energies[k] = run.out0 + k * run.dOut;
    r->xmi_input_string = input_string ? strdup(input_string) : nullptr;
  }
  if (h5) xmi_free_hdf5_F(&h5);
  if (inF) xmi_free_input_F(&inF);
  *escape_ratios = r;
}

extern "C" int xmi_main_msim(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr, int n_mpi_hosts, double **channelsPtr,
                             xmi_main_options *options, double **brute_historyPtr, double **var_red_historyPtr,
                             xmi_solid_angle *solid_angles) {
  std::lock_guard<std::mutex> guard(g_engine.lock);
  return msim_run(g_engine, inputFPtr, hdf5FPtr, n_mpi_hosts, channelsPtr, options, brute_historyPtr,
                  var_red_historyPtr, solid_angles);
}

/* ------------------------------------------------------------------------ */
/* many samples: a work queue of samples served by several engines, each with
 * its own HIP stream and buffers, so that one sample's host set-up and small
 * launches overlap other samples' kernels; the device tables of the shared
 * table handle are uploaded once */

static std::mutex g_pool_lock;
static std::vector<Engine *> g_pool;

static double sample_histories(const xmi_inputF *in, int n_mpi_hosts) {
  double n = (double)in->general.n_photons_line / n_mpi_hosts * in->n_discrete;
  if (in->n_continuous > 1) n += (double)in->general.n_photons_interval / n_mpi_hosts * (in->n_continuous - 1);
  return n;
}

extern "C" int xmi_main_msim_batch(int n_samples, xmi_inputFPtr *inputs, xmi_hdf5FPtr hdf5FPtr, int n_mpi_hosts,
                                   xmi_main_options *options, xmi_solid_angle **solid_angles, double **channels,
                                   double **var_red_history, int n_streams, double *done_ms,
                                   const uint64_t *seeds) {
  if (n_samples < 0 || (n_samples > 0 && (!inputs || !hdf5FPtr || !options || !solid_angles || !channels ||
                                          !var_red_history))) {
    xmi_set_error("xmi_main_msim_batch: NULL argument");
    return 0;
  }
  if (n_streams < 1) n_streams = 1;
  if (n_streams > 8) n_streams = 8;
  if (n_mpi_hosts < 1) n_mpi_hosts = 1;
  {
    std::lock_guard<std::mutex> g(g_pool_lock);
    while ((int)g_pool.size() < n_streams) g_pool.push_back(new Engine());
  }
  /* longest first: the queue hands out the largest budgets before the small ones */
  std::vector<int> order(n_samples);
  for (int k = 0; k < n_samples; k++) {
    order[k] = k;
    channels[k] = nullptr;
    var_red_history[k] = nullptr;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return sample_histories((const xmi_inputF *)inputs[a], n_mpi_hosts) >
           sample_histories((const xmi_inputF *)inputs[b], n_mpi_hosts);
  });
  std::atomic<int> next(0), failed(0);
  std::atomic<int64_t> histories(0), vr_events(0);
  std::string first_error;
  std::mutex err_lock;
  const auto t0 = std::chrono::steady_clock::now();
  auto worker = [&](int w) {
    Engine &E = *g_pool[w];
    std::lock_guard<std::mutex> guard(E.lock);
    for (;;) {
      const int q = next.fetch_add(1);
      if (q >= n_samples || failed.load()) break;
      const int k = order[q];
      double *brute = nullptr;
      if (!msim_run(E, inputs[k], hdf5FPtr, n_mpi_hosts, &channels[k], options, &brute, &var_red_history[k],
                    solid_angles[k], nullptr, seeds ? &seeds[k] : nullptr)) {
        std::lock_guard<std::mutex> g(err_lock);
        if (!failed.exchange(1)) first_error = xmi_msim_last_error();
        break;
      }
      free(brute);
      histories += g_stats.histories;
      vr_events += g_stats.vr_events;
      if (done_ms)
        done_ms[k] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
  };
  std::vector<std::thread> threads;
  for (int w = 0; w < n_streams; w++) threads.emplace_back(worker, w);
  for (std::thread &t : threads) t.join();
  if (failed.load()) {
    for (int k = 0; k < n_samples; k++) {
      free(channels[k]);
      free(var_red_history[k]);
      channels[k] = var_red_history[k] = nullptr;
    }
    xmi_set_error("xmi_main_msim_batch: %s", first_error.c_str());
    return 0;
  }
  /* the calling thread's statistics: totals of the batch */
  memset(&g_stats, 0, sizeof(g_stats));
  g_stats.histories = histories.load();
  g_stats.vr_events = vr_events.load();
  g_stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 1;
}

static double depth_quantile(const xmi_inputF *in, double energy, double R) {
  int nl = in->n_layers;
  std::vector<double> mu(nl);
  double my_sum = 0.0;
  for (int i = 0; i < nl; i++) {
    mu[i] = layer_mu_host(&in->layers[i], energy);
    my_sum += mu[i] * in->layers[i].density * in->layers[i].thickness_along_Z;
  }
  double Pabs = -1.0 * expm1(-1.0 * my_sum);
  double myln = -1.0 * log1p(-1.0 * R * Pabs);
  int m = nl - 1;
  my_sum = 0.0;
  for (int i = 0; i < nl; i++) {
    my_sum += mu[i] * in->layers[i].density * in->layers[i].thickness_along_Z;
    if (my_sum > myln) {
      m = i;
      break;
    }
  }
  my_sum = 0.0;
  for (int i = 0; i < m; i++)
    my_sum += (1.0 - (mu[i] * in->layers[i].density) / (mu[m] * in->layers[m].density)) * in->layers[i].thickness_along_Z;
  return my_sum + myln / (mu[m] * in->layers[m].density) + in->layers[0].Z_coord_begin;
}

extern "C" void xmi_free_solid_angle(xmi_solid_angle *sa) {
  if (!sa) return;
  free(sa->solid_angles);
  free(sa->grid_dims_r_vals);
  free(sa->grid_dims_theta_vals);
  free(sa->xmi_input_string);
  free(sa);
}

extern "C" int xmi_solid_angle_calculation_cl(xmi_inputFPtr inputFPtr, xmi_solid_angle **out, char *input_string,
                                              xmi_main_options *xmo) {
  xmi_inputF *in = (xmi_inputF *)inputFPtr;
  if (!in || !out || !in->initialised) {
    xmi_set_error("xmi_solid_angle_calculation_cl: invalid input");
    return 0;
  }
  (void)xmo;
  const long n = 1024;
  double e1, e2;
  if (in->n_continuous > 1 && in->n_discrete > 0) {
    e1 = fmin(in->continuous[0].energy, in->discrete[0].energy);
    e2 = fmax(in->continuous[in->n_continuous - 1].energy, in->discrete[in->n_discrete - 1].energy);
  } else if (in->n_continuous > 1) {
    e1 = in->continuous[0].energy;
    e2 = in->continuous[in->n_continuous - 1].energy;
  } else {
    e1 = in->discrete[0].energy;
    e2 = in->discrete[in->n_discrete - 1].energy;
  }
  double S1 = depth_quantile(in, e1, 0.00001), S2 = depth_quantile(in, e2, 0.99999);
  const double *pd = in->geometry.p_detector_window;
  double d1 = sqrt(pd[0] * pd[0] + pd[1] * pd[1] + (pd[2] - S1) * (pd[2] - S1));
  double d2 = sqrt(pd[0] * pd[0] + pd[1] * pd[1] + (pd[2] - S2) * (pd[2] - S2));
  double rmax = (d1 > d2 ? d1 : d2) * 1.25, rmin = rmax / (double)n;
  double tmin = 0.00001, tmax = M_PI / 2.0;
  xmi_solid_angle *sa = (xmi_solid_angle *)calloc(1, sizeof(xmi_solid_angle));
  sa->grid_dims_r_n = n;
  sa->grid_dims_theta_n = n;
  sa->grid_dims_r_vals = (double *)malloc(sizeof(double) * n);
  sa->grid_dims_theta_vals = (double *)malloc(sizeof(double) * n);
  sa->solid_angles = (double *)malloc(sizeof(double) * n * n);
  sa->xmi_input_string = input_string;
  for (long i = 0; i < n; i++) {
    sa->grid_dims_r_vals[i] = rmin + (rmax - rmin) * (double)i / (double)(n - 1);
    sa->grid_dims_theta_vals[i] = tmin + (tmax - tmin) * (double)i / (double)(n - 1);
  }
  std::lock_guard<std::mutex> guard(g_engine.lock);
  if (!engine_init()) {
    xmi_free_solid_angle(sa);
    return 0;
  }
  EngineParams P;
  memset(&P, 0, sizeof(P));
  fill_geometry(P, in);
  uint64_t seed;
  if (xmo && xmo->use_default_seeds) seed = 1;
  else {
    if (!g_seed_set) {
      g_seed = (uint64_t)time(NULL) * 6364136223846793005ull + (uint64_t)clock();
      g_seed_set = 1;
    }
    seed = g_seed;
  }
  hipStream_t st = g_engine.stream;
  int ok = 0;
  if (upload(g_engine.sa_r, sa->grid_dims_r_vals, (size_t)n, g_engine.stream, g_engine.stage) == hipSuccess &&
      upload(g_engine.sa_t, sa->grid_dims_theta_vals, (size_t)n, g_engine.stream, g_engine.stage) == hipSuccess &&
      g_engine.sa_grid.ensure(sizeof(double) * n * n) == hipSuccess &&
      eng_launch_grid(&P, g_engine.sa_grid.as<double>(), g_engine.sa_r.as<double>(), g_engine.sa_t.as<double>(), (int)n,
                      (int)n, xmi_hits_per_single, seed, st) == hipSuccess &&
      copy_d2h(sa->solid_angles, g_engine.sa_grid.p, sizeof(double) * n * n, st, g_engine.stage) == hipSuccess)
    ok = 1;
  if (!ok) {
    xmi_set_error("xmi_solid_angle_calculation_cl: device failure");
    g_engine.sa_key = nullptr;
    xmi_free_solid_angle(sa);
    return 0;
  }
  /* the device copy is the grid just returned: remember it */
  g_engine.sa_key = sa;
  g_engine.sa_hash = grid_hash(sa);
  g_engine.sa_nr = (int)n;
  g_engine.sa_nt = (int)n;
  *out = sa;
  return 1;
}
