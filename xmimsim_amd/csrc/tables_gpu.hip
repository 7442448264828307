/*
 * tables_gpu.hip -- the heavy parts of the interaction tables on the GPU.
 *
 * The host generator (xmi_db.c, generator of src/xmi_data_f.F90:880-1551)
 * spends nearly all its time in four loops that are independent per output
 * value; here they run one thread per value:
 *   - angle inverse CDFs (src/xmi_data_f.F90:996-1073): one thread per
 *     (element, Rayleigh|Compton, energy), the trapezoid CDF of
 *     DCS(theta) sin(theta) on 10000 points walked twice (total, inversion);
 *   - total Compton-profile inverse CDF (src/xmi_data_f.F90:1131-1237): one
 *     thread per (element, random number), 100 bisection steps;
 *   - partial-profile inversions: one thread per (element, shell, point);
 *   - azimuth inverse CDF (src/xmi_data_f.F90:1483-1549): one thread per
 *     (t, random number).
 * Every thread runs the same routine of xrl_model.h as the host generator,
 * so the tables agree with the host's up to the rounding of the device's
 * math functions.  The host builds the light parts (grids, interaction
 * probabilities, yields, cascade cross sections) and allocates the arrays;
 * xmi_tables_fill_gpu fills the heavy ones.
 */
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "xmi_internal.h"

#define TG_NTHETA_INT 10000
#define TG_MAXPZ 100.0
#define TG_CHUNK 16 /* elements per launch of the angle kernel (device memory) */

__global__ void __launch_bounds__(256) k_theta_rows(const xrl_atom *atoms, int nz, const double *energies, int n_e,
                                                    const double *rs, int n_r, double *out /* [nz][2][n_e][n_r] */) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nz * 2 * n_e) return;
  const int z = t / (2 * n_e), kind = (t / n_e) & 1, j = t % n_e;
  xrl_theta_icdf_row(&atoms[z], kind, energies[j], rs, n_r, TG_NTHETA_INT, out + (size_t)t * n_r, 1);
}

__global__ void __launch_bounds__(256) k_profile_total(const xrl_atom *atoms, int nz, const double *cmax,
                                                       const double *rn, int n, double *out /* [nz][n] */) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nz * n) return;
  const int z = t / n, i = t % n;
  double v;
  if (i == 0) v = 0.0;
  else if (i == n - 1) v = TG_MAXPZ;
  else v = xrl_bisect_total_profile(&atoms[z], rn[i] * cmax[z], 0.0, TG_MAXPZ);
  out[t] = v;
}

/* rows: (profile momentum p, CDF at the end gmax) of every (element, shell) */
__global__ void __launch_bounds__(256) k_profile_partial(const double2 *rows, int n_rows, const double *cinv, int n,
                                                         double *out /* [n_rows][n] */) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_rows * n) return;
  const int r = t / n, i = t % n;
  const double p = rows[r].x, gmax = rows[r].y;
  double v;
  if (i == 0) v = 0.0;
  else if (i == n - 1) v = TG_MAXPZ;
  else v = xrl_bisect_G3(cinv[i] / 0.5 * gmax, p, 0.0, TG_MAXPZ);
  out[t] = v;
}

__global__ void __launch_bounds__(256) k_phi(const double *thetas, int n_t, const double *rs, int n_r,
                                             double *out /* [n_t][n_r] */) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_t * n_r) return;
  const int i = t / n_r, k = t % n_r;
  double v;
  if (k == 0) v = 0.0;
  else if (k == n_r - 1) v = 2.0 * 3.14159265358979323846;
  else v = xrl_bisect_phi(thetas[i], rs[k] * 2.0 * 3.14159265358979323846);
  out[t] = v;
}

static int g_tables_mode = -1; /* -1 device when one is present, 0 host, 1 device */

extern "C" void xmi_tables_set_device(int mode) { g_tables_mode = mode < 0 ? -1 : (mode ? 1 : 0); }

/* 1: fill on the device, 0: on the host, -1: a device is demanded and none
 * is usable */
extern "C" int xmi_tables_use_device(void) {
  if (g_tables_mode == 0) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    (void)hipGetLastError();
    return g_tables_mode == 1 ? -1 : 0;
  }
  return 1;
}

namespace {
struct DBuf {
  void *p = nullptr;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
};
}  // namespace

#define TG_TRY(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      xmi_set_error("table generator: %s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__); \
      return 0;                                                                    \
    }                                                                              \
  } while (0)

static int blocks_for(size_t n) { return (int)((n + 255) / 256); }

/* fills RayleighTheta_ICDF, ComptonTheta_ICDF, cp_total_icdf and cp_Qs_inv
 * of zs[0..n-1] (built by xmi_tables_build_Z with heavy = 0, on the shared
 * grids) and, when h is given, its Phi_ICDF; returns 1 on success */
extern "C" int xmi_tables_fill_gpu(xmi_hdf5_ZF *zs, int n, xmi_hdf5F *h) {
  if (n > 0) {
    const int n_e = zs[0].n_e, n_r = zs[0].n_r, ncp = zs[0].cp_len;
    for (int i = 0; i < n; i++)
      if (zs[i].n_e != n_e || zs[i].n_r != n_r || zs[i].cp_len != ncp) {
        xmi_set_error("table generator: elements on different grids");
        return 0;
      }
    std::vector<xrl_atom> atoms(n);
    std::vector<double> cmax(n);
    for (int i = 0; i < n; i++) {
      const xrl_atom *a = xrl_atom_get(zs[i].Z);
      if (!a) {
        xmi_set_error("no physics data for Z=%d", zs[i].Z);
        return 0;
      }
      atoms[i] = *a;
      cmax[i] = xrl_total_profile_cdf(a, TG_MAXPZ);
    }
    DBuf d_atoms, d_E, d_rs, d_cmax, d_rn, d_cinv;
    TG_TRY(d_atoms.alloc(sizeof(xrl_atom) * n));
    TG_TRY(hipMemcpy(d_atoms.p, atoms.data(), sizeof(xrl_atom) * n, hipMemcpyHostToDevice));
    TG_TRY(d_E.alloc(sizeof(double) * n_e));
    TG_TRY(hipMemcpy(d_E.p, zs[0].Energies, sizeof(double) * n_e, hipMemcpyHostToDevice));
    TG_TRY(d_rs.alloc(sizeof(double) * n_r));
    TG_TRY(hipMemcpy(d_rs.p, zs[0].RandomNumbers, sizeof(double) * n_r, hipMemcpyHostToDevice));
    /* angle inverse CDFs, TG_CHUNK elements per launch */
    {
      const int chunk = n < TG_CHUNK ? n : TG_CHUNK;
      const size_t row = (size_t)n_e * n_r;
      DBuf d_out;
      TG_TRY(d_out.alloc(sizeof(double) * chunk * 2 * row));
      for (int z0 = 0; z0 < n; z0 += chunk) {
        const int nz = (n - z0) < chunk ? (n - z0) : chunk;
        hipLaunchKernelGGL(k_theta_rows, dim3(blocks_for((size_t)nz * 2 * n_e)), dim3(256), 0, 0,
                           d_atoms.as<xrl_atom>() + z0, nz, d_E.as<double>(), n_e, d_rs.as<double>(), n_r,
                           d_out.as<double>());
        TG_TRY(hipGetLastError());
        for (int z = 0; z < nz; z++) {
          TG_TRY(hipMemcpy(zs[z0 + z].RayleighTheta_ICDF, d_out.as<double>() + (size_t)(2 * z) * row,
                           sizeof(double) * row, hipMemcpyDeviceToHost));
          TG_TRY(hipMemcpy(zs[z0 + z].ComptonTheta_ICDF, d_out.as<double>() + (size_t)(2 * z + 1) * row,
                           sizeof(double) * row, hipMemcpyDeviceToHost));
        }
      }
    }
    /* total-profile inverse CDF */
    {
      TG_TRY(d_cmax.alloc(sizeof(double) * n));
      TG_TRY(hipMemcpy(d_cmax.p, cmax.data(), sizeof(double) * n, hipMemcpyHostToDevice));
      TG_TRY(d_rn.alloc(sizeof(double) * ncp));
      TG_TRY(hipMemcpy(d_rn.p, zs[0].cp_random_numbers, sizeof(double) * ncp, hipMemcpyHostToDevice));
      DBuf d_out;
      TG_TRY(d_out.alloc(sizeof(double) * n * ncp));
      hipLaunchKernelGGL(k_profile_total, dim3(blocks_for((size_t)n * ncp)), dim3(256), 0, 0,
                         d_atoms.as<xrl_atom>(), n, d_cmax.as<double>(), d_rn.as<double>(), ncp, d_out.as<double>());
      TG_TRY(hipGetLastError());
      for (int z = 0; z < n; z++)
        TG_TRY(hipMemcpy(zs[z].cp_total_icdf, d_out.as<double>() + (size_t)z * ncp, sizeof(double) * ncp,
                         hipMemcpyDeviceToHost));
    }
    /* partial-profile inversions of every (element, shell) */
    {
      std::vector<double2> rows;
      for (int z = 0; z < n; z++) {
        const xrl_atom *a = xrl_atom_get(zs[z].Z);
        for (int k = 0; k < zs[z].cp_nshells; k++) {
          const int s = zs[z].cp_shell_indices[k];
          const double p = a->profile_p[s] > 0.0 ? a->profile_p[s] : 1.0;
          rows.push_back(make_double2(p, xrl_profile_G3(TG_MAXPZ / p)));
        }
      }
      DBuf d_rows, d_out;
      TG_TRY(d_rows.alloc(sizeof(double2) * rows.size()));
      TG_TRY(hipMemcpy(d_rows.p, rows.data(), sizeof(double2) * rows.size(), hipMemcpyHostToDevice));
      TG_TRY(d_cinv.alloc(sizeof(double) * ncp));
      TG_TRY(hipMemcpy(d_cinv.p, zs[0].cp_partial_cdf_inv, sizeof(double) * ncp, hipMemcpyHostToDevice));
      TG_TRY(d_out.alloc(sizeof(double) * rows.size() * ncp));
      hipLaunchKernelGGL(k_profile_partial, dim3(blocks_for(rows.size() * ncp)), dim3(256), 0, 0,
                         d_rows.as<double2>(), (int)rows.size(), d_cinv.as<double>(), ncp, d_out.as<double>());
      TG_TRY(hipGetLastError());
      size_t r0 = 0;
      for (int z = 0; z < n; z++) {
        TG_TRY(hipMemcpy(zs[z].cp_Qs_inv, d_out.as<double>() + r0 * ncp, sizeof(double) * zs[z].cp_nshells * ncp,
                         hipMemcpyDeviceToHost));
        r0 += zs[z].cp_nshells;
      }
    }
  }
  if (h) {
    DBuf d_t, d_rs, d_out;
    const int n_t = h->phi_n_theta, n_r = h->phi_n_r;
    TG_TRY(d_t.alloc(sizeof(double) * n_t));
    TG_TRY(hipMemcpy(d_t.p, h->Thetas, sizeof(double) * n_t, hipMemcpyHostToDevice));
    TG_TRY(d_rs.alloc(sizeof(double) * n_r));
    TG_TRY(hipMemcpy(d_rs.p, h->RandomNumbers, sizeof(double) * n_r, hipMemcpyHostToDevice));
    TG_TRY(d_out.alloc(sizeof(double) * n_t * n_r));
    hipLaunchKernelGGL(k_phi, dim3(blocks_for((size_t)n_t * n_r)), dim3(256), 0, 0, d_t.as<double>(), n_t,
                       d_rs.as<double>(), n_r, d_out.as<double>());
    TG_TRY(hipGetLastError());
    TG_TRY(hipMemcpy(h->Phi_ICDF, d_out.p, sizeof(double) * n_t * n_r, hipMemcpyDeviceToHost));
  }
  TG_TRY(hipDeviceSynchronize());
  return 1;
}
