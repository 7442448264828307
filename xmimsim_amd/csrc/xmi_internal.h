/*
 * xmi_internal.h -- host-side working copies behind the opaque handles of the
 * C ABI (the equivalents of the Fortran types xmi_input / xmi_hdf5 /
 * xmi_hdf5_Z of src/xmi_aux_f.F90:43-563).  Shared by the engine, the table
 * loader and the oracle.
 */
#ifndef XMIMSIM_AMD_INTERNAL_H
#define XMIMSIM_AMD_INTERNAL_H

#include "../../include/xmimsim_amd.h"
#include "xrl_model.h"

#ifdef __cplusplus
extern "C" {
#endif

#define XMI_MAXZ 94
#define XMI_NLINES_VR 219          /* KL1..M5P5 */
#define XMI_HIST_ROWS 385          /* 383 lines + Rayleigh + Compton */
#define XMI_ESC_LINES 109          /* escape-ratio lines KL1..L3P3 */
#define XMI_ENERGY_THRESHOLD 1.0   /* keV, src/xmi_aux_f.F90:672 */
#define XMI_ENERGY_MAX 200.0       /* keV, src/xmi_aux_f.F90:673 */

/* interaction codes, src/xmi_aux_f.F90:653-656 */
#define XMI_NO_INTERACTION 0
#define XMI_RAYLEIGH_INTERACTION 1
#define XMI_COMPTON_INTERACTION 2
#define XMI_PHOTOELECTRIC_INTERACTION 3

/* reference table grids, src/xmi_data.c:150-152 and src/xmi_data_f.F90:904-969 */
#define XMI_DB_NR 2000
#define XMI_DB_NE 400
#define XMI_DB_NTHETA2 200
#define XMI_DB_NE_IP 20000
#define XMI_DB_NCP 10001

typedef struct xmi_layerF {
  int n_elements;
  int *Z;
  double *weight;          /* normalised */
  double density;
  double thickness;
  double thickness_along_Z;
  double Z_coord_begin;
  double Z_coord_end;
  int *hdf5_index;         /* 0-based index into xmi_hdf5F.Zs per element */
} xmi_layerF;

typedef struct xmi_inputF {
  xmi_general general;     /* strings not owned */
  int n_layers;
  xmi_layerF *layers;
  int reference_layer;     /* 1-based */
  xmi_geometry geometry;
  int n_discrete;
  xmi_energy_discrete *discrete;
  int n_continuous;
  xmi_energy_continuous *continuous;
  int n_exc_layers;
  xmi_layerF *exc_layers;
  int n_det_layers;
  xmi_layerF *det_layers;
  xmi_detector detector;   /* crystal_layers not owned */
  int n_crystal_layers;
  xmi_layerF *crystal_layers;
  /* derived by xmi_init_input (src/xmi_aux_f.F90:285-309) */
  int collimator_present;
  double detector_radius;
  double collimator_radius;
  double half_apex;
  double vertex[3];
  double n_detector_orientation_new_x[3];
  double n_detector_orientation_new_y[3];
  double n_detector_orientation_new_z[3];
  double n_detector_orientation_new[3][3];      /* [row][col], columns = new x,y,z */
  double n_detector_orientation_inverse[3][3];  /* [row][col] */
  double detector_solid_angle;
  double n_sample_orientation_det[3];
  int initialised;
} xmi_inputF;

typedef struct xmi_hdf5_ZF {
  int Z;
  int Zindex;                     /* 1-based position in the sorted unique list */
  int n_e, n_r;
  double *Energies;               /* [n_e] */
  double *RandomNumbers;          /* [n_r] */
  double *RayleighTheta_ICDF;     /* [n_e][n_r]  (r contiguous) */
  double *ComptonTheta_ICDF;      /* [n_e][n_r] */
  double FluorYieldsCorr[9];
  int ip_len;
  double *ip_energies;            /* [ip_len] */
  double *ip_rayl;                /* cumulative Rayleigh fraction [ip_len] */
  double *ip_rayl_compt;          /* cumulative Rayleigh+Compton fraction */
  int cp_nshells;
  int cp_len;
  int *cp_shell_indices;          /* [cp_nshells] */
  double *cp_Qs;                  /* [cp_len] */
  double *cp_partial_cdf;         /* [cp_nshells][cp_len] */
  double *cp_partial_cdf_inv;     /* [cp_len] */
  double *cp_Qs_inv;              /* [cp_nshells][cp_len] */
  double *cp_random_numbers;      /* [cp_len] */
  double *cp_total_icdf;          /* [cp_len] */
  double *precalc_xrf_cs;         /* [9][n_uniq][XMI_NLINES_VR] for the selected cascade type */
} xmi_hdf5_ZF;

typedef struct xmi_hdf5F {
  int n_Z;
  xmi_hdf5_ZF *Zs;
  int phi_n_theta, phi_n_r;
  double *Phi_ICDF;               /* [n_theta][n_r] */
  double *Thetas;                 /* [n_theta] */
  double *RandomNumbers;          /* [n_r] */
  int uniqZ[XMI_MAXZ + 1];        /* Z -> 1-based index, 0 when absent */
  int cascade_type;
  int use_M_lines;
} xmi_hdf5F;

/* table builders (xmi_db.c) */
int xmi_tables_build_Z(int Z, const int *uniqZ, int n_uniq, int cascade_type, int heavy, xmi_hdf5_ZF *out);
int xmi_tables_build_phi(xmi_hdf5F *h, int heavy);
/* the heavy loops on the GPU (tables_gpu.hip) */
int xmi_tables_fill_gpu(xmi_hdf5_ZF *zs, int n, xmi_hdf5F *h);
int xmi_tables_use_device(void);
void xmi_tables_free(xmi_hdf5F *h);

/* The CDF of the Compton-profile inverse table (rn uniform in [0, 1], icdf
 * increasing, n entries) at m equally spaced momenta 0 .. icdf[n-2], the
 * range the estimator draws from (the last interval excluded): out[j] =
 * r at which the piecewise-linear icdf reaches j * icdf[n-2] / (m - 1).  The
 * detection-weight estimator reads it to draw Doppler momenta from the
 * truncated profile directly instead of by rejection. */
static inline void xmi_profile_cdf_grid(const double *rn, const double *icdf, int n, int m, double *out) {
  const double top = icdf[n - 2];
  int k = 0;
  for (int j = 0; j < m; j++) {
    const double pz = top * j / (double)(m - 1);
    while (k < n - 3 && icdf[k + 1] < pz) k++;
    const double d = icdf[k + 1] - icdf[k];
    double r = d > 0.0 ? rn[k] + (pz - icdf[k]) / d * (rn[k + 1] - rn[k]) : rn[k + 1];
    if (r > rn[n - 2]) r = rn[n - 2];
    out[j] = r;
  }
  out[0] = 0.0;
  out[m - 1] = rn[n - 2];
}

/* geometry helpers (xmi_input.c) */
void xmi_normalize3(double v[3]);
int xmi_unique_Z(const xmi_inputF *in, int *uniqZ);  /* sorted, returns count */

/* error reporting */
void xmi_set_error(const char *fmt, ...);

#ifdef __cplusplus
}
#endif
#endif
