/*
 * xmi_db.c -- interaction tables of the photon-history engine.
 *
 * Builds, writes and reads the per-element tables that xmi_main_msim samples
 * from, in the layout of the reference's xmimsimdata.h5:
 *
 *   /"%2i"/Theta_ICDF/{RayleighTheta_ICDF[2000][400], ComptonTheta_ICDF,
 *                       Energies[400], Random numbers[2000],
 *                       Corrected fluorescence yields[9]}
 *   /"%2i"/Interaction probabilities/{Energies[n],
 *                       Rayleigh and Compton probabilities[2][n]}
 *   /"%2i"/Compton profiles/{Shell indices, Qs, Partial profile CDF,
 *                       Partial profile CDF inverted, Qs inverted,
 *                       Random numbers, Total profile ICDF}
 *   /"%2i"/Precalculated XRF cross sections/<cascade>/<shell>/"%2i"[383]
 *   /Phi/{Phi_ICDF[2000][200], Thetas[200], Random numbers[2000]}
 *
 * (writer src/xmi_data.c:149-486, generator src/xmi_data_f.F90:880-1551,
 * reader src/xmi_data_f.F90:98-848 and src/xmi_data.c:489-647).
 *
 * Generation differs from the reference in method, not in meaning: the angle
 * inverse CDFs are obtained by linear inversion of a trapezoid-integrated CDF
 * (the reference picks the first grid angle whose cumulative sum passes r), and
 * the Compton-profile and azimuth inverse CDFs are inverted by bisection on the
 * closed-form CDFs of the provider instead of 1e7-point sums.
 */
#include <hdf5.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xmi_internal.h"

#define LOWE 0.1
#define MAXE 200.0
#define NTHETA_INT 10000
#define MAXPZ 100.0

static const char *cascade_group_names[4] = {"No cascade effect", "Non-radiative cascade effect",
                                             "Radiative cascade effect", "Full cascade effect"};
static const char *shell_group_names[9] = {"K shell",  "L1 shell", "L2 shell", "L3 shell", "M1 shell",
                                           "M2 shell", "M3 shell", "M4 shell", "M5 shell"};

/* ------------------------------------------------------------------------ */
/* angle ICDFs (src/xmi_data_f.F90:996-1073), one row per energy */
static void theta_icdf(const xrl_atom *a, int compton, const double *energies, int n_e, const double *rs,
                       int n_r, double *out /* [n_e][n_r] */) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j < n_e; j++)
    xrl_theta_icdf_row(a, compton, energies[j], rs, n_r, NTHETA_INT, out + (size_t)j * n_r, 1);
}

/* heavy = 0 leaves the angle inverse CDFs, the total-profile inverse CDF and
 * the partial-profile inversions (cp_Qs_inv, except its ends) to the GPU
 * generator (xmi_tables_fill_gpu); the arrays are allocated either way */
int xmi_tables_build_Z(int Z, const int *uniqZ, int n_uniq, int cascade_type, int heavy, xmi_hdf5_ZF *out) {
  const xrl_atom *a = xrl_atom_get(Z);
  if (!a) {
    xmi_set_error("no physics data for Z=%d", Z);
    return 0;
  }
  memset(out, 0, sizeof(*out));
  out->Z = Z;
  out->n_e = XMI_DB_NE;
  out->n_r = XMI_DB_NR;
  out->Energies = (double *)malloc(sizeof(double) * XMI_DB_NE);
  out->RandomNumbers = (double *)malloc(sizeof(double) * XMI_DB_NR);
  for (int i = 0; i < XMI_DB_NE; i++) out->Energies[i] = LOWE + (MAXE - LOWE) * i / (double)(XMI_DB_NE - 1);
  for (int i = 0; i < XMI_DB_NR; i++) out->RandomNumbers[i] = i / (double)(XMI_DB_NR - 1);
  out->RayleighTheta_ICDF = (double *)malloc(sizeof(double) * XMI_DB_NE * XMI_DB_NR);
  out->ComptonTheta_ICDF = (double *)malloc(sizeof(double) * XMI_DB_NE * XMI_DB_NR);
  if (heavy) {
    theta_icdf(a, 0, out->Energies, XMI_DB_NE, out->RandomNumbers, XMI_DB_NR, out->RayleighTheta_ICDF);
    theta_icdf(a, 1, out->Energies, XMI_DB_NE, out->RandomNumbers, XMI_DB_NR, out->ComptonTheta_ICDF);
  }

  /* interaction probabilities (src/xmi_data_f.F90:1084-1125) */
  {
    int n = XMI_DB_NE_IP;
    double *e = (double *)malloc(sizeof(double) * (XMI_DB_NE_IP + 18));
    for (int j = 0; j < XMI_DB_NE_IP; j++) e[j] = LOWE + (MAXE - LOWE) * j / (double)(XMI_DB_NE_IP - 1);
    for (int s = 0; s < 9; s++) {
      double ed = a->edge[s];
      if (ed < LOWE || ed <= 0.0) continue;
      e[n++] = ed + 0.00001;
      e[n++] = ed - 0.00001;
    }
    /* insertion sort of the few appended edge points */
    for (int i = XMI_DB_NE_IP; i < n; i++) {
      double v = e[i];
      int j = i - 1;
      while (j >= 0 && e[j] > v) {
        e[j + 1] = e[j];
        j--;
      }
      e[j + 1] = v;
    }
    out->ip_len = n;
    out->ip_energies = e;
    out->ip_rayl = (double *)malloc(sizeof(double) * n);
    out->ip_rayl_compt = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n; k++) {
      double tot = xrl_cs_total(a, e[k]);
      out->ip_rayl[k] = xrl_cs_rayl(a, e[k]) / tot;
      out->ip_rayl_compt[k] = xrl_cs_compt(a, e[k]) / tot + out->ip_rayl[k];
    }
  }

  /* Compton profiles (src/xmi_data_f.F90:1131-1237) */
  {
    int nsh = 0;
    int idx[XRL_NSHELLS];
    idx[nsh++] = 0;
    for (int s = 1; s < XRL_NSHELLS; s++)
      if (a->occupancy[s] > 0.0) idx[nsh++] = s;
    const int N = XMI_DB_NCP;
    out->cp_nshells = nsh;
    out->cp_len = N;
    out->cp_shell_indices = (int *)malloc(sizeof(int) * nsh);
    memcpy(out->cp_shell_indices, idx, sizeof(int) * nsh);
    out->cp_Qs = (double *)malloc(sizeof(double) * N);
    out->cp_partial_cdf_inv = (double *)malloc(sizeof(double) * N);
    out->cp_random_numbers = (double *)malloc(sizeof(double) * N);
    out->cp_total_icdf = (double *)malloc(sizeof(double) * N);
    out->cp_partial_cdf = (double *)malloc(sizeof(double) * N * nsh);
    out->cp_Qs_inv = (double *)malloc(sizeof(double) * N * nsh);
    for (int i = 0; i < N; i++) {
      out->cp_Qs[i] = MAXPZ * i / (double)(N - 1);
      out->cp_partial_cdf_inv[i] = 0.5 * i / (double)(N - 1);
      out->cp_random_numbers[i] = i / (double)(N - 1);
    }
    double cmax = xrl_total_profile_cdf(a, MAXPZ);
    if (heavy)
      for (int i = 0; i < N; i++)
        out->cp_total_icdf[i] = xrl_bisect_total_profile(a, out->cp_random_numbers[i] * cmax, 0.0, MAXPZ);
    out->cp_total_icdf[0] = 0.0;
    out->cp_total_icdf[N - 1] = MAXPZ;
    for (int k = 0; k < nsh; k++) {
      int s = idx[k];
      double p = a->profile_p[s] > 0.0 ? a->profile_p[s] : 1.0;
      double gmax = xrl_profile_G3(MAXPZ / p);
      for (int i = 0; i < N; i++) {
        out->cp_partial_cdf[(size_t)k * N + i] = 0.5 * xrl_profile_G3(out->cp_Qs[i] / p) / gmax;
        double target = out->cp_partial_cdf_inv[i] / 0.5 * gmax;
        out->cp_Qs_inv[(size_t)k * N + i] = (i == N - 1) ? MAXPZ : heavy ? xrl_bisect_G3(target, p, 0.0, MAXPZ) : 0.0;
      }
      out->cp_Qs_inv[(size_t)k * N] = 0.0;
    }
  }

  xrl_fluor_yields_corr(a, out->FluorYieldsCorr);

  /* precalculated XRF cross sections for the selected cascade type
   * (src/xmi_data_f.F90:1307-1478): P_shell of this element at the line
   * energies of every element of the sample */
  if (uniqZ && n_uniq > 0) {
    out->precalc_xrf_cs = (double *)calloc((size_t)9 * n_uniq * XMI_NLINES_VR, sizeof(double));
    for (int j = 0; j < n_uniq; j++) {
      const xrl_atom *a2 = xrl_atom_get(uniqZ[j]);
      for (int line = 1; line <= XMI_NLINES_VR; line++) {
        double P[9];
        xrl_shell_cross_sections(a, cascade_type, a2->line_energy[line], P);
        for (int s = 0; s < 9; s++)
          out->precalc_xrf_cs[((size_t)s * n_uniq + j) * XMI_NLINES_VR + (line - 1)] = P[s];
      }
    }
  }
  return 1;
}

/* azimuth ICDF, CDF(phi) = (phi - t sin(2 phi)) / 2pi (src/xmi_data_f.F90:1483-1549) */
int xmi_tables_build_phi(xmi_hdf5F *h, int heavy) {
  h->phi_n_theta = XMI_DB_NTHETA2;
  h->phi_n_r = XMI_DB_NR;
  h->Thetas = (double *)malloc(sizeof(double) * XMI_DB_NTHETA2);
  h->RandomNumbers = (double *)malloc(sizeof(double) * XMI_DB_NR);
  h->Phi_ICDF = (double *)malloc(sizeof(double) * XMI_DB_NTHETA2 * XMI_DB_NR);
  for (int i = 0; i < XMI_DB_NTHETA2; i++) h->Thetas[i] = 0.5 * i / (double)(XMI_DB_NTHETA2 - 1);
  for (int i = 0; i < XMI_DB_NR; i++) h->RandomNumbers[i] = i / (double)(XMI_DB_NR - 1);
  if (heavy)
    for (int i = 0; i < XMI_DB_NTHETA2; i++)
      for (int k = 0; k < XMI_DB_NR; k++)
        h->Phi_ICDF[(size_t)i * XMI_DB_NR + k] = xrl_bisect_phi(h->Thetas[i], h->RandomNumbers[k] * 2.0 * M_PI);
  for (int i = 0; i < XMI_DB_NTHETA2; i++) {
    h->Phi_ICDF[(size_t)i * XMI_DB_NR] = 0.0;
    h->Phi_ICDF[(size_t)i * XMI_DB_NR + XMI_DB_NR - 1] = 2.0 * M_PI;
  }
  return 1;
}

static void free_Z(xmi_hdf5_ZF *z) {
  free(z->Energies);
  free(z->RandomNumbers);
  free(z->RayleighTheta_ICDF);
  free(z->ComptonTheta_ICDF);
  free(z->ip_energies);
  free(z->ip_rayl);
  free(z->ip_rayl_compt);
  free(z->cp_shell_indices);
  free(z->cp_Qs);
  free(z->cp_partial_cdf);
  free(z->cp_partial_cdf_inv);
  free(z->cp_Qs_inv);
  free(z->cp_random_numbers);
  free(z->cp_total_icdf);
  free(z->precalc_xrf_cs);
}

/* the engine keeps device copies of the tables per handle; it provides this
 * hook (weak, so that host-only users of the table code link without it) */
void xmi_engine_forget_tables(const void *h) __attribute__((weak));

void xmi_tables_free(xmi_hdf5F *h) {
  if (!h) return;
  if (xmi_engine_forget_tables) xmi_engine_forget_tables(h);
  for (int i = 0; i < h->n_Z; i++) free_Z(&h->Zs[i]);
  free(h->Zs);
  free(h->Phi_ICDF);
  free(h->Thetas);
  free(h->RandomNumbers);
  free(h);
}

void xmi_free_hdf5_F(xmi_hdf5FPtr *hdf5FPtr) {
  if (!hdf5FPtr || !*hdf5FPtr) return;
  xmi_tables_free((xmi_hdf5F *)*hdf5FPtr);
  *hdf5FPtr = NULL;
}

static int cascade_type_of(const xmi_main_options *o) {
  if (!o->use_cascade_auger && !o->use_cascade_radiative) return XRL_CASCADE_NONE;
  if (o->use_cascade_auger && !o->use_cascade_radiative) return XRL_CASCADE_NONRADIATIVE;
  if (!o->use_cascade_auger && o->use_cascade_radiative) return XRL_CASCADE_RADIATIVE;
  return XRL_CASCADE_FULL;
}

int xmi_init_from_provider(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr *hdf5FPtr, xmi_main_options *options) {
  if (!inputFPtr || !hdf5FPtr || !options) return 0;
  xmi_inputF *in = (xmi_inputF *)inputFPtr;
  int uniq[XMI_MAXZ + 1];
  int n = xmi_unique_Z(in, uniq);
  /* the heavy loops run on the GPU when one is usable */
  const int dev = xmi_tables_use_device();
  if (dev < 0) {
    xmi_set_error("table generator: no usable GPU");
    return 0;
  }
  xrl_atoms_preload(uniq, n);
  xmi_hdf5F *h = (xmi_hdf5F *)calloc(1, sizeof(xmi_hdf5F));
  h->cascade_type = cascade_type_of(options);
  h->use_M_lines = options->use_M_lines;
  xmi_tables_build_phi(h, !dev);
  h->n_Z = n;
  h->Zs = (xmi_hdf5_ZF *)calloc(n, sizeof(xmi_hdf5_ZF));
  for (int i = 0; i < n; i++) {
    if (!xmi_tables_build_Z(uniq[i], uniq, n, h->cascade_type, !dev, &h->Zs[i])) {
      xmi_tables_free(h);
      return 0;
    }
    h->Zs[i].Zindex = i + 1;
    h->uniqZ[uniq[i]] = i + 1;
  }
  if (dev && !xmi_tables_fill_gpu(h->Zs, n, h)) {
    xmi_tables_free(h);
    return 0;
  }
  *hdf5FPtr = h;
  return 1;
}

long xmi_tables_copy(xmi_hdf5FPtr hdf5FPtr, int Z, int which, double *out, long n) {
  const xmi_hdf5F *h = (const xmi_hdf5F *)hdf5FPtr;
  if (!h) return 0;
  const double *src = NULL;
  long len = 0;
  if (which == 4) {
    src = h->Phi_ICDF;
    len = (long)h->phi_n_theta * h->phi_n_r;
  } else {
    if (Z < 1 || Z > XMI_MAXZ || h->uniqZ[Z] == 0) return 0;
    const xmi_hdf5_ZF *z = &h->Zs[h->uniqZ[Z] - 1];
    switch (which) {
      case 0: src = z->RayleighTheta_ICDF; len = (long)z->n_e * z->n_r; break;
      case 1: src = z->ComptonTheta_ICDF; len = (long)z->n_e * z->n_r; break;
      case 2: src = z->cp_total_icdf; len = z->cp_len; break;
      case 3: src = z->cp_Qs_inv; len = (long)z->cp_nshells * z->cp_len; break;
      default: return 0;
    }
  }
  if (!src) return 0;
  if (out && n >= len) memcpy(out, src, sizeof(double) * len);
  return len;
}

int xmi_update_input_from_hdf5(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr) {
  if (!inputFPtr || !hdf5FPtr) return 0;
  xmi_inputF *in = (xmi_inputF *)inputFPtr;
  xmi_hdf5F *h = (xmi_hdf5F *)hdf5FPtr;
  for (int i = 0; i < in->n_layers; i++)
    for (int k = 0; k < in->layers[i].n_elements; k++) {
      int Z = in->layers[i].Z[k];
      if (h->uniqZ[Z] < 1) {
        xmi_set_error("xmi_update_input_from_hdf5: elements inconsistency (Z=%d not loaded)", Z);
        return 0;
      }
      in->layers[i].hdf5_index[k] = h->uniqZ[Z] - 1;
    }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* HDF5 writer */

static int write_dataset(hid_t loc, const char *name, hid_t type, int rank, const hsize_t *dims, const void *data) {
  hid_t sp = H5Screate_simple(rank, dims, dims);
  hid_t ds = H5Dcreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  herr_t st = (ds < 0) ? -1 : H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(sp);
  return st >= 0;
}

int xmi_db(char *filename, int *Zs, int nZ) {
  if (!filename || !Zs || nZ < 1) return 0;
  int *sorted = (int *)malloc(sizeof(int) * nZ);
  memcpy(sorted, Zs, sizeof(int) * nZ);
  for (int i = 1; i < nZ; i++) {
    int v = sorted[i], j = i - 1;
    while (j >= 0 && sorted[j] > v) {
      sorted[j + 1] = sorted[j];
      j--;
    }
    sorted[j + 1] = v;
  }
  hid_t gcpl = H5Pcreate(H5P_FILE_CREATE);
  H5Pset_link_creation_order(gcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  hid_t file = H5Fcreate(filename, H5F_ACC_TRUNC, gcpl, H5P_DEFAULT);
  H5Pclose(gcpl);
  if (file < 0) {
    xmi_set_error("could not create %s", filename);
    free(sorted);
    return 0;
  }
  {
    hid_t root = H5Gopen2(file, "/", H5P_DEFAULT);
    double version = 8.1;
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t at = H5Acreate2(root, "version", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(at, H5T_NATIVE_DOUBLE, &version);
    H5Aclose(at);
    hid_t st = H5Tcopy(H5T_C_S1);
    H5Tset_size(st, strlen("XMI_HDF5_DATA") + 1);
    H5Tset_strpad(st, H5T_STR_NULLTERM);
    at = H5Acreate2(root, "kind", st, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(at, st, "XMI_HDF5_DATA");
    H5Aclose(at);
    H5Tclose(st);
    H5Sclose(sp);
    H5Gclose(root);
  }
  const int dev = xmi_tables_use_device();
  int ok = dev >= 0;
  if (!ok) xmi_set_error("table generator: no usable GPU");
  for (int i = 0; i < nZ && ok; i++) {
    xmi_hdf5_ZF z;
    if (!xmi_tables_build_Z(sorted[i], NULL, 0, XRL_CASCADE_FULL, !dev, &z)) {
      ok = 0;
      break;
    }
    if (dev && !xmi_tables_fill_gpu(&z, 1, NULL)) {
      free_Z(&z);
      ok = 0;
      break;
    }
    char name[8];
    snprintf(name, sizeof(name), "%2i", sorted[i]);
    hid_t g = H5Gcreate2(file, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hid_t g2 = H5Gcreate2(g, "Theta_ICDF", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    /* on disk [r][e]; in memory [e][r] */
    double *tr = (double *)malloc(sizeof(double) * XMI_DB_NE * XMI_DB_NR);
    hsize_t d2[2] = {XMI_DB_NR, XMI_DB_NE};
    for (int e = 0; e < XMI_DB_NE; e++)
      for (int r = 0; r < XMI_DB_NR; r++) tr[(size_t)r * XMI_DB_NE + e] = z.RayleighTheta_ICDF[(size_t)e * XMI_DB_NR + r];
    ok &= write_dataset(g2, "RayleighTheta_ICDF", H5T_NATIVE_DOUBLE, 2, d2, tr);
    for (int e = 0; e < XMI_DB_NE; e++)
      for (int r = 0; r < XMI_DB_NR; r++) tr[(size_t)r * XMI_DB_NE + e] = z.ComptonTheta_ICDF[(size_t)e * XMI_DB_NR + r];
    ok &= write_dataset(g2, "ComptonTheta_ICDF", H5T_NATIVE_DOUBLE, 2, d2, tr);
    free(tr);
    hsize_t d1 = XMI_DB_NE;
    ok &= write_dataset(g2, "Energies", H5T_NATIVE_DOUBLE, 1, &d1, z.Energies);
    d1 = XMI_DB_NR;
    ok &= write_dataset(g2, "Random numbers", H5T_NATIVE_DOUBLE, 1, &d1, z.RandomNumbers);
    d1 = 9;
    ok &= write_dataset(g2, "Corrected fluorescence yields", H5T_NATIVE_DOUBLE, 1, &d1, z.FluorYieldsCorr);
    H5Gclose(g2);

    g2 = H5Gcreate2(g, "Interaction probabilities", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    d1 = z.ip_len;
    ok &= write_dataset(g2, "Energies", H5T_NATIVE_DOUBLE, 1, &d1, z.ip_energies);
    double *rc = (double *)malloc(sizeof(double) * 2 * z.ip_len);
    memcpy(rc, z.ip_rayl, sizeof(double) * z.ip_len);
    memcpy(rc + z.ip_len, z.ip_rayl_compt, sizeof(double) * z.ip_len);
    hsize_t dip[2] = {2, (hsize_t)z.ip_len};
    ok &= write_dataset(g2, "Rayleigh and Compton probabilities", H5T_NATIVE_DOUBLE, 2, dip, rc);
    free(rc);
    H5Gclose(g2);

    g2 = H5Gcreate2(g, "Compton profiles", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    d1 = z.cp_nshells;
    ok &= write_dataset(g2, "Shell indices", H5T_NATIVE_INT, 1, &d1, z.cp_shell_indices);
    d1 = z.cp_len;
    ok &= write_dataset(g2, "Qs", H5T_NATIVE_DOUBLE, 1, &d1, z.cp_Qs);
    hsize_t dcp[2] = {(hsize_t)z.cp_nshells, (hsize_t)z.cp_len};
    ok &= write_dataset(g2, "Partial profile CDF", H5T_NATIVE_DOUBLE, 2, dcp, z.cp_partial_cdf);
    ok &= write_dataset(g2, "Partial profile CDF inverted", H5T_NATIVE_DOUBLE, 1, &d1, z.cp_partial_cdf_inv);
    ok &= write_dataset(g2, "Qs inverted", H5T_NATIVE_DOUBLE, 2, dcp, z.cp_Qs_inv);
    ok &= write_dataset(g2, "Random numbers", H5T_NATIVE_DOUBLE, 1, &d1, z.cp_random_numbers);
    ok &= write_dataset(g2, "Total profile ICDF", H5T_NATIVE_DOUBLE, 1, &d1, z.cp_total_icdf);
    H5Gclose(g2);

    g2 = H5Gcreate2(g, "Precalculated XRF cross sections", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const xrl_atom *a = xrl_atom_get(sorted[i]);
    double *buf = (double *)malloc(sizeof(double) * XRL_NLINES * 9);
    for (int c = 0; c < 4; c++) {
      hid_t g3 = H5Gcreate2(g2, cascade_group_names[c], H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
      hid_t g4s[9];
      for (int s = 0; s < 9; s++) g4s[s] = H5Gcreate2(g3, shell_group_names[s], H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
      for (int j = 0; j < nZ; j++) {
        const xrl_atom *a2 = xrl_atom_get(sorted[j]);
        for (int line = 1; line <= XRL_NLINES; line++) {
          double P[9];
          xrl_shell_cross_sections(a, c + 1, a2->line_energy[line], P);
          for (int s = 0; s < 9; s++) buf[(size_t)s * XRL_NLINES + line - 1] = P[s];
        }
        char n2[8];
        snprintf(n2, sizeof(n2), "%2i", sorted[j]);
        hsize_t dl = XRL_NLINES;
        for (int s = 0; s < 9; s++) ok &= write_dataset(g4s[s], n2, H5T_NATIVE_DOUBLE, 1, &dl, buf + (size_t)s * XRL_NLINES);
      }
      for (int s = 0; s < 9; s++) H5Gclose(g4s[s]);
      H5Gclose(g3);
    }
    free(buf);
    H5Gclose(g2);
    H5Gclose(g);
    free_Z(&z);
  }
  if (ok) {
    xmi_hdf5F ph;
    memset(&ph, 0, sizeof(ph));
    xmi_tables_build_phi(&ph, !dev);
    if (dev && !xmi_tables_fill_gpu(NULL, 0, &ph)) ok = 0;
    hid_t g = H5Gcreate2(file, "Phi", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    /* on disk [r][theta] */
    double *tr = (double *)malloc(sizeof(double) * XMI_DB_NTHETA2 * XMI_DB_NR);
    for (int t = 0; t < XMI_DB_NTHETA2; t++)
      for (int r = 0; r < XMI_DB_NR; r++) tr[(size_t)r * XMI_DB_NTHETA2 + t] = ph.Phi_ICDF[(size_t)t * XMI_DB_NR + r];
    hsize_t d2[2] = {XMI_DB_NR, XMI_DB_NTHETA2};
    ok &= write_dataset(g, "Phi_ICDF", H5T_NATIVE_DOUBLE, 2, d2, tr);
    free(tr);
    hsize_t d1 = XMI_DB_NTHETA2;
    ok &= write_dataset(g, "Thetas", H5T_NATIVE_DOUBLE, 1, &d1, ph.Thetas);
    d1 = XMI_DB_NR;
    ok &= write_dataset(g, "Random numbers", H5T_NATIVE_DOUBLE, 1, &d1, ph.RandomNumbers);
    H5Gclose(g);
    free(ph.Phi_ICDF);
    free(ph.Thetas);
    free(ph.RandomNumbers);
  }
  H5Fclose(file);
  free(sorted);
  return ok;
}

/* ------------------------------------------------------------------------ */
/* HDF5 reader */

static double *read_doubles(hid_t loc, const char *name, int rank, hsize_t *dims) {
  hid_t ds = H5Dopen2(loc, name, H5P_DEFAULT);
  if (ds < 0) {
    xmi_set_error("dataset %s not found", name);
    return NULL;
  }
  hid_t sp = H5Dget_space(ds);
  int r = H5Sget_simple_extent_ndims(sp);
  if (r != rank) {
    xmi_set_error("dataset %s has rank %d, expected %d", name, r, rank);
    H5Sclose(sp);
    H5Dclose(ds);
    return NULL;
  }
  H5Sget_simple_extent_dims(sp, dims, NULL);
  size_t n = 1;
  for (int i = 0; i < rank; i++) n *= dims[i];
  double *buf = (double *)malloc(sizeof(double) * n);
  if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0) {
    free(buf);
    buf = NULL;
  }
  H5Sclose(sp);
  H5Dclose(ds);
  return buf;
}

int xmi_init_from_hdf5(char *hdf5_file, xmi_inputFPtr inputFPtr, xmi_hdf5FPtr *hdf5FPtr, xmi_main_options *options) {
  if (!hdf5_file || !inputFPtr || !hdf5FPtr || !options) return 0;
  xmi_inputF *in = (xmi_inputF *)inputFPtr;
  H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
  hid_t file = H5Fopen(hdf5_file, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) {
    xmi_set_error("could not open %s", hdf5_file);
    return 0;
  }
  int uniq[XMI_MAXZ + 1];
  int n = xmi_unique_Z(in, uniq);
  xrl_atoms_preload(uniq, n);
  xmi_hdf5F *h = (xmi_hdf5F *)calloc(1, sizeof(xmi_hdf5F));
  h->cascade_type = cascade_type_of(options);
  h->use_M_lines = options->use_M_lines;
  int ok = 1;
  hsize_t dims[2];
  {
    hid_t g = H5Gopen2(file, "Phi", H5P_DEFAULT);
    if (g < 0) ok = 0;
    double *phi = ok ? read_doubles(g, "Phi_ICDF", 2, dims) : NULL;
    if (!phi) ok = 0;
    if (ok) {
      h->phi_n_r = (int)dims[0];
      h->phi_n_theta = (int)dims[1];
      h->Phi_ICDF = (double *)malloc(sizeof(double) * dims[0] * dims[1]);
      for (int t = 0; t < h->phi_n_theta; t++)
        for (int r = 0; r < h->phi_n_r; r++)
          h->Phi_ICDF[(size_t)t * h->phi_n_r + r] = phi[(size_t)r * h->phi_n_theta + t];
      h->Thetas = read_doubles(g, "Thetas", 1, dims);
      h->RandomNumbers = read_doubles(g, "Random numbers", 1, dims);
      if (!h->Thetas || !h->RandomNumbers) ok = 0;
    }
    free(phi);
    if (g >= 0) H5Gclose(g);
  }
  h->n_Z = n;
  h->Zs = (xmi_hdf5_ZF *)calloc(n, sizeof(xmi_hdf5_ZF));
  for (int i = 0; i < n && ok; i++) {
    xmi_hdf5_ZF *z = &h->Zs[i];
    z->Z = uniq[i];
    z->Zindex = i + 1;
    h->uniqZ[uniq[i]] = i + 1;
    char name[8];
    snprintf(name, sizeof(name), "%2i", uniq[i]);
    hid_t g = H5Gopen2(file, name, H5P_DEFAULT);
    if (g < 0) {
      xmi_set_error("element %d not present in %s", uniq[i], hdf5_file);
      ok = 0;
      break;
    }
    hid_t g2 = H5Gopen2(g, "Theta_ICDF", H5P_DEFAULT);
    double *ray = read_doubles(g2, "RayleighTheta_ICDF", 2, dims);
    double *com = read_doubles(g2, "ComptonTheta_ICDF", 2, dims);
    z->n_r = (int)dims[0];
    z->n_e = (int)dims[1];
    z->Energies = read_doubles(g2, "Energies", 1, dims);
    z->RandomNumbers = read_doubles(g2, "Random numbers", 1, dims);
    double *fy = read_doubles(g2, "Corrected fluorescence yields", 1, dims);
    if (!ray || !com || !z->Energies || !z->RandomNumbers || !fy) ok = 0;
    if (ok) {
      z->RayleighTheta_ICDF = (double *)malloc(sizeof(double) * z->n_e * z->n_r);
      z->ComptonTheta_ICDF = (double *)malloc(sizeof(double) * z->n_e * z->n_r);
      for (int e = 0; e < z->n_e; e++)
        for (int r = 0; r < z->n_r; r++) {
          z->RayleighTheta_ICDF[(size_t)e * z->n_r + r] = ray[(size_t)r * z->n_e + e];
          z->ComptonTheta_ICDF[(size_t)e * z->n_r + r] = com[(size_t)r * z->n_e + e];
        }
      memcpy(z->FluorYieldsCorr, fy, sizeof(double) * 9);
    }
    free(ray);
    free(com);
    free(fy);
    H5Gclose(g2);

    g2 = H5Gopen2(g, "Interaction probabilities", H5P_DEFAULT);
    z->ip_energies = read_doubles(g2, "Energies", 1, dims);
    z->ip_len = (int)dims[0];
    double *rc = read_doubles(g2, "Rayleigh and Compton probabilities", 2, dims);
    if (!z->ip_energies || !rc || (int)dims[1] != z->ip_len) ok = 0;
    if (ok) {
      z->ip_rayl = (double *)malloc(sizeof(double) * z->ip_len);
      z->ip_rayl_compt = (double *)malloc(sizeof(double) * z->ip_len);
      memcpy(z->ip_rayl, rc, sizeof(double) * z->ip_len);
      memcpy(z->ip_rayl_compt, rc + z->ip_len, sizeof(double) * z->ip_len);
    }
    free(rc);
    H5Gclose(g2);

    g2 = H5Gopen2(g, "Compton profiles", H5P_DEFAULT);
    {
      hid_t ds = H5Dopen2(g2, "Shell indices", H5P_DEFAULT);
      hid_t sp = H5Dget_space(ds);
      H5Sget_simple_extent_dims(sp, dims, NULL);
      z->cp_nshells = (int)dims[0];
      z->cp_shell_indices = (int *)malloc(sizeof(int) * z->cp_nshells);
      if (H5Dread(ds, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, z->cp_shell_indices) < 0) ok = 0;
      H5Sclose(sp);
      H5Dclose(ds);
    }
    z->cp_Qs = read_doubles(g2, "Qs", 1, dims);
    z->cp_len = (int)dims[0];
    z->cp_partial_cdf = read_doubles(g2, "Partial profile CDF", 2, dims);
    z->cp_partial_cdf_inv = read_doubles(g2, "Partial profile CDF inverted", 1, dims);
    z->cp_Qs_inv = read_doubles(g2, "Qs inverted", 2, dims);
    z->cp_random_numbers = read_doubles(g2, "Random numbers", 1, dims);
    z->cp_total_icdf = read_doubles(g2, "Total profile ICDF", 1, dims);
    if (!z->cp_Qs || !z->cp_partial_cdf || !z->cp_partial_cdf_inv || !z->cp_Qs_inv || !z->cp_random_numbers ||
        !z->cp_total_icdf)
      ok = 0;
    H5Gclose(g2);

    if (options->use_variance_reduction && ok) {
      z->precalc_xrf_cs = (double *)calloc((size_t)9 * n * XMI_NLINES_VR, sizeof(double));
      g2 = H5Gopen2(g, "Precalculated XRF cross sections", H5P_DEFAULT);
      hid_t g3 = H5Gopen2(g2, cascade_group_names[h->cascade_type - 1], H5P_DEFAULT);
      int last_shell = options->use_M_lines ? 8 : 3;
      for (int s = 0; s <= last_shell && ok; s++) {
        hid_t g4 = H5Gopen2(g3, shell_group_names[s], H5P_DEFAULT);
        for (int j = 0; j < n && ok; j++) {
          char n2[8];
          snprintf(n2, sizeof(n2), "%2i", uniq[j]);
          double *v = read_doubles(g4, n2, 1, dims);
          if (!v) {
            ok = 0;
            break;
          }
          memcpy(z->precalc_xrf_cs + ((size_t)s * n + j) * XMI_NLINES_VR, v, sizeof(double) * XMI_NLINES_VR);
          free(v);
        }
        H5Gclose(g4);
      }
      H5Gclose(g3);
      H5Gclose(g2);
    }
    H5Gclose(g);
  }
  H5Fclose(file);
  if (!ok) {
    xmi_tables_free(h);
    return 0;
  }
  *hdf5FPtr = h;
  return 1;
}

int xmi_provider_active_lines(int Z, int line_last) {
  const xrl_atom *a = xrl_atom_get(Z);
  if (!a) return 0;
  int n = 0;
  for (int line = 1; line <= line_last && line <= XMI_NLINES_VR; line++) {
    if (a->line_energy[line] < XMI_ENERGY_THRESHOLD) continue;
    int s = xrl_line_shell(line);
    if (a->fluor_yield[s] * a->rad_rate[line] == 0.0) continue;
    n++;
  }
  return n;
}

/* ------------------------------------------------------------------------ */
/* physics provider queries */

const char *xmi_line_name(int line) { return xrl_line_name(line); }

int xmi_provider_atom(int Z, void *out_) {
  xrl_atom *out = (xrl_atom *)out_;
  const xrl_atom *a = xrl_atom_get(Z);
  if (!a || !out) return 0;
  *out = *a;
  return 1;
}

double xmi_provider_cs(int kind, int Z, double E) {
  const xrl_atom *a = xrl_atom_get(Z);
  if (!a) return -1.0;
  switch (kind) {
    case 0: return xrl_cs_total(a, E);
    case 1: return xrl_cs_photo_total(a, E);
    case 2: return xrl_cs_rayl(a, E);
    case 3: return xrl_cs_compt(a, E);
    default: return -1.0;
  }
}

double xmi_provider_dcs(int kind, int Z, double E, double theta) {
  const xrl_atom *a = xrl_atom_get(Z);
  if (!a) return -1.0;
  return kind == 0 ? xrl_dcs_rayl(a, E, theta) : xrl_dcs_compt(a, E, theta);
}

/* which: 0 fluorescence yield of shell `index` (0 = K .. 8 = M5), 1 radiative
 * rate of line `index`, 2 energy of line `index` (keV) */
double xmi_provider_line_data(int Z, int which, int index) {
  const xrl_atom *a = xrl_atom_get(Z);
  if (!a) return 0.0;
  switch (which) {
    case 0: return index >= 0 && index < 9 ? a->fluor_yield[index] : 0.0;
    case 1: return index >= 1 && index <= XRL_NLINES ? a->rad_rate[index] : 0.0;
    case 2: return index >= 1 && index <= XRL_NLINES ? a->line_energy[index] : 0.0;
    default: return 0.0;
  }
}

double xmi_provider_auger(int Z, int trans, int *shells3) {
  const xrl_atom *a = xrl_atom_get(Z);
  if (!a || trans < 0 || trans >= XRL_AUGER_N) return -1.0;
  if (shells3) {
    shells3[0] = trans < 64 ? XRL_K_SHELL : XRL_L1_SHELL + (trans - 64) / 30;
    xrl_auger_shells(trans, &shells3[1], &shells3[2]);
  }
  return xrl_auger_rate(a, trans);
}

int xmi_provider_shell_cross_sections(int Z, int cascade_type, double E, double *P9) {
  const xrl_atom *a = xrl_atom_get(Z);
  if (!a) return 0;
  xrl_shell_cross_sections(a, cascade_type, E, P9);
  return 1;
}
