/*
 * xmi_cache.c -- the HDF5 caches of solid-angle grids and escape ratios.
 *
 * Reference interfaces replaced (XMI-MSIM 8.1):
 *   xmi_create_empty_solid_angle_hdf5_file  src/xmi_solid_angle.c:193-233
 *   xmi_update_solid_angle_hdf5_file        src/xmi_solid_angle.c:236-312
 *   xmi_check_solid_angle_match             src/xmi_solid_angle.c:422-673
 *   xmi_find_solid_angle_match              src/xmi_solid_angle.c:675-785
 *   xmi_get_solid_angle_file                src/xmi_solid_angle.c:796-821
 *   xmi_create_empty_escape_ratios_hdf5_file src/xmi_detector.c:40-89
 *   xmi_check_escape_ratios_match           src/xmi_detector.c:143-172
 *   xmi_update_escape_ratios_hdf5_file      src/xmi_detector.c:174-283
 *   xmi_find_escape_ratios_match            src/xmi_detector.c:437-555
 *   xmi_get_escape_ratios_file              src/xmi_detector.c:582-618
 *
 * File layout (kept as the reference writes it, so that either program can
 * read the other's cache): root attributes "version" (double) and "kind"
 * (string XMI_HDF5_SOLID_ANGLES / XMI_HDF5_ESCAPE_RATIOS); one group per
 * stored result, named "<user> <time>", created with link creation order
 * tracked; each group holds the result's arrays and "xmi_input_string", the
 * XMSI text of the input it was computed for.  A lookup walks the groups
 * newest first and returns the first whose input matches.
 */
#include <errno.h>
#include <stdio.h>
#include <math.h>
#include <pwd.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <time.h>
#include <unistd.h>

#include <hdf5.h>

#include "xmi_internal.h"

#define CACHE_VERSION 8.1
#define CACHE_MIN_VERSION 7.2 /* include/xmi_solid_angle.h:52, include/xmi_detector.h:75 */
#define COMPARE_THRESHOLD 1e-10 /* include/xmi_data_structs.h:517 */
#define KIND_SA "XMI_HDF5_SOLID_ANGLES"
#define KIND_ESC "XMI_HDF5_ESCAPE_RATIOS"

/* ---- file-level helpers -------------------------------------------------- */

static int create_empty(const char *file, const char *kind) {
  hid_t fcpl = H5Pcreate(H5P_FILE_CREATE);
  H5Pset_link_creation_order(fcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  hid_t f = H5Fcreate(file, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT);
  H5Pclose(fcpl);
  if (f < 0) {
    xmi_set_error("could not create cache file %s", file);
    return 0;
  }
  const double version = CACHE_VERSION;
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(f, "version", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
  H5Awrite(a, H5T_NATIVE_DOUBLE, &version);
  H5Aclose(a);
  hid_t st = H5Tcopy(H5T_C_S1);
  H5Tset_size(st, strlen(kind) + 1);
  H5Tset_strpad(st, H5T_STR_NULLTERM);
  a = H5Acreate2(f, "kind", st, sp, H5P_DEFAULT, H5P_DEFAULT);
  H5Awrite(a, st, kind);
  H5Aclose(a);
  H5Tclose(st);
  H5Sclose(sp);
  H5Fclose(f);
  return 1;
}

/* "<user> <local time with microseconds>": unique per stored result */
static void group_name(char *buf, size_t n) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  struct tm lt;
  localtime_r(&tv.tv_sec, &lt);
  char t[64];
  strftime(t, sizeof t, "%Y-%m-%dT%H:%M:%S", &lt);
  const struct passwd *pw = getpwuid(getuid());
  snprintf(buf, n, "%s %s.%06ldZ", pw && pw->pw_name ? pw->pw_name : "unknown", t, (long)tv.tv_usec);
}

static hid_t new_group(hid_t f) {
  char name[256];
  hid_t gcpl = H5Pcreate(H5P_GROUP_CREATE);
  H5Pset_link_creation_order(gcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
  hid_t g = -1;
  for (int attempt = 0; attempt < 4 && g < 0; attempt++) {
    group_name(name, sizeof name);
    H5E_BEGIN_TRY { g = H5Gcreate2(f, name, H5P_DEFAULT, gcpl, H5P_DEFAULT); }
    H5E_END_TRY;
    if (g < 0) usleep(2);
  }
  H5Pclose(gcpl);
  return g;
}

static int write_ds(hid_t g, const char *name, hid_t type, int rank, const hsize_t *dims, const void *data) {
  hid_t sp = H5Screate_simple(rank, dims, dims);
  hid_t d = H5Dcreate2(g, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  const herr_t rc = d < 0 ? -1 : H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  if (d >= 0) H5Dclose(d);
  H5Sclose(sp);
  return rc >= 0;
}

/* a dataset of the expected rank into a malloc'ed buffer; dims out */
static void *read_ds(hid_t g, const char *name, hid_t type, size_t esize, int rank, hsize_t *dims) {
  hid_t d = H5Dopen2(g, name, H5P_DEFAULT);
  if (d < 0) return NULL;
  hid_t sp = H5Dget_space(d);
  void *buf = NULL;
  if (H5Sget_simple_extent_ndims(sp) == rank) {
    H5Sget_simple_extent_dims(sp, dims, NULL);
    size_t n = 1;
    for (int k = 0; k < rank; k++) n *= (size_t)dims[k];
    buf = malloc(n * esize + 1);
    if (buf && H5Dread(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0) {
      free(buf);
      buf = NULL;
    }
  }
  H5Sclose(sp);
  H5Dclose(d);
  return buf;
}

/* default cache location: $XDG_DATA_HOME/XMI-MSIM/<name>, else
 * ~/.local/share/XMI-MSIM/<name> (GLib's user data directory); with
 * create_file the file and its directory are made when missing */
static int cache_file(char **filePtr, int create_file, const char *basename, const char *kind) {
  char *file = *filePtr;
  if (!file) {
    const char *xdg = getenv("XDG_DATA_HOME"), *home = getenv("HOME");
    char dir[4096];
    if (xdg && xdg[0]) snprintf(dir, sizeof dir, "%s/XMI-MSIM", xdg);
    else snprintf(dir, sizeof dir, "%s/.local/share/XMI-MSIM", home && home[0] ? home : "/tmp");
    const size_t n = strlen(dir) + strlen(basename) + 2;
    file = (char *)malloc(n);
    snprintf(file, n, "%s/%s", dir, basename);
    *filePtr = file;
  }
  struct stat sb;
  if (stat(file, &sb) == 0 || !create_file) return 1;
  /* make the parent directories */
  char *dir = strdup(file);
  for (char *p = dir + 1; *p; p++) {
    if (*p != '/') continue;
    *p = '\0';
    if (mkdir(dir, 0700) != 0 && errno != EEXIST) {
      xmi_set_error("could not create directory %s", dir);
      free(dir);
      return 0;
    }
    *p = '/';
  }
  free(dir);
  return create_empty(file, kind);
}

/* opens a cache file for reading after its kind and version checks:
 * returns the file id; 0 with *rebuilt set when the file lacked the
 * attributes or was too old and has been recreated empty; -1 on error */
static hid_t open_checked(const char *file, const char *kind, const char *basename, int *rebuilt) {
  *rebuilt = 0;
  hid_t f;
  H5E_BEGIN_TRY { f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT); }
  H5E_END_TRY;
  if (f < 0) {
    xmi_set_error("cannot open %s for reading", file);
    return -1;
  }
  int stale = 0;
  if (H5Aexists(f, "kind") <= 0 || H5Aexists(f, "version") <= 0) {
    stale = 1;
  } else {
    hid_t a = H5Aopen(f, "kind", H5P_DEFAULT);
    hid_t t = H5Aget_type(a), tm = H5Tget_native_type(t, H5T_DIR_ASCEND);
    char k[512] = "";
    if (H5Tget_size(tm) < sizeof k) H5Aread(a, tm, k);
    H5Tclose(tm);
    H5Tclose(t);
    H5Aclose(a);
    if (strcmp(k, kind) != 0) {
      H5Fclose(f);
      xmi_set_error("%s is a cache of kind %s, expected %s", file, k, kind);
      return -1;
    }
    double version = 0.0;
    a = H5Aopen(f, "version", H5P_DEFAULT);
    H5Aread(a, H5T_NATIVE_DOUBLE, &version);
    H5Aclose(a);
    stale = version < CACHE_MIN_VERSION;
  }
  if (!stale) return f;
  H5Fclose(f);
  if (unlink(file) != 0) {
    xmi_set_error("could not delete %s", file);
    return -1;
  }
  char *p = (char *)file;
  if (!cache_file(&p, 1, basename, kind)) return -1;
  *rebuilt = 1;
  return 0;
}

/* the XMSI text stored with a result, and its input */
static xmi_input *stored_input(hid_t g, char **text_out) {
  hsize_t n = 0;
  char *text = (char *)read_ds(g, "xmi_input_string", H5T_NATIVE_CHAR, 1, 1, &n);
  if (!text) return NULL;
  text[n] = '\0';
  xmi_input *in = xmi_input_read_from_xml_string(text, NULL);
  if (!in) {
    free(text);
    return NULL;
  }
  *text_out = text;
  return in;
}

/* ---- solid angles --------------------------------------------------------- */

int xmi_create_empty_solid_angle_hdf5_file(char *hdf5_file) { return create_empty(hdf5_file, KIND_SA); }

int xmi_get_solid_angle_file(char **filePtr, int create_file) {
  return cache_file(filePtr, create_file, "xmimsim-solid-angles.h5", KIND_SA);
}

int xmi_update_solid_angle_hdf5_file(char *hdf5_file, xmi_solid_angle *sa) {
  if (!hdf5_file || !sa || !sa->solid_angles || !sa->xmi_input_string) {
    xmi_set_error("xmi_update_solid_angle_hdf5_file: missing grid or input string");
    return 0;
  }
  hid_t f = H5Fopen(hdf5_file, H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0) {
    xmi_set_error("cannot open %s for read/write", hdf5_file);
    return 0;
  }
  hid_t g = new_group(f);
  int ok = g >= 0;
  if (ok) {
    hsize_t dims[2] = {(hsize_t)sa->grid_dims_theta_n, (hsize_t)sa->grid_dims_r_n};
    hsize_t ns = strlen(sa->xmi_input_string) + 1;
    ok = write_ds(g, "solid_angles", H5T_NATIVE_DOUBLE, 2, dims, sa->solid_angles) &&
         write_ds(g, "grid_dims_r_vals", H5T_NATIVE_DOUBLE, 1, &dims[1], sa->grid_dims_r_vals) &&
         write_ds(g, "grid_dims_theta_vals", H5T_NATIVE_DOUBLE, 1, &dims[0], sa->grid_dims_theta_vals) &&
         write_ds(g, "xmi_input_string", H5T_NATIVE_CHAR, 1, &ns, sa->xmi_input_string);
    H5Gclose(g);
  }
  H5Fflush(f, H5F_SCOPE_GLOBAL);
  H5Fclose(f);
  if (!ok) xmi_set_error("could not write a solid-angle grid to %s", hdf5_file);
  return ok;
}

/* mass attenuation of each layer at E, sum_i w_i sigma_tot(Z_i, E) */
static double layer_mu(const xmi_layer *l, double E) {
  double mu = 0.0;
  for (int j = 0; j < l->n_elements; j++) mu += xmi_provider_cs(0, l->Z[j], E) * l->weight[j];
  return mu;
}

/* depth along z below which a fraction R of the beam of energy E has
 * interacted (src/xmi_solid_angle.c:487-556 for one of S1, S2); z is
 * measured from the top of the reference layer */
static double depth_quantile(const xmi_input *in, const double *tz, double E, double R) {
  const int nl = in->composition->n_layers;
  if (nl < 1) return 0.0;
  double *mu = (double *)malloc(sizeof(double) * nl), sum = 0.0;
  for (int i = 0; i < nl; i++) {
    mu[i] = layer_mu(&in->composition->layers[i], E);
    sum += mu[i] * in->composition->layers[i].density * tz[i];
  }
  const double Pabs = -expm1(-sum), myln = -log1p(-R * Pabs);
  int m = 0;
  sum = 0.0;
  for (m = 0; m < nl; m++) {
    sum += mu[m] * in->composition->layers[m].density * tz[m];
    if (sum > myln) break;
  }
  if (m == nl) m = nl - 1;
  const double mrm = mu[m] * in->composition->layers[m].density;
  double s = 0.0;
  for (int i = 0; i < m; i++) s += (1.0 - mu[i] * in->composition->layers[i].density / mrm) * tz[i];
  /* z of the top of the first layer: minus the thickness above the
   * reference layer */
  double z0 = 0.0;
  for (int i = 0; i < in->composition->reference_layer - 1; i++) z0 -= tz[i];
  free(mu);
  return s + myln / mrm + z0;
}

static void energy_range(const xmi_input *in, double *lo, double *hi) {
  const xmi_excitation *x = in->excitation;
  if (x->n_continuous > 1 && x->n_discrete > 0) {
    *lo = fmin(x->continuous[0].energy, x->discrete[0].energy);
    *hi = fmax(x->continuous[x->n_continuous - 1].energy, x->discrete[x->n_discrete - 1].energy);
  } else if (x->n_continuous > 1) {
    *lo = x->continuous[0].energy;
    *hi = x->continuous[x->n_continuous - 1].energy;
  } else {
    *lo = x->discrete[0].energy;
    *hi = x->discrete[x->n_discrete - 1].energy;
  }
}

static void unit(const double *v, double *u) {
  const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int k = 0; k < 3; k++) u[k] = n > 0.0 ? v[k] / n : v[k];
}

/* A grid computed for input A serves input B when B's interaction depths
 * (from 1e-5 to 0.99999 of the absorbed beam, at the lowest and highest
 * source energies) lie inside A's, and the detector geometry is the same */
int xmi_check_solid_angle_match(xmi_input *A, xmi_input *B) {
  double S[2][2];
  const xmi_input *in[2] = {A, B};
  for (int k = 0; k < 2; k++) {
    double ns[3], lo, hi;
    unit(in[k]->geometry->n_sample_orientation, ns);
    const int nl = in[k]->composition->n_layers;
    double *tz = (double *)malloc(sizeof(double) * nl);
    for (int i = 0; i < nl; i++) tz[i] = fabs(in[k]->composition->layers[i].thickness / ns[2]);
    energy_range(in[k], &lo, &hi);
    S[k][0] = depth_quantile(in[k], tz, lo, 0.00001);
    S[k][1] = depth_quantile(in[k], tz, hi, 0.99999);
    free(tz);
  }
  if (S[0][1] - S[1][1] < -0.0001 || S[0][0] - S[1][0] > 0.0001) return 0;
  const xmi_geometry *a = A->geometry, *b = B->geometry;
  if (fabs(a->p_detector_window[0] - b->p_detector_window[0]) > COMPARE_THRESHOLD) return 0;
  if (fabs(a->p_detector_window[1] - b->p_detector_window[1]) > COMPARE_THRESHOLD) return 0;
  if (fabs((a->p_detector_window[2] - a->d_sample_source) - (b->p_detector_window[2] - b->d_sample_source)) >
      COMPARE_THRESHOLD)
    return 0;
  double da[3], db[3];
  unit(a->n_detector_orientation, da);
  unit(b->n_detector_orientation, db);
  for (int k = 0; k < 3; k++)
    if (fabs(da[k] - db[k]) > COMPARE_THRESHOLD) return 0;
  if (fabs(a->area_detector - b->area_detector) / a->area_detector > COMPARE_THRESHOLD) return 0;
  if (fabs(a->collimator_height - b->collimator_height) > COMPARE_THRESHOLD) return 0;
  if (fabs(a->collimator_diameter - b->collimator_diameter) > COMPARE_THRESHOLD) return 0;
  return 1;
}

typedef struct {
  xmi_input *input;
  void **result;
  int verbose;
  int kind; /* 0 solid angles, 1 escape ratios */
} lookup_t;

static herr_t read_sa_group(hid_t g, xmi_solid_angle **rv) {
  hsize_t d2[2], dr, dt;
  xmi_solid_angle *sa = (xmi_solid_angle *)calloc(1, sizeof(xmi_solid_angle));
  sa->solid_angles = (double *)read_ds(g, "solid_angles", H5T_NATIVE_DOUBLE, 8, 2, d2);
  sa->grid_dims_r_vals = (double *)read_ds(g, "grid_dims_r_vals", H5T_NATIVE_DOUBLE, 8, 1, &dr);
  sa->grid_dims_theta_vals = (double *)read_ds(g, "grid_dims_theta_vals", H5T_NATIVE_DOUBLE, 8, 1, &dt);
  if (!sa->solid_angles || !sa->grid_dims_r_vals || !sa->grid_dims_theta_vals || dr != d2[1] || dt != d2[0]) {
    xmi_free_solid_angle(sa);
    xmi_set_error("malformed solid-angle group");
    return -1;
  }
  sa->grid_dims_theta_n = (long)d2[0];
  sa->grid_dims_r_n = (long)d2[1];
  *rv = sa;
  return 1;
}

static herr_t read_esc_group(hid_t g, xmi_escape_ratios **rv) {
  hsize_t d1, d3[3], d2[2], dfi, dci, dco;
  xmi_escape_ratios *e = (xmi_escape_ratios *)calloc(1, sizeof(xmi_escape_ratios));
  e->Z = (int *)read_ds(g, "elements", H5T_NATIVE_INT, sizeof(int), 1, &d1);
  e->fluo_escape_ratios = (double *)read_ds(g, "fluo escape ratios", H5T_NATIVE_DOUBLE, 8, 3, d3);
  e->fluo_escape_input_energies = (double *)read_ds(g, "fluo escape input energies", H5T_NATIVE_DOUBLE, 8, 1, &dfi);
  e->compton_escape_ratios = (double *)read_ds(g, "compton escape ratios", H5T_NATIVE_DOUBLE, 8, 2, d2);
  e->compton_escape_input_energies =
      (double *)read_ds(g, "compton escape input energies", H5T_NATIVE_DOUBLE, 8, 1, &dci);
  e->compton_escape_output_energies =
      (double *)read_ds(g, "compton escape output energies", H5T_NATIVE_DOUBLE, 8, 1, &dco);
  if (!e->Z || !e->fluo_escape_ratios || !e->fluo_escape_input_energies || !e->compton_escape_ratios ||
      !e->compton_escape_input_energies || !e->compton_escape_output_energies || d3[2] != d1 ||
      d3[1] != XMI_ESC_LINES || dfi != d3[0] || dci != d2[1] || dco != d2[0]) {
    xmi_free_escape_ratios(e);
    xmi_set_error("malformed escape-ratio group");
    return -1;
  }
  e->n_elements = (int)d1;
  e->n_fluo_input_energies = (int)d3[0];
  e->n_compton_input_energies = (int)d2[1];
  e->n_compton_output_energies = (int)d2[0];
  *rv = e;
  return 1;
}

int xmi_check_escape_ratios_match(xmi_input *A, xmi_input *B);

static herr_t visit(hid_t file, const char *name, const H5L_info_t *info, void *op) {
  (void)info;
  lookup_t *lk = (lookup_t *)op;
  hid_t g = H5Gopen2(file, name, H5P_DEFAULT);
  if (g < 0) return -1;
  char *text = NULL;
  xmi_input *stored = stored_input(g, &text);
  if (!stored) {
    H5Gclose(g);
    xmi_set_error("group %s: unreadable xmi_input_string", name);
    return -1;
  }
  const int match = lk->kind == 0 ? xmi_check_solid_angle_match(stored, lk->input)
                                  : xmi_check_escape_ratios_match(stored, lk->input);
  xmi_input_free(stored);
  if (lk->verbose) fprintf(stdout, "%s in group %s\n", match ? "Match" : "No match", name);
  herr_t rc = 0;
  if (match) {
    rc = lk->kind == 0 ? read_sa_group(g, (xmi_solid_angle **)lk->result)
                       : read_esc_group(g, (xmi_escape_ratios **)lk->result);
    if (rc == 1) {
      if (lk->kind == 0) (*(xmi_solid_angle **)lk->result)->xmi_input_string = text;
      else (*(xmi_escape_ratios **)lk->result)->xmi_input_string = text;
      text = NULL;
    }
  }
  free(text);
  H5Gclose(g);
  return rc;
}

static int find_match(char *file, xmi_input *A, void **rv, xmi_main_options *options, int kind) {
  *rv = NULL;
  if (!file || !A) {
    xmi_set_error("cache lookup: missing file or input");
    return 0;
  }
  int rebuilt;
  const char *basename = kind == 0 ? "xmimsim-solid-angles.h5" : "xmimsim-escape-ratios.h5";
  hid_t f = open_checked(file, kind == 0 ? KIND_SA : KIND_ESC, basename, &rebuilt);
  if (f < 0) return 0;
  if (rebuilt) return 1;
  lookup_t lk = {A, rv, options ? options->extra_verbose : 0, kind};
  const herr_t rc = H5Literate(f, H5_INDEX_CRT_ORDER, H5_ITER_DEC, NULL, visit, &lk);
  H5Fclose(f);
  if (rc < 0) {
    if (*rv) {
      if (kind == 0) xmi_free_solid_angle((xmi_solid_angle *)*rv);
      else xmi_free_escape_ratios((xmi_escape_ratios *)*rv);
      *rv = NULL;
    }
    return 0;
  }
  return 1;
}

int xmi_find_solid_angle_match(char *hdf5_file, xmi_input *A, xmi_solid_angle **rv, xmi_main_options *options) {
  return find_match(hdf5_file, A, (void **)rv, options, 0);
}

/* ---- escape ratios -------------------------------------------------------- */

int xmi_create_empty_escape_ratios_hdf5_file(char *hdf5_file) { return create_empty(hdf5_file, KIND_ESC); }

int xmi_get_escape_ratios_file(char **filePtr, int create_file) {
  return cache_file(filePtr, create_file, "xmimsim-escape-ratios.h5", KIND_ESC);
}

/* the crystal's layers must agree: count, thickness and density to 1e-10
 * relative, the same elements with weights to 1e-10 */
int xmi_check_escape_ratios_match(xmi_input *A, xmi_input *B) {
  const xmi_detector *a = A->detector, *b = B->detector;
  if (a->n_crystal_layers != b->n_crystal_layers) return 0;
  for (int i = 0; i < a->n_crystal_layers; i++) {
    const xmi_layer *la = &a->crystal_layers[i], *lb = &b->crystal_layers[i];
    if (fabs(la->thickness - lb->thickness) / la->thickness > COMPARE_THRESHOLD) return 0;
    if (fabs(la->density - lb->density) / la->density > COMPARE_THRESHOLD) return 0;
    if (la->n_elements != lb->n_elements) return 0;
    for (int j = 0; j < la->n_elements; j++) {
      if (la->Z[j] != lb->Z[j]) return 0;
      if (fabs(la->weight[j] - lb->weight[j]) > COMPARE_THRESHOLD) return 0;
    }
  }
  return 1;
}

int xmi_update_escape_ratios_hdf5_file(char *hdf5_file, xmi_escape_ratios *e) {
  if (!hdf5_file || !e || !e->Z || !e->fluo_escape_ratios || !e->compton_escape_ratios || !e->xmi_input_string) {
    xmi_set_error("xmi_update_escape_ratios_hdf5_file: missing ratios or input string");
    return 0;
  }
  hid_t f = H5Fopen(hdf5_file, H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0) {
    xmi_set_error("cannot open %s for read/write", hdf5_file);
    return 0;
  }
  hid_t g = new_group(f);
  int ok = g >= 0;
  if (ok) {
    hsize_t d1 = (hsize_t)e->n_elements;
    hsize_t d3[3] = {(hsize_t)e->n_fluo_input_energies, (hsize_t)XMI_ESC_LINES, (hsize_t)e->n_elements};
    hsize_t d2[2] = {(hsize_t)e->n_compton_output_energies, (hsize_t)e->n_compton_input_energies};
    hsize_t ns = strlen(e->xmi_input_string) + 1;
    ok = write_ds(g, "elements", H5T_NATIVE_INT, 1, &d1, e->Z) &&
         write_ds(g, "fluo escape ratios", H5T_NATIVE_DOUBLE, 3, d3, e->fluo_escape_ratios) &&
         write_ds(g, "fluo escape input energies", H5T_NATIVE_DOUBLE, 1, &d3[0], e->fluo_escape_input_energies) &&
         write_ds(g, "compton escape ratios", H5T_NATIVE_DOUBLE, 2, d2, e->compton_escape_ratios) &&
         write_ds(g, "compton escape input energies", H5T_NATIVE_DOUBLE, 1, &d2[1],
                  e->compton_escape_input_energies) &&
         write_ds(g, "compton escape output energies", H5T_NATIVE_DOUBLE, 1, &d2[0],
                  e->compton_escape_output_energies) &&
         write_ds(g, "xmi_input_string", H5T_NATIVE_CHAR, 1, &ns, e->xmi_input_string);
    H5Gclose(g);
  }
  H5Fflush(f, H5F_SCOPE_GLOBAL);
  H5Fclose(f);
  if (!ok) xmi_set_error("could not write escape ratios to %s", hdf5_file);
  return ok;
}

int xmi_find_escape_ratios_match(char *hdf5_file, xmi_input *A, xmi_escape_ratios **rv, xmi_main_options *options) {
  return find_match(hdf5_file, A, (void **)rv, options, 1);
}
