/*
 * xmi_input.c -- input handling of the C ABI: options, the working copy of
 * xmi_input behind xmi_inputFPtr and its geometry set-up.
 *
 *   xmi_main_options_new    src/xmi_data_structs.c:2531-2566
 *   xmi_input_C2F           src/xmi_aux_f.F90:782-1025 (weights normalised, :857-858)
 *   xmi_init_input          src/xmi_main.F90:1741-1918
 *   xmi_free_input_F        src/xmi_aux_f.F90:1027
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xmi_internal.h"

static __thread char g_error[512];

void xmi_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

const char *xmi_msim_last_error(void) { return g_error; }

void xmi_free(void *ptr) { free(ptr); }

xmi_main_options *xmi_main_options_new(void) {
  xmi_main_options *o = (xmi_main_options *)calloc(1, sizeof(xmi_main_options));
  if (!o) return NULL;
  o->use_M_lines = 1;
  o->use_cascade_auger = 1;
  o->use_cascade_radiative = 1;
  o->use_variance_reduction = 1;
  o->use_sum_peaks = 0;
  o->use_escape_peaks = 1;
  o->escape_ratios_mode = 0;
  o->verbose = 0;
  o->use_poisson = 0;
  o->use_gpu = 1;
  o->omp_num_threads = 1;
  o->extra_verbose = 0;
  o->custom_detector_response = NULL;
  o->use_advanced_compton = 0;
  o->use_default_seeds = 0;
  return o;
}

void xmi_normalize3(double v[3]) {
  double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= n;
  v[1] /= n;
  v[2] /= n;
}

static int copy_layer(const xmi_layer *src, xmi_layerF *dst, int normalise) {
  memset(dst, 0, sizeof(*dst));
  if (src->n_elements < 1 || !src->Z || !src->weight) {
    xmi_set_error("layer without elements");
    return 0;
  }
  dst->n_elements = src->n_elements;
  dst->Z = (int *)malloc(sizeof(int) * src->n_elements);
  dst->weight = (double *)malloc(sizeof(double) * src->n_elements);
  dst->hdf5_index = (int *)calloc(src->n_elements, sizeof(int));
  if (!dst->Z || !dst->weight || !dst->hdf5_index) return 0;
  double sum = 0.0;
  for (int i = 0; i < src->n_elements; i++) {
    if (src->Z[i] < 1 || src->Z[i] > XMI_MAXZ) {
      xmi_set_error("atomic number %d out of range 1..%d", src->Z[i], XMI_MAXZ);
      return 0;
    }
    dst->Z[i] = src->Z[i];
    dst->weight[i] = src->weight[i];
    sum += src->weight[i];
  }
  if (normalise) {
    if (!(sum > 0.0)) {
      xmi_set_error("layer weight fractions sum to %g", sum);
      return 0;
    }
    for (int i = 0; i < src->n_elements; i++) dst->weight[i] /= sum;
  }
  dst->density = src->density;
  dst->thickness = src->thickness;
  return 1;
}

static void free_layer(xmi_layerF *l) {
  free(l->Z);
  free(l->weight);
  free(l->hdf5_index);
}

void xmi_free_input_F(xmi_inputFPtr *inputFPtr) {
  if (!inputFPtr || !*inputFPtr) return;
  xmi_inputF *in = (xmi_inputF *)*inputFPtr;
  for (int i = 0; i < in->n_layers; i++) free_layer(&in->layers[i]);
  for (int i = 0; i < in->n_exc_layers; i++) free_layer(&in->exc_layers[i]);
  for (int i = 0; i < in->n_det_layers; i++) free_layer(&in->det_layers[i]);
  for (int i = 0; i < in->n_crystal_layers; i++) free_layer(&in->crystal_layers[i]);
  free(in->layers);
  free(in->exc_layers);
  free(in->det_layers);
  free(in->crystal_layers);
  free(in->discrete);
  free(in->continuous);
  free(in);
  *inputFPtr = NULL;
}

int xmi_input_C2F(xmi_input *c, xmi_inputFPtr *out) {
  if (!c || !out || !c->general || !c->composition || !c->geometry || !c->excitation ||
      !c->absorbers || !c->detector) {
    xmi_set_error("xmi_input_C2F: incomplete input");
    return 0;
  }
  xmi_inputF *in = (xmi_inputF *)calloc(1, sizeof(xmi_inputF));
  if (!in) return 0;
  in->general = *c->general;
  const xmi_composition *comp = c->composition;
  if (comp->n_layers < 1 || comp->reference_layer < 1 || comp->reference_layer > comp->n_layers) {
    xmi_set_error("xmi_input_C2F: invalid composition (n_layers=%d reference_layer=%d)", comp->n_layers,
                  comp->reference_layer);
    free(in);
    return 0;
  }
  in->n_layers = comp->n_layers;
  in->reference_layer = comp->reference_layer;
  in->layers = (xmi_layerF *)calloc(comp->n_layers, sizeof(xmi_layerF));
  for (int i = 0; i < comp->n_layers; i++)
    if (!copy_layer(&comp->layers[i], &in->layers[i], 1)) goto fail;
  in->geometry = *c->geometry;
  in->n_discrete = c->excitation->n_discrete;
  in->n_continuous = c->excitation->n_continuous;
  if (in->n_discrete > 0) {
    in->discrete = (xmi_energy_discrete *)malloc(sizeof(xmi_energy_discrete) * in->n_discrete);
    memcpy(in->discrete, c->excitation->discrete, sizeof(xmi_energy_discrete) * in->n_discrete);
  }
  if (in->n_continuous > 0) {
    in->continuous = (xmi_energy_continuous *)malloc(sizeof(xmi_energy_continuous) * in->n_continuous);
    memcpy(in->continuous, c->excitation->continuous, sizeof(xmi_energy_continuous) * in->n_continuous);
  }
  if (in->n_discrete + in->n_continuous < 1) {
    xmi_set_error("xmi_input_C2F: no excitation");
    goto fail;
  }
  in->n_exc_layers = c->absorbers->n_exc_layers;
  if (in->n_exc_layers > 0) {
    in->exc_layers = (xmi_layerF *)calloc(in->n_exc_layers, sizeof(xmi_layerF));
    for (int i = 0; i < in->n_exc_layers; i++)
      if (!copy_layer(&c->absorbers->exc_layers[i], &in->exc_layers[i], 1)) goto fail;
  }
  in->n_det_layers = c->absorbers->n_det_layers;
  if (in->n_det_layers > 0) {
    in->det_layers = (xmi_layerF *)calloc(in->n_det_layers, sizeof(xmi_layerF));
    for (int i = 0; i < in->n_det_layers; i++)
      if (!copy_layer(&c->absorbers->det_layers[i], &in->det_layers[i], 1)) goto fail;
  }
  in->detector = *c->detector;
  in->n_crystal_layers = c->detector->n_crystal_layers;
  if (in->n_crystal_layers > 0) {
    in->crystal_layers = (xmi_layerF *)calloc(in->n_crystal_layers, sizeof(xmi_layerF));
    for (int i = 0; i < in->n_crystal_layers; i++)
      if (!copy_layer(&c->detector->crystal_layers[i], &in->crystal_layers[i], 1)) goto fail;
  }
  *out = in;
  return 1;
fail : {
  xmi_inputFPtr p = in;
  xmi_free_input_F(&p);
  return 0;
}
}

static void cross3(const double a[3], const double b[3], double r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

/* 3x3 inverse by cofactors (src/xmi_aux_f.F90:2076-2106) */
static void inverse3(const double m[3][3], double inv[3][3]) {
  double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
               m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
  inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
  inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
  inv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
  inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
  inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
  inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
  inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
  inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
}

static void layer_coordinates(xmi_inputF *in) {
  const double nz = in->geometry.n_sample_orientation[2];
  for (int j = 0; j < in->n_layers; j++) in->layers[j].thickness_along_Z = fabs(in->layers[j].thickness / nz);
  int r = in->reference_layer - 1;
  in->layers[r].Z_coord_begin = in->geometry.d_sample_source;
  in->layers[r].Z_coord_end = in->layers[r].thickness_along_Z + in->geometry.d_sample_source;
  for (int j = r + 1; j < in->n_layers; j++) {
    in->layers[j].Z_coord_begin = in->layers[j - 1].Z_coord_end;
    in->layers[j].Z_coord_end = in->layers[j].Z_coord_begin + in->layers[j].thickness_along_Z;
  }
  for (int j = r - 1; j >= 0; j--) {
    in->layers[j].Z_coord_end = in->layers[j + 1].Z_coord_begin;
    in->layers[j].Z_coord_begin = in->layers[j].Z_coord_end - in->layers[j].thickness_along_Z;
  }
}

int xmi_init_input(xmi_inputFPtr *inputFPtr) {
  if (!inputFPtr || !*inputFPtr) return 0;
  xmi_inputF *in = (xmi_inputF *)*inputFPtr;
  xmi_geometry *g = &in->geometry;
  xmi_normalize3(g->n_sample_orientation);
  if (g->n_sample_orientation[2] < 0.0)
    for (int i = 0; i < 3; i++) g->n_sample_orientation[i] = -g->n_sample_orientation[i];
  xmi_normalize3(g->n_detector_orientation);

  in->detector_radius = sqrt(g->area_detector / M_PI);
  if (g->collimator_height > 0.0 && g->collimator_diameter > 0.0) {
    in->collimator_present = 1;
    in->collimator_radius = g->collimator_diameter / 2.0;
    if (in->collimator_radius >= in->detector_radius) {
      xmi_set_error("Non conical collimator found");
      return 0;
    }
    in->half_apex = atan((in->detector_radius - in->collimator_radius) / g->collimator_height);
    in->vertex[0] = in->detector_radius / tan(in->half_apex);
    in->vertex[1] = 0.0;
    in->vertex[2] = 0.0;
  } else {
    in->collimator_present = 0;
  }

  double *nx = in->n_detector_orientation_new_x, *ny = in->n_detector_orientation_new_y,
         *nz = in->n_detector_orientation_new_z;
  for (int i = 0; i < 3; i++) nx[i] = g->n_detector_orientation[i];
  const double ex[3] = {1.0, 0.0, 0.0}, ey[3] = {0.0, 1.0, 0.0};
  if (fabs(nx[0]) > 1.0e-6) cross3(ey, nx, ny);
  else cross3(ex, nx, ny);
  xmi_normalize3(ny);
  cross3(nx, ny, nz);
  for (int i = 0; i < 3; i++) {
    in->n_detector_orientation_new[i][0] = nx[i];
    in->n_detector_orientation_new[i][1] = ny[i];
    in->n_detector_orientation_new[i][2] = nz[i];
  }
  inverse3(in->n_detector_orientation_new, in->n_detector_orientation_inverse);

  double dx = g->p_detector_window[0], dy = g->p_detector_window[1], dz = g->p_detector_window[2] - g->d_sample_source;
  double distance = sqrt(dx * dx + dy * dy + dz * dz);
  double half_apex = atan(in->detector_radius / distance);
  in->detector_solid_angle = 2.0 * M_PI * (1.0 - cos(half_apex));
  for (int i = 0; i < 3; i++) {
    double s = 0.0;
    for (int k = 0; k < 3; k++) s += in->n_detector_orientation_inverse[i][k] * g->n_sample_orientation[k];
    in->n_sample_orientation_det[i] = s;
  }
  layer_coordinates(in);
  in->initialised = 1;
  return 1;
}

static int cmp_int(const void *a, const void *b) { return (*(const int *)a) - (*(const int *)b); }

int xmi_unique_Z(const xmi_inputF *in, int *uniqZ) {
  int n = 0;
  for (int i = 0; i < in->n_layers; i++)
    for (int j = 0; j < in->layers[i].n_elements; j++) {
      int Z = in->layers[i].Z[j], seen = 0;
      for (int k = 0; k < n; k++)
        if (uniqZ[k] == Z) seen = 1;
      if (!seen) uniqZ[n++] = Z;
    }
  qsort(uniqZ, n, sizeof(int), cmp_int);
  return n;
}
