/*
 * xmimsim_chain.c -- a C caller of the library's drop-in chain, compiled
 * against include/xmimsim_amd.h only.  It runs the steps of the reference
 * harness bin/xmimsim.c in its order:
 *
 *   read the .xmsi                       bin/xmimsim.c:262   xmi_input_read_from_xml_file
 *   C -> F copy, derived geometry        :275-283            xmi_input_C2F, xmi_init_input
 *   interaction tables (HDF5 file)       :249,287-294        xmi_db (when the file is missing),
 *                                                            xmi_init_from_hdf5, xmi_update_input_from_hdf5
 *   solid-angle grid, through its cache  :300-356            xmi_find_solid_angle_match, else
 *                                                            xmi_solid_angle_calculation_cl +
 *                                                            xmi_update_solid_angle_hdf5_file
 *   the photon histories                 :361                xmi_main_msim
 *   escape ratios, through their cache   :457-495            xmi_find_escape_ratios_match, else
 *                                                            xmi_escape_ratios_calculation +
 *                                                            xmi_update_escape_ratios_hdf5_file
 *   detector response                    :496-530            xmi_detector_convolute_all
 *   output record and .xmso              :642-651            xmi_output_raw2struct,
 *                                                            xmi_output_write_to_xml_file
 *
 * usage: xmimsim_chain in.xmsi out.xmso tables.h5 solid_angles.h5 escape_ratios.h5
 *                      n_photons_line seed_grid seed_msim hits_per_single escape_photons
 * (n_photons_line > 0 overrides the input's; escape_photons is the escape
 * run's photons per input energy).  Prints one line per step and exits 0 on
 * success.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../../include/xmimsim_amd.h"

#define FAIL(...)                        \
  do {                                   \
    fprintf(stderr, "xmimsim_chain: ");  \
    fprintf(stderr, __VA_ARGS__);        \
    fprintf(stderr, "\n");               \
    return 1;                            \
  } while (0)

static int add_Z(const xmi_layer *layers, int n_layers, int *Zs, int n) {
  for (int l = 0; l < n_layers; l++)
    for (int k = 0; k < layers[l].n_elements; k++) {
      int z = layers[l].Z[k], seen = 0;
      for (int j = 0; j < n; j++) seen |= Zs[j] == z;
      if (!seen) Zs[n++] = z;
    }
  return n;
}

/* the elements of the sample and of the detector crystal (the escape-ratio
 * run reads the crystal's tables from the same file) */
static int unique_Z(const xmi_input *in, int *Zs) {
  int n = add_Z(in->composition->layers, in->composition->n_layers, Zs, 0);
  return add_Z(in->detector->crystal_layers, in->detector->n_crystal_layers, Zs, n);
}

int main(int argc, char **argv) {
  if (argc != 11) FAIL("usage: %s in.xmsi out.xmso tables.h5 grids.h5 escape.h5 n_photons seed_grid seed_msim hits escape_photons", argv[0]);
  char *xmsi = argv[1], *xmso = argv[2], *tables = argv[3], *grids = argv[4], *escapes = argv[5];
  const long n_photons = atol(argv[6]);
  const uint64_t seed_grid = strtoull(argv[7], NULL, 10), seed_msim = strtoull(argv[8], NULL, 10);
  const long hits = atol(argv[9]), escape_photons = atol(argv[10]);
  xmi_error *error = NULL;
  setvbuf(stdout, NULL, _IONBF, 0); /* each step's line reaches the caller even if a later one fails */

  xmi_main_options *options = xmi_main_options_new();
  xmi_input *input = xmi_input_read_from_xml_file(xmsi, &error);
  if (!input) FAIL("reading %s: %s", xmsi, error ? error->message : "?");
  if (n_photons > 0) input->general->n_photons_line = n_photons;
  printf("read %s: %d layers, %d lines\n", xmsi, input->composition->n_layers, input->excitation->n_discrete);

  xmi_inputFPtr inputFPtr = NULL;
  xmi_hdf5FPtr hdf5FPtr = NULL;
  if (!xmi_input_C2F(input, &inputFPtr) || !xmi_init_input(&inputFPtr)) FAIL("input: %s", xmi_msim_last_error());

  if (access(tables, R_OK) != 0) {
    int Zs[256];
    int nZ = unique_Z(input, Zs);
    if (!xmi_db(tables, Zs, nZ)) FAIL("xmi_db: %s", xmi_msim_last_error());
    printf("tables of %d elements written to %s\n", nZ, tables);
  }
  if (!xmi_init_from_hdf5(tables, inputFPtr, &hdf5FPtr, options) || !xmi_update_input_from_hdf5(inputFPtr, hdf5FPtr))
    FAIL("tables: %s", xmi_msim_last_error());

  xmi_solid_angle *sa = NULL;
  if (access(grids, R_OK) != 0 && !xmi_create_empty_solid_angle_hdf5_file(grids)) FAIL("grid cache %s", grids);
  if (!xmi_find_solid_angle_match(grids, input, &sa, options)) FAIL("grid lookup: %s", xmi_msim_last_error());
  if (sa) {
    printf("solid-angle grid found in %s\n", grids);
  } else {
    char *s = NULL;
    if (!xmi_input_write_to_xml_string(input, &s, &error)) FAIL("input string: %s", error ? error->message : "?");
    xmi_hits_per_single = hits;
    xmi_set_seed(seed_grid);
    /* the grid keeps s and frees it with itself: the plugin entry takes the
     * copy that the reference's wrapper xmi_solid_angle_calculation makes
     * (src/xmi_solid_angle.c:183) */
    if (!xmi_solid_angle_calculation_cl(inputFPtr, &sa, s, options)) FAIL("grid: %s", xmi_msim_last_error());
    if (!xmi_update_solid_angle_hdf5_file(grids, sa)) FAIL("grid store: %s", xmi_msim_last_error());
    printf("solid-angle grid computed and stored in %s\n", grids);
  }

  double *channels = NULL, *brute = NULL, *var_red = NULL;
  xmi_set_seed(seed_msim);
  if (!xmi_main_msim(inputFPtr, hdf5FPtr, 1, &channels, options, &brute, &var_red, sa))
    FAIL("xmi_main_msim: %s", xmi_msim_last_error());
  xmi_msim_stats st;
  xmi_msim_get_stats(&st);
  printf("histories %lld in %.1f ms\n", (long long)st.histories, st.total_ms);

  const int n_int = input->general->n_interactions_trajectory;
  const int nch = (int)input->detector->nchannels;
  xmi_escape_ratios *er = NULL;
  if (options->use_escape_peaks) {
    if (access(escapes, R_OK) != 0 && !xmi_create_empty_escape_ratios_hdf5_file(escapes)) FAIL("escape cache");
    if (!xmi_find_escape_ratios_match(escapes, input, &er, options)) FAIL("escape lookup: %s", xmi_msim_last_error());
    if (er) {
      printf("escape ratios found in %s\n", escapes);
    } else {
      char *s = NULL;
      if (!xmi_input_write_to_xml_string(input, &s, &error)) FAIL("input string");
      xmi_escape_ratios_options ero = xmi_get_default_escape_ratios_options();
      ero.n_photons = escape_photons;
      xmi_escape_ratios_calculation(input, &er, s, tables, options, ero);
      if (!er) FAIL("escape ratios: %s", xmi_msim_last_error());
      if (!xmi_update_escape_ratios_hdf5_file(escapes, er)) FAIL("escape store: %s", xmi_msim_last_error());
      xmi_free(s);
      printf("escape ratios computed and stored in %s\n", escapes);
    }
  }

  double **noconv = malloc(sizeof(double *) * (n_int + 1));
  double **conv = malloc(sizeof(double *) * (n_int + 1));
  for (int n = 0; n <= n_int; n++) noconv[n] = channels + (size_t)n * nch;
  xmi_detector_convolute_all(inputFPtr, noconv, conv, brute, var_red, options, er, n_int, 0);
  printf("detector response done\n");
  for (int n = 0; n <= n_int; n++)
    if (!conv[n]) FAIL("detector response: %s", xmi_msim_last_error());

  xmi_output *out = xmi_output_raw2struct(input, brute, var_red, conv, channels, xmsi, 0);
  if (!out || !xmi_output_write_to_xml_file(out, xmso, &error)) FAIL("writing %s", xmso);
  printf("wrote %s\n", xmso);

  xmi_output_free(out);
  for (int n = 0; n <= n_int; n++) xmi_free(conv[n]);
  free(conv);
  free(noconv);
  if (er) xmi_free_escape_ratios(er);
  xmi_free(channels);
  xmi_free(brute);
  xmi_free(var_red);
  xmi_free_solid_angle(sa);
  xmi_free_hdf5_F(&hdf5FPtr);
  xmi_free_input_F(&inputFPtr);
  xmi_input_free(input);
  xmi_free(options);
  return 0;
}
