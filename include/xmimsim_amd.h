/*
 * xmimsim_amd.h -- C ABI of the MI355X photon-history engine.
 *
 * This is a drop-in for the xmi_main_msim path of XMI-MSIM 8.1: the data
 * structures below have the same member order and types as the reference's
 * public C structs (include/xmi_data_structs.h:36-495,
 * include/xmi_solid_angle.h:27-35), and the entry points take the same
 * arguments, return 1 on success / 0 on failure, and allocate their outputs
 * with malloc (release with xmi_free).  Nothing in this header depends on GLib,
 * torch or HIP.
 */
#ifndef XMIMSIM_AMD_H
#define XMIMSIM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- data model: include/xmi_data_structs.h ---------------------------- */

typedef struct _xmi_general {
  float version;
  char *outputfile;
  long n_photons_interval;
  long n_photons_line;
  int n_interactions_trajectory;
  char *comments;
} xmi_general;

typedef struct _xmi_layer {
  int n_elements;
  int *Z;
  double *weight;
  double density;
  double thickness;
} xmi_layer;

typedef struct _xmi_composition {
  int n_layers;
  xmi_layer *layers;
  int reference_layer; /* 1-based */
} xmi_composition;

typedef struct _xmi_geometry {
  double d_sample_source;
  double n_sample_orientation[3];
  double p_detector_window[3];
  double n_detector_orientation[3];
  double area_detector;
  double collimator_height;
  double collimator_diameter;
  double d_source_slit;
  double slit_size_x;
  double slit_size_y;
} xmi_geometry;

typedef enum {
  XMI_ENERGY_DISCRETE_DISTRIBUTION_MONOCHROMATIC,
  XMI_ENERGY_DISCRETE_DISTRIBUTION_GAUSSIAN,
  XMI_ENERGY_DISCRETE_DISTRIBUTION_LORENTZIAN
} XmiEnergyDiscreteDistribution;

typedef struct _xmi_energy_discrete {
  double energy;
  double horizontal_intensity;
  double vertical_intensity;
  double sigma_x;
  double sigma_xp;
  double sigma_y;
  double sigma_yp;
  XmiEnergyDiscreteDistribution distribution_type;
  double scale_parameter;
} xmi_energy_discrete;

typedef struct _xmi_energy_continuous {
  double energy;
  double horizontal_intensity;
  double vertical_intensity;
  double sigma_x;
  double sigma_xp;
  double sigma_y;
  double sigma_yp;
} xmi_energy_continuous;

typedef struct _xmi_excitation {
  int n_discrete;
  xmi_energy_discrete *discrete;
  int n_continuous;
  xmi_energy_continuous *continuous;
} xmi_excitation;

typedef struct _xmi_absorbers {
  int n_exc_layers;
  xmi_layer *exc_layers;
  int n_det_layers;
  xmi_layer *det_layers;
} xmi_absorbers;

typedef enum {
  XMI_DETECTOR_CONVOLUTION_PROFILE_SILI,
  XMI_DETECTOR_CONVOLUTION_PROFILE_GE,
  XMI_DETECTOR_CONVOLUTION_PROFILE_SI_SDD
} XmiDetectorConvolutionProfile;

typedef struct _xmi_detector {
  XmiDetectorConvolutionProfile detector_type;
  double live_time;
  double pulse_width;
  double gain;
  double zero;
  double fano;
  double noise;
  int nchannels;
  int n_crystal_layers;
  xmi_layer *crystal_layers;
} xmi_detector;

typedef struct _xmi_input {
  xmi_general *general;
  xmi_composition *composition;
  xmi_geometry *geometry;
  xmi_excitation *excitation;
  xmi_absorbers *absorbers;
  xmi_detector *detector;
} xmi_input;

/* include/xmi_data_structs.h:479-495 */
typedef struct _xmi_main_options {
  int use_M_lines;
  int use_cascade_auger;
  int use_cascade_radiative;
  int use_variance_reduction;
  int use_sum_peaks;
  int use_escape_peaks;
  int escape_ratios_mode;
  int verbose;
  int use_poisson;
  int use_gpu;
  int omp_num_threads;
  int extra_verbose;
  char *custom_detector_response;
  int use_advanced_compton;
  int use_default_seeds;
} xmi_main_options;

/* include/xmi_solid_angle.h:27-35; solid_angles is double[theta_n][r_n] */
typedef struct _xmi_solid_angle {
  double *solid_angles;
  long int grid_dims_r_n;
  long int grid_dims_theta_n;
  double *grid_dims_r_vals;
  double *grid_dims_theta_vals;
  char *xmi_input_string;
} xmi_solid_angle;

/* opaque handles: include/xmi_data.h:71 and src/xmi_aux_f.F90 (xmi_inputFPtr) */
typedef void *xmi_inputFPtr;
typedef void *xmi_hdf5FPtr;

/* ---- entry points ------------------------------------------------------ */

/* Default options; replaces xmi_main_options_new (src/xmi_data_structs.c:2531-2566).
 * The returned struct is malloc'ed. */
xmi_main_options *xmi_main_options_new(void);

/* Deep copy of an xmi_input with weight fractions normalised per layer and the
 * derived geometry left for xmi_init_input.  Replaces xmi_input_C2F
 * (src/xmi_aux_f.F90:782-1025).  Returns 1 on success. */
int xmi_input_C2F(xmi_input *xmi_inputC, xmi_inputFPtr *xmi_inputFPtr_out);

/* Geometry set-up: detector radius and frames, collimator cone, layer
 * coordinates.  Replaces xmi_init_input (src/xmi_main.F90:1741-1918). */
int xmi_init_input(xmi_inputFPtr *inputFPtr);

/* Releases a handle made by xmi_input_C2F (src/xmi_aux_f.F90:1027). */
void xmi_free_input_F(xmi_inputFPtr *inputFPtr);

/* Writes the interaction tables for the listed elements in the xmimsimdata.h5
 * layout.  Replaces xmi_db (include/xmi_data.h:85, src/xmi_data.c:149-486);
 * the physics comes from the built-in analytic provider. */
int xmi_db(char *filename, int *Zs, int nZ);

/* Reads the tables of the input's elements.  Replaces xmi_init_from_hdf5
 * (include/xmi_data.h:88, src/xmi_data_f.F90:164-848). */
int xmi_init_from_hdf5(char *hdf5_file, xmi_inputFPtr inputFPtr, xmi_hdf5FPtr *hdf5FPtr,
                       xmi_main_options *options);

/* Same tables without a file round trip (generated in memory). */
int xmi_init_from_provider(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr *hdf5FPtr, xmi_main_options *options);

/* Where xmi_init_from_provider and xmi_db run the heavy loops of the table
 * generator (angle, Compton-profile and azimuth inverse CDFs;
 * src/xmi_data_f.F90:996-1237,1483-1549): -1 (default) on the GPU when one is
 * usable, else on the host; 0 on the host; 1 on the GPU (an error without
 * one).  No counterpart in the reference, whose generator runs on the host. */
void xmi_tables_set_device(int mode);

/* Copies one table of a table handle: which = 0 Rayleigh and 1 Compton angle
 * inverse CDF of element Z ([400][2000]), 2 its total Compton-profile inverse
 * CDF, 3 its partial-profile inversions ([shells][10001]), 4 the azimuth
 * inverse CDF ([200][2000], Z ignored).  Returns the number of values (and
 * copies them when out holds at least n), 0 when absent.  No counterpart in
 * the reference (which reads the same tables from its HDF5 file). */
long xmi_tables_copy(xmi_hdf5FPtr hdf5FPtr, int Z, int which, double *out, long n);

/* Binds the input's layers to the loaded tables.  Replaces
 * xmi_update_input_from_hdf5 (include/xmi_data.h:93, src/xmi_data_f.F90:98-162). */
int xmi_update_input_from_hdf5(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr);

/* Releases a table handle (include/xmi_data.h:91). */
void xmi_free_hdf5_F(xmi_hdf5FPtr *hdf5FPtr);

/* Solid-angle grid of the detector, computed on the GPU (1024 x 1024 cells,
 * xmi_hits_per_single trials per cell).  Drop-in for the OpenCL plugin entry
 * xmi_solid_angle_calculation_cl (src/xmi_solid_angle_cl.c:120-439), which
 * itself replaces xmi_solid_angle_calculation_f (src/xmi_solid_angle_f.F90:303).
 * *solid_angle is malloc'ed; release with xmi_free_solid_angle.  The grid
 * keeps input_string (the XMSI text of the input, malloc'ed by the caller, as
 * the reference's wrapper hands the plugin a copy) as its xmi_input_string
 * and frees it with the grid; NULL is allowed. */
int xmi_solid_angle_calculation_cl(xmi_inputFPtr inputFPtr, xmi_solid_angle **solid_angle,
                                   char *input_string, xmi_main_options *xmo);
void xmi_free_solid_angle(xmi_solid_angle *solid_angle);

/* HDF5 cache of solid-angle grids (include/xmi_solid_angle.h:41-50,
 * src/xmi_solid_angle.c:193-821), in the reference's file layout.  A stored
 * grid needs its xmi_input_string.  xmi_find_solid_angle_match sets *rv to
 * the newest stored grid whose input matches A (or NULL) and returns 1; a
 * file without the kind/version attributes, or older than 7.2, is recreated
 * empty.  xmi_get_solid_angle_file: *filePtr NULL selects
 * $XDG_DATA_HOME (or ~/.local/share)/XMI-MSIM/xmimsim-solid-angles.h5. */
int xmi_create_empty_solid_angle_hdf5_file(char *hdf5_file);
int xmi_update_solid_angle_hdf5_file(char *hdf5_file, xmi_solid_angle *solid_angle);
int xmi_check_solid_angle_match(xmi_input *input_in, xmi_input *input_h5);
int xmi_find_solid_angle_match(char *hdf5_file, xmi_input *A, xmi_solid_angle **rv, xmi_main_options *options);
int xmi_get_solid_angle_file(char **filePtr, int create_file);

/* Trials per solid-angle cell; the reference's global hits_per_single
 * (src/xmi_solid_angle_f.F90:43-44), default 5000. */
extern long xmi_hits_per_single;

/* The photon-history engine on the GPU.  Replaces xmi_main_msim
 * (include/xmi_main.h:29, src/xmi_main.F90:66-954).
 * Outputs (malloc'ed, owned by the caller, release with xmi_free):
 *   *channels        double[n_interactions+1][nchannels], cumulative over
 *                    interaction order, scaled by live_time;
 *   *brute_history   double[100][385][n_interactions]: the weight of the
 *                    photons that reached the detector, by element, row of
 *                    their last interaction (line, 384 Rayleigh, 385 Compton)
 *                    and its order; zero with variance reduction, which
 *                    never produces detector hits;
 *   *var_red_history double[100][385][n_interactions], or NULL when variance
 *                    reduction is off.
 * With options->use_variance_reduction = 0 the brute-force path runs
 * (analogue transport, detector and collimator tests of every path segment,
 * Auger and radiative cascades with their offspring photons per
 * use_cascade_auger / use_cascade_radiative) and solid_angles may be NULL.
 * Random numbers: Philox4x32-10, one stream per photon history, keyed by the
 * seed (options->use_default_seeds selects seed 1, otherwise
 * xmi_set_seed / time-based).  Returns 1 on success, 0 on failure. */
int xmi_main_msim(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr, int n_mpi_hosts, double **channels,
                  xmi_main_options *options, double **brute_history, double **var_red_history,
                  xmi_solid_angle *solid_angles);

/* Many samples in one call; no counterpart in the reference, whose batch mode
 * (src/xmi_batch.c) runs one xmimsim process per sample.  A work queue of
 * samples, largest photon budget first, is served by n_streams engines (at
 * most 8), each with its own HIP stream, so that samples overlap on the GPU.
 * All inputs must be bound to the one table handle; each sample has its
 * solid-angle grid (one grid may serve every sample of one geometry).
 * Outputs per sample as xmi_main_msim's (channels[k], var_red_history[k],
 * malloc'ed); done_ms[k], when given, receives the sample's completion time in
 * ms after the start.  seeds[k], when seeds is given, is sample k's seed;
 * otherwise every sample uses the seed xmi_main_msim would.  Each sample's
 * result equals xmi_main_msim's for the same seed.  Afterwards xmi_msim_get_stats on the calling thread holds the
 * batch's totals (histories, detection-weight events, wall time).  Returns 1
 * on success, 0 (no outputs) on failure. */
int xmi_main_msim_batch(int n_samples, xmi_inputFPtr *inputs, xmi_hdf5FPtr hdf5FPtr, int n_mpi_hosts,
                        xmi_main_options *options, xmi_solid_angle **solid_angles, double **channels,
                        double **var_red_history, int n_streams, double *done_ms, const uint64_t *seeds);

/* Releases memory returned by this library (src/xmi_aux.c:217). */
void xmi_free(void *ptr);

/* Seed used by the next xmi_main_msim / xmi_solid_angle_calculation_cl call
 * when options->use_default_seeds is 0 (stands in for the /dev/urandom seed
 * acquisition of src/xmi_random.c:37-120). */
void xmi_set_seed(uint64_t seed);

/* Timing and counters of the last xmi_main_msim call on this thread. */
typedef struct _xmi_msim_stats {
  double transport_ms;        /* device time of all transport kernels (events) */
  double total_ms;            /* wall time of xmi_main_msim */
  int64_t histories;          /* primary photons simulated */
  int64_t launches;           /* with profiling on: launches of each kernel class */
  double vr_kernel_ms;        /* with profiling on: device time of the detection-weight kernels */
  int64_t vr_events;          /* detection-weight evaluations (photon-interactions) */
  double step_kernel_ms;      /* with profiling on: forced-step kernel (incl. queue reset); brute force:
                                 the analogue step with its detector tests and tallies */
  double interaction_kernel_ms; /* with profiling on: Rayleigh + Compton + photo kernels */
  int64_t solid_angle_fallbacks; /* estimator points off the solid-angle grid */
  double vr_geom_ms;          /* with profiling on: the estimator's four kernels */
  double vr_cone_ms;
  double vr_scatter_ms;
  double vr_lines_ms;
  int64_t grid_uploads;       /* solid-angle grid copies to the device (0 when the resident copy was current) */
  int64_t photon_steps;       /* brute force: analogue steps taken (every order, offspring included) */
  int64_t detector_hits;      /* brute force: photons that reached the detector */
  int64_t batches;            /* batches of the run order simulated */
  int64_t graph_replays;      /* of them, replays of the run's captured graph (0: launched one by one) */
} xmi_msim_stats;
void xmi_msim_get_stats(xmi_msim_stats *out);

/* Times every kernel class with HIP events (synchronising once per batch)
 * when on; used by the benchmark to measure the detection-weight kernel. */
void xmi_msim_set_profiling(int on);

/* Launch shape of the transport (no counterpart in the reference, whose
 * OpenMP loop has no launch shape; for callers that tune the engine to a
 * workload, such as xmimsim_amd.autotune).  Results do not depend on
 * batch_log2 or use_graph beyond the last bits of the estimator's sums;
 * lines_kp and wg_per_cu change only the order of those sums.  Zero fields
 * take the defaults.  Applies to the calls that start after it; not to be
 * changed while a call runs. */
typedef struct _xmi_msim_launch {
  int batch_log2; /* histories per batch 2^batch_log2, 12..21 (default 21) */
  int use_graph;  /* 1: capture one batch's launches in a HIP graph, replayed per batch; each replay pops its
                     batch from a device batch queue */
  int lines_kp;   /* photons per lane of the estimator's line kernel: 2..4 up to 4 layers (default 3),
                     1..2 up to 8 (default 2), 1 beyond */
  int wg_per_cu;  /* resident workgroups per CU of the estimator's grid-stride kernels (default: occupancy) */
} xmi_msim_launch;
int xmi_msim_set_launch(const xmi_msim_launch *launch); /* 0 and an error message for a shape out of range */
void xmi_msim_get_launch(xmi_msim_launch *launch);

/* Detection-weight evaluations per layer of the last xmi_main_msim call on
 * this thread; writes min(n, 16) values, returns how many. */
int xmi_msim_get_layer_events(int64_t *out, int n);

/* ---- physics provider queries ----------------------------------------
 * No counterpart in the reference, which calls xraylib directly; these expose
 * the built-in provider that replaces xraylib (line names of
 * src/xmi_lines.c's numbering, cross sections in cm2/g, energies in keV). */
const char *xmi_line_name(int line);              /* 1..383, 384 Rayleigh, 385 Compton */
int xmi_provider_atom(int Z, void *xrl_atom_out);  /* copies the provider's per-element record */
double xmi_provider_cs(int kind, int Z, double E); /* 0 total, 1 photo, 2 Rayleigh, 3 Compton */
double xmi_provider_dcs(int kind, int Z, double E, double theta); /* 0 Rayleigh, 1 Compton, cm2/g/sr */
int xmi_provider_shell_cross_sections(int Z, int cascade_type, double E, double *P9);
/* 0: fluorescence yield of shell index (0 K .. 8 M5); 1: radiative rate and
 * 2: energy (keV) of line index (xmi_line_name numbering) */
double xmi_provider_line_data(int Z, int which, int index);
/* The Auger transition trans (0 .. 153) of element Z, the equivalent of
 * xraylib's AugerRate: writes the vacancy shell it empties and the two it
 * fills (0 K .. 8 M5, 9 any shell beyond M5) and returns its rate within
 * that shell's Auger decays; -1 for an invalid index. */
double xmi_provider_auger(int Z, int trans, int *shells3);

/* Number of fluorescence lines (1..line_last, line_last <= 219) of element Z
 * that the detection-weight estimator evaluates: energy >= 1 keV and a
 * non-zero yield x radiative rate in the built-in physics provider. */
int xmi_provider_active_lines(int Z, int line_last);

/* Last error message of this thread, or "" */
const char *xmi_msim_last_error(void);

/* ---- detector response ------------------------------------------------ */

/* Escape ratios of the detector crystal (include/xmi_detector.h:28-40). */
typedef struct _xmi_escape_ratios {
  int n_elements;
  int n_fluo_input_energies;
  int n_compton_input_energies;
  int n_compton_output_energies;
  int *Z;
  double *fluo_escape_ratios;          /* Fortran (element, line 1..109, input energy) */
  double *fluo_escape_input_energies;
  double *compton_escape_ratios;       /* Fortran (input energy, output energy) */
  double *compton_escape_input_energies;
  double *compton_escape_output_energies;
  char *xmi_input_string;
} xmi_escape_ratios;

/* Options of the escape-ratio Monte Carlo (include/xmi_detector.h:42-51);
 * defaults 1990 input energies from 1 keV by 0.1 keV, 1999 Compton output
 * energies from 0.1 keV by 0.1 keV, 500000 photons per input energy. */
typedef struct _xmi_escape_ratios_options {
  long n_input_energies;
  long n_compton_output_energies;
  long n_photons;
  double input_energy_min;
  double input_energy_delta;
  double compton_output_energy_min;
  double compton_output_energy_delta;
} xmi_escape_ratios_options;

/* include/xmi_detector.h:53 */
xmi_escape_ratios_options xmi_get_default_escape_ratios_options(void);

/* Escape ratios of the input's detector crystal on the GPU.  Replaces
 * xmi_escape_ratios_calculation (include/xmi_detector.h:57,
 * src/xmi_detector.c:91-141, src/xmi_main.F90:5473-5801): the crystal layers
 * are the sample, hit at normal incidence by a pencil beam of each input
 * energy; each photon's one forced interaction is followed by the analogue
 * test of whether its secondary leaves the crystal.  Fluorescence ratios per
 * (element, line KL1..L3P3, input energy) and Compton ratios per (input,
 * output energy), normalised by the weight of the photons that interacted.
 * hdf5_file NULL takes the tables from the built-in provider.  *escape_ratios
 * is NULL on failure (xmi_msim_last_error holds the reason). */
void xmi_escape_ratios_calculation(xmi_input *input, xmi_escape_ratios **escape_ratios, char *input_string,
                                   char *hdf5_file, xmi_main_options *options, xmi_escape_ratios_options ero);

/* include/xmi_detector.h:67 */
void xmi_free_escape_ratios(xmi_escape_ratios *escape_ratios);

/* HDF5 cache of escape ratios (include/xmi_detector.h:59-73,
 * src/xmi_detector.c:40-618), in the reference's file layout; same
 * conventions as the solid-angle cache (xmimsim-escape-ratios.h5).  A stored
 * result matches an input whose detector crystal has the same layers. */
int xmi_create_empty_escape_ratios_hdf5_file(char *hdf5_file);
int xmi_update_escape_ratios_hdf5_file(char *hdf5_file, xmi_escape_ratios *escape_ratios);
int xmi_check_escape_ratios_match(xmi_input *input_in, xmi_input *input_h5);
int xmi_find_escape_ratios_match(char *hdf5_file, xmi_input *A, xmi_escape_ratios **rv, xmi_main_options *options);
int xmi_get_escape_ratios_file(char **filePtr, int create_file);

/* Detector response of every interaction order: absorption in the window and
 * crystal, escape peaks (when options->use_escape_peaks and escape_ratios is
 * given), pile-up (use_sum_peaks), Gaussian response with low-energy tail and
 * Poisson noise (use_poisson); the histories get the detector absorption at
 * their line energies in place.  channels_noconv[n] are the spectra of
 * xmi_main_msim per order; channels_conv[n] receive malloc'ed results (order
 * 0 is zero-filled when zero_interaction is 0).  Replaces
 * xmi_detector_convolute_all (include/xmi_main.h:33,
 * src/xmi_detector_f.F90:219-580); xmi_detector_convolute_all_custom is the
 * same function under the name the reference loads a custom detector
 * response by (bin/xmimsim.c:516). */
void xmi_detector_convolute_all(xmi_inputFPtr inputFPtr, double **channels_noconv, double **channels_conv,
                                double *brute_history, double *var_red_history, xmi_main_options *options,
                                xmi_escape_ratios *escape_ratios, int n_interactions_all, int zero_interaction);
void xmi_detector_convolute_all_custom(xmi_inputFPtr inputFPtr, double **channels_noconv, double **channels_conv,
                                       double *brute_history, double *var_red_history, xmi_main_options *options,
                                       xmi_escape_ratios *escape_ratios, int n_interactions_all,
                                       int zero_interaction);

/* The same detector response without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (harnesses that forbid hipMalloc, tight serving loops).  The caller owns
 * the outputs: channels_conv[0..n_interactions_all] must each point to
 * nchannels doubles.  Device work runs on the null stream in a file-scope
 * device array of the library (one per device, at most 65536 channels:
 * 2.6 MB); calls that overlap on one device share it and race.  Returns 1 on
 * success, 0 with xmi_msim_last_error set (more channels than the maximum,
 * a NULL output, a device error, a pile-up train over 100 pulses). */
int xmi_detector_convolute_all_noalloc(xmi_inputFPtr inputFPtr, double **channels_noconv, double **channels_conv,
                                       double *brute_history, double *var_red_history, xmi_main_options *options,
                                       xmi_escape_ratios *escape_ratios, int n_interactions_all,
                                       int zero_interaction);

/* ---- net-line surrogate ------------------------------------------------
 * No counterpart in the reference.  An ensemble of multilayer perceptrons
 * (dims[0] inputs, GELU hidden layers, linear outputs) trained on the GPU
 * with BF16 matrix-core products, FP32 master weights and Adam on a masked
 * mean-square loss.  Layer widths dims[1..n_layers] must be multiples of 64,
 * batches multiples of 64 up to max_batch; inputs are padded internally. */
typedef struct xmi_surrogate xmi_surrogate;
xmi_surrogate *xmi_surrogate_create(int n_nets, int n_layers, const int *dims, int max_batch, uint64_t seed);
void xmi_surrogate_free(xmi_surrogate *s);
long xmi_surrogate_param_count(const xmi_surrogate *s);      /* per net */
int xmi_surrogate_padded_width(const xmi_surrogate *s, int layer);
/* one Adam step of every net on the batch; loss[n_nets] before the step */
int xmi_surrogate_train_step(xmi_surrogate *s, const float *X, const float *T, const float *mask, int batch,
                             float lr, float *loss);
/* Training data kept on the device: n rows of inputs X[n][dims[0]], targets
 * T[n][dims[n_layers]] and mask[n][dims[n_layers]]; replaces an earlier set. */
int xmi_surrogate_set_data(xmi_surrogate *s, const float *X, const float *T, const float *mask, long n);
/* steps Adam steps of every net on mini-batches of the resident data: step k
 * uses rows idx[k][0..batch) and learning rate lr[k]; loss[steps][n_nets]
 * receives each step's loss before its update.  The same arithmetic as
 * xmi_surrogate_train_step on those rows, without a host transfer or a
 * synchronisation per step. */
int xmi_surrogate_train_resident(xmi_surrogate *s, const int32_t *idx, int steps, int batch, const float *lr,
                                 float *loss);
/* loss and parameter gradients without an update; grad[n_nets][param_count] */
int xmi_surrogate_gradients(xmi_surrogate *s, const float *X, const float *T, const float *mask, int batch,
                            float *loss, float *grad);
/* Y[n_nets][batch][dims[n_layers]] */
int xmi_surrogate_predict(xmi_surrogate *s, const float *X, int batch, float *Y);
/* f32 parameters per net: for each layer W[N][K padded] then b[N] */
int xmi_surrogate_get_params(const xmi_surrogate *s, float *out);
int xmi_surrogate_set_params(xmi_surrogate *s, const float *in);

/* ---- XMSI / XMSO files: include/xmi_xml.h, include/xmi_data_structs.h ----
 * The input file (.xmsi) and the output file (.xmso) of xml/xmimsim-1.0.dtd.
 * Errors: the functions return NULL / 0 and, when error is not NULL and
 * *error is NULL, set *error to a malloc'ed xmi_error (release with
 * xmi_error_free).  Its leading members have the layout of GLib's GError
 * (domain, code, message), so that callers written against the reference's
 * GError ** can read error->message unchanged. */

typedef struct _xmi_error {
  uint32_t domain;
  int code;
  char *message;
} xmi_error;

void xmi_error_free(xmi_error *error);

/* include/xmi_data_structs.h:335-366 */
typedef struct _xmi_counts {
  double counts;
  int interaction_number;
} xmi_counts;

typedef struct _xmi_fluorescence_line {
  char *line_type;
  double energy;
  double total_counts;
  int n_interactions;
  xmi_counts *interactions;
} xmi_fluorescence_line;

typedef struct _xmi_fluorescence_line_counts {
  int atomic_number;
  double total_counts;
  int n_lines;
  xmi_fluorescence_line *lines;
} xmi_fluorescence_line_counts;

/* include/xmi_data_structs.h:386-404 */
typedef struct _xmi_output {
  float version;
  char *inputfile;
  char *outputfile;
  xmi_input *input;
  xmi_fluorescence_line_counts *brute_force_history;
  xmi_fluorescence_line_counts *var_red_history;
  int nbrute_force_history;
  int nvar_red_history;
  double **channels_conv;   /* [ninteractions + 1][nchannels] */
  double **channels_unconv; /* [ninteractions + 1][nchannels] */
  int ninteractions;
  int use_zero_interactions;
} xmi_output;

/* include/xmi_data_structs.h:508-515 */
typedef enum {
  XMI_INPUT_GENERAL = 1,
  XMI_INPUT_COMPOSITION = 2,
  XMI_INPUT_GEOMETRY = 4,
  XMI_INPUT_EXCITATION = 8,
  XMI_INPUT_ABSORBERS = 16,
  XMI_INPUT_DETECTOR = 32
} XmiInputFlags;

/* src/xmi_xml.c:1286 and :1814.  Weights are normalised and sorted by Z,
 * zero weights dropped, discrete energies given twice merged; the input must
 * pass xmi_input_validate.  Release with xmi_input_free. */
xmi_input *xmi_input_read_from_xml_file(const char *xmsifile, xmi_error **error);
xmi_input *xmi_input_read_from_xml_string(const char *xmsistring, xmi_error **error);
/* src/xmi_xml.c:1404 and :1363 (the string is malloc'ed; release with xmi_free) */
int xmi_input_write_to_xml_file(xmi_input *input, const char *xmsifile, xmi_error **error);
int xmi_input_write_to_xml_string(xmi_input *input, char **xmlstring, xmi_error **error);
/* src/xmi_data_structs.c: 0 when valid, else the OR of the failing sections */
int xmi_input_validate(xmi_input *input);
/* frees an xmi_input built by this library's readers (every member malloc'ed) */
void xmi_input_free(xmi_input *input);

/* src/xmi_data_structs.c:1371-1519: the output record of one simulation.
 * brute_history and var_red_history are double[100][385][n_int] (the
 * layout xmi_main_msim returns; var_red_history may be NULL),
 * channels_conv[n_int + 1][nchannels] (xmi_detector_convolute_all) and
 * channels_unconv double[n_int + 1][nchannels] (xmi_main_msim).  Line
 * energies are the physics provider's.  The input is copied.  Release with
 * xmi_output_free. */
xmi_output *xmi_output_raw2struct(xmi_input *input, double *brute_history, double *var_red_history,
                                  double **channels_conv, double *channels_unconv, char *inputfile,
                                  int use_zero_interactions);
/* src/xmi_xml.c:1638 (with the preview graphs) and :2033 */
int xmi_output_write_to_xml_file(xmi_output *output, const char *xmsofile, xmi_error **error);
xmi_output *xmi_output_read_from_xml_file(const char *xmsofile, xmi_error **error);
void xmi_output_free(xmi_output *output);
/* element symbol of Z (1..103), NULL outside; xraylib's AtomicNumberToSymbol */
const char *xmi_element_symbol(int Z);

#ifdef __cplusplus
}
#endif
#endif
