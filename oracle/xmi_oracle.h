/*
 * xmi_oracle.h -- CPU restatement of the reference's photon-history engine.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the product never does.
 */
#ifndef XMI_ORACLE_H
#define XMI_ORACLE_H

#include <stdint.h>

#include "../include/xmimsim_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rng_mode */
#define XMI_ORACLE_RNG_PER_HISTORY 0 /* Philox substreams per history, as on the GPU */
#define XMI_ORACLE_RNG_PER_THREAD 1  /* one sequential stream per thread, reference structure */

/* xmi_main_msim restated (src/xmi_main.F90:66-954, VR path).  Same outputs and
 * layouts as the product's xmi_main_msim.  n_threads OpenMP threads. */
int xmi_oracle_main_msim(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr, int n_mpi_hosts, double **channels,
                         xmi_main_options *options, double **brute_history, double **var_red_history,
                         xmi_solid_angle *solid_angles, int rng_mode, uint64_t seed, int n_threads);

/* per-run counters of the last xmi_oracle_main_msim call */
typedef struct {
  int64_t histories;
  int64_t interactions;
  int64_t vr_calls;
  int64_t solid_angle_fallbacks;
  double seconds;
} xmi_oracle_stats;
void xmi_oracle_get_stats(xmi_oracle_stats *s);

/* work of the detection-weight estimator in the last xmi_oracle_main_msim
 * call, summed over the events that reach its element loop: the inputs of the
 * algorithmic byte and flop accounting of the GPU kernels */
typedef struct {
  int64_t events;           /* estimator evaluations scored */
  int64_t elements;         /* sum of E, the elements of the event's layer */
  int64_t element_layers;   /* sum of E * L, L = layers on the escape path */
  int64_t element2_layers;  /* sum of E * E * L (attenuation at the Compton energy) */
  int64_t doppler_trials;   /* Doppler words drawn by the Compton terms */
  int64_t lines;            /* fluorescence lines evaluated (energy above threshold, shell open) */
  int64_t line_layers;      /* sum of L over those lines */
} xmi_oracle_work;
void xmi_oracle_get_work(xmi_oracle_work *w);

/* solid-angle grid restated (src/xmi_solid_angle_f.F90:62-710), one Philox
 * stream per cell; *out malloc'ed (free with xmi_oracle_free_solid_angle). */
int xmi_oracle_solid_angle_calculation(xmi_inputFPtr inputFPtr, xmi_solid_angle **out, long hits_per_single,
                                       uint64_t seed, int n_threads);
void xmi_oracle_free_solid_angle(xmi_solid_angle *sa);

/* grid extents only (xmi_solid_angle_inputs_f), for checking the GPU's */
int xmi_oracle_solid_angle_extents(xmi_inputFPtr inputFPtr, double *r_min, double *r_max, double *theta_min,
                                   double *theta_max);

/* one cell, reference algorithm (xmi_single_solid_angle_calculation) */
double xmi_oracle_single_solid_angle(xmi_inputFPtr inputFPtr, double r1, double theta1, long hits_per_single,
                                     uint64_t seed, uint64_t cell);

/* closed-form first-order intensity, tests/libxmimsim-test.c:153-180 (fpm) */
double xmi_oracle_fpm_first_order(xmi_inputFPtr inputFPtr, int Z, int line, int cascade_type);
/* first-order Rayleigh (kind 0) / Compton (kind 1) scatter of element Z, closed form */
double xmi_oracle_scatter_first_order(xmi_inputFPtr inputFPtr, int Z, int kind);
/* detector absorption and Gaussian response of one order (no escape, pile-up, noise) */
void xmi_oracle_detector_convolute(xmi_inputFPtr inputFPtr, const double *noconv, double *conv);

/* small numerical helpers exposed for unit tests */
int xmi_oracle_poly_solve_quadratic(double a, double b, double c, double *r1, double *r2);

/* escape ratios of a crystal (the escape input prepared by the caller),
 * src/xmi_main.F90:5473-5801; outputs in the reference's Fortran orders */
int xmi_oracle_escape_ratios(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr, xmi_main_options *options, long nE,
                             double E0, double dE, long nOut, double out0, double dOut, long n_photons, uint64_t seed,
                             int n_threads, double *fluo, double *compton);
double xmi_oracle_energy_from_q(double init_energy, double Q, double theta);
double xmi_oracle_q_from_energy(double old_energy, double new_energy, double theta);
double xmi_oracle_ran_trap(double x1, double x2, double y1, double y2, double u);
void xmi_oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

void xmi_oracle_free(void *p);

/* ---- probes of single restated routines (tests only) --------------------
 * Each runs one routine n times from the given state, run i on the Philox
 * stream (seed, substream 14, history i).  The photon sits in the middle of
 * layer 1 and interacts with element element_index of that layer. */

/* Rayleigh (kind 1, src/xmi_main.F90:1986-2101) or Compton (kind 2,
 * :2103-2229 with :4985-5069): out[7i..7i+6] = new dirv, new elecv, energy */
int xmi_oracle_probe_scatter(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr, xmi_main_options *options, int kind,
                             int element_index, double energy, const double *dirv, const double *elecv, long n,
                             uint64_t seed, double *out);
/* photo-absorption + fluorescence (:2231-2411, :5184-5437): emitted line
 * (0 when none), weight factor, and optionally dir[6i..] = dirv, elecv */
int xmi_oracle_probe_fluorescence(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr, xmi_main_options *options,
                                  int element_index, double energy, long n, uint64_t seed, int *line,
                                  double *weight, double *dir);
/* xmi_update_photon_dirv then xmi_update_photon_elecv (:5071-5182): out = dirv, elecv */
void xmi_oracle_probe_rotate(const double *dirv, const double *elecv, double theta, double phi, double *out);
/* azimuth of the polarisation in the photon's frame (:2034-2060) */
double xmi_oracle_probe_phi0(const double *dirv, const double *elecv);
/* primary photons of discrete line `which` (>= 0) or continuous interval
 * -which-1 (< 0), history i on stream (seed, 0, i) as in the engine:
 * out[10i..] = entry point, dirv, elecv, energy (NaN when outside the window) */
int xmi_oracle_probe_source(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr, xmi_main_options *options, int which,
                            long n, uint64_t seed, double *out);
/* the estimator's Doppler-broadened Compton energy (its truncated-profile
 * draw, one word each) of element element_index of layer 1 at (energy,
 * theta), n draws from one stream */
int xmi_oracle_probe_doppler(xmi_inputFPtr inputFPtr, xmi_hdf5FPtr hdf5FPtr, xmi_main_options *options,
                             int element_index, double energy, double theta, long n, uint64_t seed, double *out);
/* derived fields after xmi_init_input: out[0..34] detector radius, collimator
 * present/radius, half apex, vertex[3], n_sample_orientation_det[3],
 * n_sample_orientation[3], n_detector_orientation[3], new[3][3], inverse[3][3]
 * (row-major), detector solid angle; then per layer thickness along z,
 * z begin, z end */
void xmi_oracle_input_derived(xmi_inputFPtr inputFPtr, double *out);
/* the brute-force run's detector tests on n given cases: out[i] = 1 when a
 * photon at coords[i] moving along dirv[i] (lab frame) that has left the
 * sample reaches the detector (src/xmi_main.F90:1920-1984) */
void xmi_oracle_probe_detector_hit(xmi_inputFPtr inputFPtr, const double *coords, const double *dirv, long n,
                                   int *out);
/* out[i] = the code of src/xmi_aux_f.F90:1622-1833 for the path segment
 * begin[i] -> end[i] (lab frame): 0 none, 1 collimator, 2 detector window,
 * 3 window from behind */
void xmi_oracle_probe_detector_segment(xmi_inputFPtr inputFPtr, const double *begin, const double *end, long n,
                                       int *out);

/* planted mistakes: a test sets one, shows that its check rejects the
 * result, and resets it to XMI_ORACLE_PLANT_NONE */
#define XMI_ORACLE_PLANT_NONE 0
#define XMI_ORACLE_PLANT_RAYLEIGH_ENERGY 1     /* Rayleigh angle looked up at 1.15 E */
#define XMI_ORACLE_PLANT_PHI0_SIGN 2           /* polarisation azimuth with the wrong sign */
#define XMI_ORACLE_PLANT_COMPTON_FULL_ANGLE 3  /* Doppler shift with sin(theta) for sin(theta/2) */
#define XMI_ORACLE_PLANT_DOPPLER_ONE_SIDED 4   /* momentum sign never flipped */
#define XMI_ORACLE_PLANT_CK_L1_SWAP 5          /* FL12 and FL13 exchanged */
#define XMI_ORACLE_PLANT_LINE_SHIFT 6          /* radiative rate of the next line */
#define XMI_ORACLE_PLANT_ELECV_STALE 7         /* polarisation not re-orthogonalised */
#define XMI_ORACLE_PLANT_DIRV_PHI_SIGN 8       /* rotation by -phi */
#define XMI_ORACLE_PLANT_SLIT_FULL 9           /* slit half-angle not halved */
#define XMI_ORACLE_PLANT_GAUSS_XY 10           /* x divergence used for y */
#define XMI_ORACLE_PLANT_COMPTON_T_RAYLEIGH 11 /* Compton azimuth with the Rayleigh parameter */
#define XMI_ORACLE_PLANT_SHELL_NEXT 12         /* shell drawn with the next shell's cross section */
#define XMI_ORACLE_PLANT_QIMAX_IGNORED 13      /* advanced Compton: subshell profiles not truncated at Qimax */
#define XMI_ORACLE_PLANT_DOPPLER_NO_NEGATIVE 14 /* estimator: momenta of negative E' not drawn */
#define XMI_ORACLE_PLANT_BRUTE_BLBS 15         /* analogue step: transmission of the crossed layers dropped */
#define XMI_ORACLE_PLANT_AUGER_ONE_VACANCY 16  /* Auger cascade: second vacancy dropped */
#define XMI_ORACLE_PLANT_RADIATIVE_SHELL 17    /* radiative cascade: K line's vacancy one shell too low */
#define XMI_ORACLE_PLANT_NO_COLLIMATOR 18      /* detector hit test ignores the collimator */
void xmi_oracle_set_plant(int id);

/* trials of the Monte Carlo solid-angle fallback inside the estimator
 * (the reference's hits_per_single), default 5000 */
void xmi_oracle_set_hits_per_single(long hits);

#ifdef __cplusplus
}
#endif
#endif
