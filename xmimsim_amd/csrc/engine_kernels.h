/*
 * engine_kernels.h -- interface between the host driver (engine_host.cpp) and
 * the gfx950 kernels (engine_kernels.hip) of the photon-history engine.
 */
#ifndef XMIMSIM_AMD_ENGINE_KERNELS_H
#define XMIMSIM_AMD_ENGINE_KERNELS_H

#include "engine_device.h"

#define ENG_BLOCK 256
#define ENG_LDS_BYTES 65536        /* LDS budget of the estimator's tallies */
#define XMI_NLINES_VR_DEV 219

/* counters[] slots; every slot is ENG_SPREAD sub-counters and a wavefront
 * adds to sub-counter (global wave index mod ENG_SPREAD), so that same-address
 * atomics from all waves of a launch do not serialise on one location.  The
 * queue slots (0..ENG_NQUEUES-1) are ENG_SPREAD sub-queues of P.qreg entries each. */
#define ENG_SPREAD 64
#define ENG_CNT_Q_RAYL 0
#define ENG_CNT_Q_COMPT 1
#define ENG_CNT_Q_PHOTO 2
#define ENG_CNT_Q_SA 3              /* off-grid solid angles of the estimator */
#define ENG_NQUEUES 4
#define ENG_CNT_VR 4
#define ENG_CNT_FALLBACKS 5
#define ENG_CNT_ERRORS 6
#define ENG_CNT_HISTORIES 7
#define ENG_CNT_LAYER0 8            /* detection-weight events per layer */
#define ENG_CNT_ERR_SOURCE (ENG_CNT_LAYER0 + ENG_MAX_LAYERS)
#define ENG_CNT_ERR_VR_DOPPLER (ENG_CNT_ERR_SOURCE + 1)
#define ENG_CNT_ERR_DOPPLER (ENG_CNT_ERR_SOURCE + 2)
#define ENG_CNT_STEPS (ENG_CNT_ERR_SOURCE + 3)    /* brute force: photons that took a step */
#define ENG_CNT_HITS (ENG_CNT_ERR_SOURCE + 4)     /* brute force: photons that reached the detector */
#define ENG_NCOUNTERS (ENG_CNT_ERR_SOURCE + 5)

/* events recorded per order when profiling: start, after the step, after
 * each of the four estimator kernels, after the interactions */
#define ENG_ORDER_EVENTS 7

/* m_e c^2 in keV as the reference computes it (src/xmi_main.F90:48-55) */
#define ENG_MEC2 (9.10938188e-31 * 2.99792458e8 * 2.99792458e8 / 1.602176487e-19 / 1000.0)

struct SourceDesc {
  int kind;               /* 0 = continuous interval, 1 = discrete line */
  int valid;
  int dist;               /* XmiEnergyDiscreteDistribution for lines */
  int pol;                /* line polarisation: 0 from the intensities, 1 uniform
                             azimuth in the x-y plane (escape-ratio runs,
                             src/xmi_main.F90:5672-5687) */
  long count;             /* histories of this source */
  uint64_t hid0;          /* first history id */
  uint64_t run0;          /* position of its first history in the run order
                             (all sources back to back, empty ones skipped) */
  double x1, x2, y1, y2;  /* interval: energies and total intensities at the ends */
  double h1, h2;          /* interval: horizontal intensities at the ends */
  double energy, scale;   /* line */
  double hor_ver_ratio;   /* line: photons with 1-based index <= this are horizontal */
  double weight;          /* interval: before excitation-absorber correction; line: final */
  double sx, sxp, sy, syp;
};

struct SourceTable {
  const SourceDesc *d;
  int n;
  const double *line_mus;     /* [n_sources][n_layers] mus at the line energy */
  const double *exc_layers;   /* per absorber: density, thickness, n_el, Z[n_el], w[n_el] */
  int n_exc;
  const xrl_atom *exc_atoms;  /* [95] indexed by Z */
};

/* device batch queue of a run whose batches replay one captured graph: the
 * batches are the run order's positions in steps of the batch capacity;
 * start and count describe the batch being simulated */
struct BatchQueue {
  unsigned long long next, total, start;
  int count, pad;
};

/* escape-ratio runs (src/xmi_main.F90:5473-5801): fixed-point tallies of
 * the photons that leave the crystal after their one forced interaction */
#define ENG_ESC_LINES 109 /* K and L lines: KL1 (1) .. L3P3 (109) */
struct EscParams {
  unsigned long long *fluo;       /* [nE][ENG_ESC_LINES][nZ][2] */
  unsigned long long *compton;    /* [nOut][nE][2] */
  unsigned long long *interacted; /* [nE][2] */
  long n_photons;                 /* per input energy: history id / n_photons = energy index */
  int nE, nOut;
  double out0, dOut;              /* Compton output energy grid */
};

#ifdef __cplusplus
extern "C" {
#endif
hipError_t eng_launch_escape(const EngineParams *P, const PhotonSoA *S, const VrSoA *V, const EscParams *Q,
                             int count, hipStream_t st);
hipError_t eng_launch_grid(const EngineParams *P, double *grid, const double *rv, const double *tv, int n_r, int n_t,
                           long hits, uint64_t seed, hipStream_t st);
hipError_t eng_launch_source(const EngineParams *P, const PhotonSoA *S, const SourceTable *T, uint64_t start,
                             int count, hipStream_t st);
/* the batch queue: set up for `total` positions; then per batch the pop and
 * the source kernel over `cap` slots (those beyond the batch marked empty) */
hipError_t eng_launch_batch_init(BatchQueue *Q, unsigned long long total, hipStream_t st);
hipError_t eng_launch_source_queued(const EngineParams *P, const PhotonSoA *S, const SourceTable *T, BatchQueue *Q,
                                    int cap, hipStream_t st);
/* order n of a brute-force run: the analogue step with its detector tests
 * and tallies over a fixed grid of `blocks` workgroups (eng_brute_blocks)
 * that strides over the live photons, then (n <= n_int) the interactions */
hipError_t eng_launch_order_brute(const EngineParams *P, const PhotonSoA *S, int n, int count, int blocks,
                                  hipStream_t st, hipEvent_t *ev /* 3 events (before, between, after) or NULL */);
/* workgroups of the brute-force grid: one resident round of the step kernel */
int eng_brute_blocks(int n_cu);
hipError_t eng_launch_order(const EngineParams *P, const PhotonSoA *S, const VrSoA *V, int n, int count, int n_cu,
                            hipStream_t st, hipEvent_t *ev);
#ifdef __cplusplus
}
#endif
#endif
