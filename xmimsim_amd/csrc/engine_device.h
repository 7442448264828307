/*
 * engine_device.h -- device-side data model and helpers of the photon-history
 * engine (gfx950).
 *
 * Layout in HBM (all arrays structure-of-arrays, double precision):
 *   photon state       PhotonSoA, one slot per history of the current batch
 *   angle tables       ray/comp ICDF [nZ][400][2000] (r contiguous: one
 *                      bilinear gather = two 16-byte pairs)
 *   azimuth table      Phi ICDF [200][2000]
 *   solid-angle grid   [theta][r] 1024 x 1024
 *   per-element data   xrl_atom[nZ] (edges, yields, rates, cross-section
 *                      tables), interaction-probability tables with a uniform
 *                      index accelerator, Compton profile ICDF [nZ][10001]
 *   detection tables   active fluorescence lines per element (line, shell,
 *                      yield*rate, channel) and mu*rho per layer
 *   tallies            128-bit fixed point (two u64 words), order independent
 */
#ifndef XMIMSIM_AMD_ENGINE_DEVICE_H
#define XMIMSIM_AMD_ENGINE_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.h"
#include "xrl_model.h"

#define ENG_MAX_LAYERS 16
/* interaction orders per trajectory: the Philox counter word of a history
 * stream keeps the order in 8 bits (philox.h) */
#define ENG_MAX_INTERACTIONS 255
#define ENG_MAX_ACT 96          /* active fluorescence lines per element */
#define ENG_IP_ACC 32768        /* interaction-probability accelerator cells (about one per table point) */
#define ENG_PI 3.14159265358979323846

struct PhotonSoA {
  double *x, *y, *z;
  double *dx, *dy, *dz;
  double *ex, *ey, *ez;
  double *theta, *phi;
  double *energy, *weight;
  double *mus;                  /* [n_layers][cap] */
  int *layer;                   /* 1-based, 0 = dead */
  int *elem;                    /* selected element index within layer */
  int *prev_line;               /* line of the previous interaction if fluorescence, else 0 */
  int *prev_u;                  /* hdf5 index of the element that emitted it */
  uint64_t *hid;                /* history id */
  int *cflags;                  /* brute-force runs: cascades left (bit 0 Auger, bit 1 radiative) */
};

/* per-photon records of the estimator at the current order, written by the
 * geometry kernel and read by the cone, scattering and line kernels */
/* one ln(E) cell of a layer's attenuation table: the Rayleigh + Compton
 * value at the cell's start and its increment to the next cell, and, when no
 * absorption edge of the layer lies in the cell, the photo-absorption
 * coefficient of every energy in it (edge_free; else K, the number of edges
 * above the cell, starts the edge walk) */
struct LayCell {
  double T0, dT, pb;
  int K, edge_free;
};

struct VrSoA {
  double *theta, *phi;          /* scattering angle to the detector point, azimuth to the polarisation */
  double *omega;                /* solid angle of the detector */
  double *pesc;                 /* escape probability at the photon energy */
  double *dist;                 /* [NL][cap] path in the k-th layer crossed (k = 0: own layer) */
  double *sar, *sat;            /* off-grid point (r, theta) for the cone estimate */
  int *info;                    /* bit 0 active, bit 1 upwards, bits 8.. layers crossed */
  unsigned int *blk;            /* estimator stream position (Philox block) after the solid angle */
};

struct EngineParams {
  /* geometry */
  double n_sample[3];
  double p_det[3];
  double n_det[3];
  double det_new[3][3];
  double det_inv[3][3];
  double det_radius, coll_radius, coll_height;
  int coll_present;
  double det_vertex[3];         /* collimator cone in the detector frame: vertex */
  double det_cos2;              /* cos^2 of its half apex */
  /* brute-force runs (use_variance_reduction = 0): analogue transport, a
   * photon is tallied when it reaches the detector; a cascade's offspring
   * photon of slot i lives in slot i + off_half (0: no cascades) */
  int brute;
  int casc;                     /* cascades of a primary: bit 0 Auger, bit 1 radiative */
  int off_half;
  double slit_x, slit_y, d_source_slit;
  int n_layers;
  int n_int;
  int nch;
  double gain, zero;
  int line_last;
  int cascade;
  int use_M_lines;
  long hits_per_single;
  uint64_t seed;
  uint32_t sample;
  int nZ;
  /* layers */
  double layer_density[ENG_MAX_LAYERS];
  double layer_zb[ENG_MAX_LAYERS];
  double layer_ze[ENG_MAX_LAYERS];
  int layer_nel[ENG_MAX_LAYERS];
  int layer_off[ENG_MAX_LAYERS];
  const int *el_Z;              /* flattened per layer */
  const double *el_w;
  const int *el_u;              /* hdf5 index */
  /* attenuation of each layer as one lookup: mu_l(E) = E^-p * pb[K] + rc(E),
   * K = number of the layer's edges above E (edges sorted descending), rc the
   * weighted sum of the elements' Rayleigh+Compton tables on the shared ln E
   * grid; kstart[ci] = edges certainly above any E of ln E cell ci */
  const double *lay_rc;         /* [n_layers][XRL_NCS] */
  const double *lay_edge;       /* [lay_eoff[l] .. lay_eoff[l+1]) */
  const double *lay_pb;         /* [lay_eoff[l] + l .. lay_eoff[l+1] + l] */
  const int *lay_kstart;        /* [n_layers][XRL_NCS] */
  const struct LayCell *lay_cell; /* [n_layers][XRL_NCS] */
  const struct LayCell *el_cell;  /* [elements of the layers, layer_off order][XRL_NCS]: each element's own
                                     total cross section (K counts its own edges) */
  int lay_eoff[ENG_MAX_LAYERS + 1];
  /* per element */
  const xrl_atom *atoms;        /* [nZ] */
  const double *ray_icdf;       /* [nZ][n_e][n_r] */
  const double *comp_icdf;
  const double *tab_E;          /* [n_e] shared energy grid */
  const double *tab_R;          /* [n_r] */
  int n_e, n_r;
  double tab_E0, tab_Einv, tab_R0, tab_Rinv;
  const double *fycorr;         /* [nZ][9] */
  const double *ip_E;           /* concatenated */
  const double *ip_r;
  const double *ip_rc;
  const int *ip_off;            /* [nZ+1] */
  const int *ip_acc;            /* [nZ][ENG_IP_ACC] */
  double ip_acc_E0, ip_acc_inv;
  const double *cp_icdf;        /* [nZ][cp_len] */
  const double *cp_rn;          /* [cp_len] shared */
  int cp_len;
  const double *cp_cdfg;        /* [nZ][XMI_CP_GRID] profile CDF on a uniform momentum grid */
  const double *cp_gscale;      /* [nZ] grid points per unit momentum */
  /* the advanced Compton model (use_advanced_compton): the occupied
   * subshells of element u are g = cp_sh_off[u] .. cp_sh_off[u+1]-1, with
   * the provider subshell cp_sh_idx[g], the partial-profile CDF
   * cp_pcdf[g][cp_len] on the Q grid cp_Qs and its inverse cp_qinv[g][cp_len]
   * on the CDF grid cp_inv (both grids uniform and shared) */
  int adv_compton;
  const int *cp_sh_off, *cp_sh_idx;
  const double *cp_pcdf, *cp_qinv, *cp_Qs, *cp_inv;
  const double *precalc_xrf;    /* [nZ][9][nZ][219] */
  const double *line_E;         /* [nZ][219] */
  const double *precalc_mu;     /* [n_layers][nZ][219] */
  /* detection tables */
  const int *act_n;             /* [nZ] */
  const int *act_line;          /* [nZ][ENG_MAX_ACT] */
  const int *act_shell;
  const double *act_factor;     /* fluor_yield(shell) * rad_rate(line) */
  const double *act_E;
  const int *act_ch;
  const double *act_murho;      /* [n_layers][nZ][ENG_MAX_ACT] */
  /* the same lines flattened per layer: entry t of layer l (fl_off[l] <= t <
   * fl_off[l+1]) is active line q of element u of that layer */
  int fl_off[ENG_MAX_LAYERS + 1];
  const int4 *fl_info;          /* {u, q, shell, compact bin} */
  const double *fl_wf;          /* weight fraction of u in the layer * yield * rate */
  const double *fl_mr;          /* [entry][layer bound of the kernels] mu*rho of the line per layer, 0 past the stack */
  const int *lbin_g;            /* compact bin -> u * ENG_MAX_ACT + q */
  int n_lbins;
  /* azimuth and solid angle */
  const double *phi_icdf;       /* [n_t][n_pr] */
  const double *phi_T;
  const double *phi_R;
  int phi_nt, phi_nr;
  double phi_T0, phi_Tinv, phi_R0, phi_Rinv;
  const double *sa_grid;        /* [n_theta][n_r] */
  const double *sa_r;
  const double *sa_t;
  int sa_nr, sa_nt;
  double sa_r0, sa_rinv, sa_t0, sa_tinv;
  /* tallies */
  unsigned long long *t_ch;     /* [n_int][nch][2] */
  unsigned long long *t_line;   /* [n_int][nZ][ENG_MAX_ACT][2] */
  unsigned long long *t_rc;     /* [n_int][nZ][2][2] */
  unsigned long long *t_bch;    /* brute force: [n_int + 1][nch][2] detector hits by interactions so far */
  unsigned long long *t_bhist;  /* brute force: [n_int][nZ][385][2] by last interaction's order, element, row */
  int tally_shift;              /* fixed point: value * 2^tally_shift */
  unsigned long long *counters; /* [slot ENG_CNT_*][ENG_SPREAD]: queue sizes, events, fallbacks, errors */
  int *queue;                   /* [ENG_NQUEUES][ENG_SPREAD][qreg] */
  unsigned long long *qcount;   /* sizes of queue's sub-queues [ENG_NQUEUES][ENG_SPREAD]: counters' first slots,
                                   or (brute force) those of the queue set of this order */
  int lines_kp;                 /* launch shape (host only): photons per lane of the lines kernel, 0 = default */
  int wg_per_cu;                /* launch shape (host only): estimator workgroups per CU, 0 = occupancy */
  int qreg;                     /* entries per sub-queue */
  /* brute force: the photons alive at the start of order n are those of
   * order n - 1's interaction queues (pqueue, pqcount: the other queue set)
   * and the offspring that order's cascades started (olist_in, ocount_in);
   * this order's offspring go to olist_out / ocount_out */
  const int *pqueue;
  const unsigned long long *pqcount;
  const int *olist_in;
  const unsigned int *ocount_in;
  int *olist_out;
  unsigned int *ocount_out;
  /* layer lists of an order's estimator events: the slots of the photons
   * whose interaction lies in layer l, in slot order, so that every
   * wavefront of the scatter and line kernels holds photons of one layer */
  unsigned long long *cone_acc; /* [ENG_SPREAD * qreg] off-grid cone estimates: completed segments << 40 | hits */
  int *lay_list;                /* [n_layers][cap] */
  int *lay_woff;                /* [n_layers][n_waves]: per-wave counts, then exclusive offsets */
  int *lay_total;               /* [n_layers] entries of each list */
  int *lay_segbase;             /* [16][32] segment totals, then offsets of the per-wave scan */
  int n_waves;                  /* row length of lay_woff: waves of a one-thread-per-slot launch over cap */
  int cap;
};

/* ------------------------------------------------------------------------ */
/* 128-bit fixed-point accumulation: order independent, hence bit-reproducible */

__device__ __forceinline__ void fx_split(double v, int shift, unsigned long long *lo, unsigned long long *hi) {
  if (!(v > 0.0)) {
    *lo = 0ull;
    *hi = 0ull;
    return;
  }
  double t = ldexp(v, shift - 64);
  double th = trunc(t);
  *hi = (unsigned long long)th;
  *lo = (unsigned long long)ldexp(t - th, 64);
}

__device__ __forceinline__ void fx_add_global(unsigned long long *bin, double v, int shift) {
  unsigned long long lo, hi;
  fx_split(v, shift, &lo, &hi);
  if (lo) {
    unsigned long long old = atomicAdd(bin, lo);
    if (old + lo < old) hi += 1ull;
  }
  if (hi) atomicAdd(bin + 1, hi);
}

__device__ __forceinline__ void fx_add_shared(unsigned long long *bin, double v, int shift) {
  unsigned long long lo, hi;
  fx_split(v, shift, &lo, &hi);
  if (lo) {
    unsigned long long old = atomicAdd(bin, lo);
    if (old + lo < old) hi += 1ull;
  }
  if (hi) atomicAdd(bin + 1, hi);
}

/* adds a 128-bit (lo,hi) pair held elsewhere into a global bin */
__device__ __forceinline__ void fx_merge_global(unsigned long long *bin, unsigned long long lo, unsigned long long hi) {
  if (lo) {
    unsigned long long old = atomicAdd(bin, lo);
    if (old + lo < old) hi += 1ull;
  }
  if (hi) atomicAdd(bin + 1, hi);
}

/* ------------------------------------------------------------------------ */
/* vector helpers */

__device__ __forceinline__ double d_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ double d_norm(const double *a) { return sqrt(d_dot(a, a)); }
__device__ __forceinline__ void d_normalize(double *a) {
  double n = d_norm(a);
  a[0] /= n;
  a[1] /= n;
  a[2] /= n;
}
__device__ __forceinline__ void d_matvec(const double m[3][3], const double *v, double *r) {
  for (int i = 0; i < 3; i++) r[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
}
__device__ __forceinline__ double d_distance(const double *a, const double *b) {
  return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}
__device__ __forceinline__ int d_intersect(const double *pp, const double *pn, const double *lp, const double *ld,
                                           double *out) {
  double itn = d_dot(ld, pn);
  if (itn == 0.0) return 0;
  double diff[3] = {pp[0] - lp[0], pp[1] - lp[1], pp[2] - lp[2]};
  double d = d_dot(diff, pn) / itn;
  for (int i = 0; i < 3; i++) out[i] = d * ld[i] + lp[i];
  return 1;
}
__device__ __forceinline__ double d_interp(double x0, double y0, double x1, double y1, double x) {
  return y0 + ((y1 - y0) * (x - x0) / (x1 - x0));
}

/* index on a (nearly) uniform grid with the semantics of the reference's
 * findpos: a[p] < x <= a[p+1]; -1 when x > a[n-1] */
__device__ __forceinline__ int d_findpos_uniform(const double *a, int n, double a0, double inv_h, double x) {
  if (fabs(x - a[0]) < 1e-10) return 0;
  if (x > a[n - 1]) return -1;
  int p = (int)floor((x - a0) * inv_h);
  if (p < 0) p = 0;
  if (p > n - 2) p = n - 2;
  while (p > 0 && x <= a[p]) p--;
  while (p < n - 2 && x > a[p + 1]) p++;
  return p;
}

__device__ __forceinline__ double d_bilinear(const double *arr, const double *a1, int n1, double a10, double a1inv,
                                             const double *a2, int n2, double a20, double a2inv, double x1, double x2) {
  int p1 = d_findpos_uniform(a1, n1, a10, a1inv, x1);
  int p2 = d_findpos_uniform(a2, n2, a20, a2inv, x2);
  if (p1 < 0) p1 = n1 - 2;
  if (p2 < 0) p2 = n2 - 2;
  double a1l = a1[p1], a1h = a1[p1 + 1], a2l = a2[p2], a2h = a2[p2 + 1];
  double denom = (a1h - a1l) * (a2h - a2l);
  double c1 = (a1h - x1) * (a2h - x2) / denom;
  double c2 = (x1 - a1l) * (a2h - x2) / denom;
  double c3 = (a1h - x1) * (x2 - a2l) / denom;
  double c4 = (x1 - a1l) * (x2 - a2l) / denom;
  const double *r0 = arr + (size_t)p1 * n2 + p2;
  const double *r1 = arr + (size_t)(p1 + 1) * n2 + p2;
  return c1 * r0[0] + c2 * r1[0] + c3 * r0[1] + c4 * r1[1];
}

/* phx stream positioned at draw index k */
__device__ __forceinline__ void d_stream_seek(phx_stream *s, uint32_t k) {
  s->ctr[0] = k >> 1;
  s->avail = 0;
  if (k & 1u) {
    uint32_t o[4];
    phx_4x32_10(s->ctr, s->key, o);
    s->ctr[0]++;
    s->buf[0] = phx_u01(o[0], o[1]);
    s->buf[1] = phx_u01(o[2], o[3]);
    s->avail = 1;
  }
}

#endif
