/*
 * engine_kernels.hip -- gfx950 kernels of the photon-history engine.
 *
 * Event-based transport of one batch of histories (reference loop
 * src/xmi_main.F90:1223-1681, variance-reduction branch):
 *
 *   k_source        primary energy, direction, polarisation, entry point
 *                   (src/xmi_main.F90:319-716, 957-1186)
 *   per interaction order n = 1..n_int:
 *     k_step        forced interaction (src/xmi_main.F90:1414-1518), atom and
 *                   interaction-type selection (:1558-1652); appends the
 *                   photon to one of three queues (Rayleigh, Compton, photo),
 *                   counts each wave's events per layer, and computes the
 *                   estimator's geometry of the new point (d_vr_geom)
 *     k_layer_scan, k_layer_base   offsets of the per-layer event lists
 *     k_vr_list, k_vr_cone, k_vr_scatter, k_vr_lines
 *                   detection-weight estimator of every photon at order n
 *                   (src/xmi_variance_reduction.F90:29-726, 949-1101)
 *     k_interact    one launch over the three interaction queues:
 *                   Rayleigh src/xmi_main.F90:1986-2101, Compton :2103-2229
 *                   and 4985-5069, photo-absorption :2231-2411, 5184-5437
 *
 * Random numbers: Philox streams per history (philox.h); the transport stream
 * of order n is consumed as: step, atom, type, then the interaction kernel
 * continues at draw 3.  The estimator has its own stream per order.
 *
 * Regrouping by interaction type keeps every wavefront of the interaction
 * kernels on one code path.  All kernels are FP64 VALU work; nothing here is
 * a dense contraction, so no MFMA is used.
 */
#include <stdlib.h>

#include "engine_kernels.h"

/* segments of the per-wave layer-count scan (k_layer_scan, d_layer_list) */
#define ENG_SCAN_SEG 1024
#define ENG_SCAN_NSEG 32

/* ------------------------------------------------------------------------ */
/* physics helpers on the device */

/* mass attenuation of a layer, sum_i w_i CS_Total(Z_i, E) (src/xmi_aux_f.F90:1109-1141),
 * from the per-layer tables: equal to the element sum up to rounding, since
 * the layer sums are formed once on the host */
__device__ __forceinline__ double d_layer_mu_tab(const EngineParams &P, int layer0, const xrl_ectx *c) {
  if (!(c->E > 0.0)) return 0.0;
  const LayCell r = P.lay_cell[(size_t)layer0 * XRL_NCS + c->ci];
  double pb = r.pb;
  if (!r.edge_free) {
    /* an edge inside the cell: count the edges above E */
    const int eo = P.lay_eoff[layer0], ne = P.lay_eoff[layer0 + 1] - eo;
    int K = r.K;
    while (K < ne && P.lay_edge[eo + K] > c->E) K++;
    pb = P.lay_pb[eo + layer0 + K];
  }
  return c->pe * pb + (r.T0 + c->cf * r.dT);
}

__device__ __forceinline__ void d_mu_all(const EngineParams &P, double E, double *mus, size_t stride, size_t idx) {
  xrl_ectx c = xrl_energy(E);
  for (int l = 0; l < P.n_layers; l++) mus[(size_t)l * stride + idx] = d_layer_mu_tab(P, l, &c);
}

/* this wavefront's sub-counter of a counter slot (1-D launches) */
__device__ __forceinline__ unsigned long long *d_ctr(const EngineParams &P, int slot) {
  const unsigned int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  return &P.counters[(size_t)slot * ENG_SPREAD + (wave & (ENG_SPREAD - 1))];
}

/* wavefront-aggregated counter: call from converged code only */
__device__ __forceinline__ void d_wave_count(unsigned long long *ctr, int pred) {
  unsigned long long m = __ballot(pred);
  if (m != 0ull && __lane_id() == 0) atomicAdd(ctr, (unsigned long long)__popcll(m));
}

__device__ __forceinline__ void d_update_dirv(double *dirv, double *theta, double *phi, double theta_i, double phi_i) {
  double phi_new = phi_i;
  if (phi_i > 2.0 * ENG_PI) phi_new = phi_i - 2.0 * ENG_PI;
  else if (phi_i < 0.0) phi_new = phi_i + 2.0 * ENG_PI;
  double cp = cos(*phi), sp = sin(*phi), ct = cos(*theta), st = sin(*theta);
  double ts = sin(theta_i);
  double v0 = ts * cos(phi_new), v1 = ts * sin(phi_new), v2 = cos(theta_i);
  double d0 = ct * cp * v0 + (-sp) * v1 + st * cp * v2;
  double d1 = ct * sp * v0 + cp * v1 + st * sp * v2;
  double d2 = (-st) * v0 + 0.0 * v1 + ct * v2;
  double n = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  dirv[0] = d0 / n;
  dirv[1] = d1 / n;
  dirv[2] = d2 / n;
  *theta = acos(dirv[2]);
  double ph = atan2(dirv[1], dirv[0]);
  if (ph > 2.0 * ENG_PI) ph -= 2.0 * ENG_PI;
  else if (ph < 0.0) ph += 2.0 * ENG_PI;
  *phi = ph;
}

__device__ __forceinline__ void d_update_elecv(const double *dirv, double *elecv) {
  double cosalfa = d_dot(dirv, elecv);
  double c_alfa = acos(cosalfa);
  double sinalfa = sin(c_alfa);
  double c_ae = 1.0 / sinalfa;
  double c_be = -c_ae * cosalfa;
  for (int i = 0; i < 3; i++) elecv[i] = c_ae * elecv[i] + c_be * dirv[i];
  d_normalize(elecv);
}

__device__ __forceinline__ double d_phi0(double theta, double phi, const double *elecv) {
  double cp = cos(phi), sp = sin(phi), ct = cos(theta), st = sin(theta);
  double cosphi0 = elecv[0] * (cp * ct) + elecv[1] * (ct * sp) + elecv[2] * (-st);
  double sinphi0 = elecv[0] * sp + elecv[1] * (-cp) + elecv[2] * 0.0;
  if (fabs(cosphi0) > 1.0) cosphi0 = cosphi0 > 0 ? 1.0 : -1.0;
  double phi0 = acos(cosphi0);
  if (sinphi0 > 0.0) phi0 = -phi0;
  return phi0;
}

/* cone-sampling estimate of the detector solid angle seen from (r1, theta1),
 * src/xmi_solid_angle_f.F90:432-710 (same estimator as src/xmi_kernels.cl:219-452).
 * Split in a set-up and a per-trial test: trial i consumes exactly the
 * uniforms 2i and 2i+1 of the stream, so a wavefront can run the trials of one
 * estimate on all its lanes (trial i = Philox block first_block + i) and count
 * exactly the hits of the sequential loop. */
struct SaCone {
  double cone_sa, cos_apex, st, ct, lp1, lp2, R, Rc, H;
  int outside;
};

/* detector window radius and collimator (radius, height, present) */
struct SaGeom {
  double R, Rc, H;
  int coll;
};

__device__ __forceinline__ SaGeom d_sa_geom(const EngineParams &P) {
  SaGeom g;
  g.R = P.det_radius;
  g.Rc = P.coll_radius;
  g.H = P.coll_height;
  g.coll = P.coll_present;
  return g;
}

/* returns 0 when the collimator hides the detector (estimate 0, no draws) */
__device__ __forceinline__ int d_sa_setup(const SaGeom g, double r1, double theta1, SaCone *c) {
  double r, theta, fcbr;
  int outside = 0;
  const double R = g.R, Rc = g.Rc, H = g.H;
  if (!g.coll) {
    r = r1;
    theta = theta1;
    fcbr = R;
  } else if (fabs(Rc - R) < 0.000001) {
    if (r1 * cos(theta1) <= R) {
      r = r1;
      theta = theta1;
      fcbr = R;
    } else {
      r = sqrt(r1 * r1 - 2.0 * r1 * sin(theta1) * H + H * H);
      theta = acos(r1 * cos(theta1) / r);
      fcbr = Rc;
    }
    outside = r1 * sin(theta1) > H;
  } else {
    if (r1 * cos(theta1) <= R && r1 * sin(theta1) <= H * (r1 * cos(theta1) - R) / (Rc - R)) {
      r = r1;
      theta = theta1;
      fcbr = R;
    } else if (r1 * sin(theta1) <= H) {
      return 0;
    } else {
      r = sqrt(r1 * r1 - 2.0 * r1 * sin(theta1) * H + H * H);
      theta = acos(r1 * cos(theta1) / r);
      fcbr = Rc;
    }
    outside = r1 * sin(theta1) > H;
  }
  double beta = atan(fcbr / r);
  double alpha1 = atan(fcbr * sin(theta) / (r - fcbr * cos(theta)));
  if (alpha1 <= 0.0) alpha1 += ENG_PI;
  double apex = beta > alpha1 ? beta : alpha1;
  c->cos_apex = cos(apex);
  c->cone_sa = 2.0 * ENG_PI * (1.0 - c->cos_apex);
  c->st = sin(theta);
  c->ct = cos(theta);
  c->lp1 = r1 * cos(theta1);
  c->lp2 = r1 * sin(theta1);
  c->R = R;
  c->Rc = Rc;
  c->H = H;
  c->outside = outside;
  return 1;
}

/* one cone direction from two uniforms; 1 if it reaches the detector window.
 * theta = acos(1 - u1 (1 - cos apex)) enters only through its cosine and
 * sine, taken directly (cos theta is the argument, sin theta = sqrt(1 - cos^2)
 * >= 0 on [0, pi]); phi = 2 pi u2 through sincospi (no argument reduction) */
__device__ __forceinline__ int d_sa_trial(const SaCone &c, double u1, double u2) {
  const double x = 1.0 - u1 * (1.0 - c.cos_apex);
  const double sth = sqrt((1.0 - x) * (1.0 + x));
  double sp, cp;
  sincospi(2.0 * u2, &sp, &cp);
  double c0 = sth * cp, c1 = sth * sp, c2 = x;
  /* rotation matrix rows (1,0,0), (0,-sin,-cos), (0,cos,-sin) */
  double v0 = 1.0 * c0 + 0.0 * c1 + 0.0 * c2;
  double v1 = 0.0 * c0 + (-c.st) * c1 + (-c.ct) * c2;
  double v2 = 0.0 * c0 + c.ct * c1 + (-c.st) * c2;
  if (0.0 * v0 + 0.0 * v1 + 1.0 * v2 >= 0.0) return 0;
  if (c.outside) {
    double d = (c.H - c.lp2) / v2;
    double ix = d * v0 + 0.0, iy = d * v1 + c.lp1;
    if (sqrt(ix * ix + iy * iy + 0.0) > c.Rc) return 0;
  }
  double d = (0.0 - c.lp2) / v2;
  double ix = d * v0 + 0.0, iy = d * v1 + c.lp1;
  return sqrt(ix * ix + iy * iy + 0.0) <= c.R;
}

/* cone trials 2 b and 2 b + 1 from block b of a stream (its counter
 * word 0 advanced by b): words 0, 1 and 2, 3, one uniform each */
__device__ __forceinline__ int d_sa_trial_pair(const SaCone &c, const uint32_t *ctr, const uint32_t *key, uint32_t b,
                                               int both) {
  uint32_t o[4];
  const uint32_t cc[4] = {ctr[0] + b, ctr[1], ctr[2], ctr[3]};
  phx_4x32_10(cc, key, o);
  int n = d_sa_trial(c, phx_u32(o[0]), phx_u32(o[1]));
  if (both) n += d_sa_trial(c, phx_u32(o[2]), phx_u32(o[3]));
  return n;
}

/* sequential estimate (one thread per grid cell): hits trials from blocks
 * 0 .. ceil(hits / 2) - 1 of the cell's stream */
__device__ __forceinline__ double d_single_solid_angle_impl(const SaGeom g, double r1, double theta1, long hits,
                                                           phx_stream *rng) {
  SaCone c;
  if (!d_sa_setup(g, r1, theta1, &c)) return 0.0;
  long nhit = 0;
  for (long i = 0; i < hits; i += 2) nhit += d_sa_trial_pair(c, rng->ctr, rng->key, (uint32_t)(i >> 1), i + 1 < hits);
  return c.cone_sa * (double)nhit / (double)hits;
}


/* ------------------------------------------------------------------------ */
/* solid-angle grid: one thread per (r, theta) cell, Philox stream per cell */

__global__ void __launch_bounds__(ENG_BLOCK) k_solid_angle_grid(EngineParams P, double *grid, const double *rv,
                                                                 const double *tv, int n_r, int n_t, long hits,
                                                                 uint64_t seed) {
  long total = (long)n_r * n_t;
  for (long cell = blockIdx.x * (long)blockDim.x + threadIdx.x; cell < total; cell += (long)gridDim.x * blockDim.x) {
    int jt = (int)(cell / n_r), ir = (int)(cell % n_r);
    phx_stream s = {};
    phx_stream_init(&s, seed, 0u, 0u, PHX_SUB_GRID, (uint64_t)cell);
    grid[cell] = d_single_solid_angle_impl(d_sa_geom(P), rv[ir], tv[jt], hits, &s);
  }
}

/* ------------------------------------------------------------------------ */
/* source */

__device__ double d_gauss(phx_stream *s, double sigma) {
  double u1 = phx_uniform(s), u2 = phx_uniform(s);
  return sigma * sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * ENG_PI * u2);
}

__device__ double d_exc_corr(const SourceTable &S, double E) {
  double corr = 1.0;
  const double *p = S.exc_layers;
  for (int k = 0; k < S.n_exc; k++) {
    double dens = p[0], thick = p[1];
    int nel = (int)p[2];
    double mu = 0.0;
    for (int i = 0; i < nel; i++) mu += xrl_cs_total(&S.exc_atoms[(int)p[3 + i]], E) * p[3 + nel + i];
    corr *= exp(-1.0 * dens * thick * mu);
    p += 3 + 2 * nel;
  }
  return corr;
}

/* trapezoid sampler, src/xmi_aux_f.F90:1836-1941; returns 0 on failure */
__device__ int d_ran_trap(double x1, double x2, double y1, double y2, double u, double *out) {
  double m = (y2 - y1) / (x2 - x1);
  double denom = (x2 - x1) * (y1 - x1 * m) + m * (x2 * x2 - x1 * x1) / 2.0;
  double a = m / 2.0, b = y1 - x1 * m, c = -x1 * y1 + m * x1 * x1 / 2.0 - denom * u;
  double r1 = 0.0, r2 = 0.0;
  int nr;
  if (a == 0.0) {
    if (b == 0.0) return 0;
    r1 = -1.0 * c / b;
    nr = 1;
  } else {
    double delta = b * b - 4.0 * a * c;
    if (delta < 0.0) return 0;
    if (delta == 0.0) {
      r1 = r2 = -b / 2.0 / a;
    } else {
      double sd = sqrt(delta);
      double t1 = (-b + sd) / 2.0 / a, t2 = (-b - sd) / 2.0 / a;
      r1 = t1 < t2 ? t1 : t2;
      r2 = t1 < t2 ? t2 : t1;
    }
    nr = 2;
  }
  if (x1 <= r1 && r1 <= x2) {
    *out = r1;
    return 1;
  }
  if (nr == 2 && x1 <= r2 && r2 <= x2) {
    *out = r2;
    return 1;
  }
  return 0;
}

/* one primary photon; returns 1 when it enters the sample */
__device__ int d_source_one(const EngineParams &P, const PhotonSoA &S, const SourceTable &T, uint64_t batch_start,
                            int i) {
  const uint64_t run = batch_start + (uint64_t)i;
  int lo = 0, hi = T.n - 1;
  while (lo < hi) { /* last descriptor with run0 <= run: empty sources share
                       their successor's run0 and are skipped */
    int mid = (lo + hi + 1) >> 1;
    if (T.d[mid].run0 <= run) lo = mid;
    else hi = mid - 1;
  }
  const SourceDesc &d = T.d[lo];
  const uint64_t hid = d.hid0 + (run - d.run0);
  const size_t cap = (size_t)P.cap;
  S.layer[i] = 0;
  S.hid[i] = hid;
  S.prev_line[i] = 0;
  S.prev_u[i] = -1;
  if (P.brute) {
    S.cflags[i] = P.casc;
    if (P.off_half) S.layer[i + P.off_half] = 0; /* no offspring yet */
  }
  if (!d.valid) return 0;
  const long j = (long)(hid - d.hid0);
  phx_stream s;
  phx_stream_init(&s, P.seed, P.sample, 0u, PHX_SUB_SOURCE, hid);

  double E, w, hor = 0.0;
  if (d.kind == 0) {
    if (!d_ran_trap(d.x1, d.x2, d.y1, d.y2, phx_uniform(&s), &E)) {
      atomicAdd(d_ctr(P, ENG_CNT_ERRORS), 1ull);
      atomicAdd(d_ctr(P, ENG_CNT_ERR_SOURCE), 1ull);
      return 0;
    }
    hor = d_interp(d.x1, d.h1, d.x2, d.h2, E) / d_interp(d.x1, d.y1, d.x2, d.y2, E);
    w = d.weight * d_exc_corr(T, E);
    d_mu_all(P, E, S.mus, cap, i);
  } else {
    if (d.dist == 1) {
      E = d_gauss(&s, d.scale) + d.energy;
      if (E <= 1.0 || E > 200.0) return 0;
      d_mu_all(P, E, S.mus, cap, i);
    } else if (d.dist == 2) {
      E = d.scale * tan(ENG_PI * (phx_uniform(&s) - 0.5)) + d.energy;
      if (E <= 1.0 || E > 200.0) return 0;
      d_mu_all(P, E, S.mus, cap, i);
    } else {
      E = d.energy;
      for (int l = 0; l < P.n_layers; l++) S.mus[(size_t)l * cap + i] = T.line_mus[(size_t)lo * P.n_layers + l];
    }
    w = d.weight;
  }

  /* xmi_coords_dir, src/xmi_main.F90:957-1136 */
  double dirv[3], coords[3], x1, y1;
  if (fabs(d.sx * d.sy) < 1.0e-20) {
    double x1m = atan(P.slit_x / P.d_source_slit / 2.0);
    double y1m = atan(P.slit_y / P.d_source_slit / 2.0);
    x1 = x1m * (-1.0 + 2.0 * phx_uniform(&s));
    y1 = y1m * (-1.0 + 2.0 * phx_uniform(&s));
    coords[0] = coords[1] = coords[2] = 0.0;
  } else {
    x1 = d_gauss(&s, d.sxp);
    y1 = d_gauss(&s, d.syp);
    coords[0] = d_gauss(&s, d.sx) - P.d_source_slit * sin(x1);
    coords[1] = d_gauss(&s, d.sy) - P.d_source_slit * sin(y1);
    coords[2] = 0.0;
  }
  dirv[0] = tan(x1);
  dirv[1] = tan(y1);
  dirv[2] = 1.0;
  d_normalize(dirv);

  double ev[3];
  if (d.pol == 1) {
    /* escape-ratio runs: a random polarisation azimuth in the x-y plane */
    const double te = phx_uniform(&s) * ENG_PI * 2.0;
    ev[0] = cos(te);
    ev[1] = sin(te);
    ev[2] = 0.0;
  } else {
    int horizontal;
    if (d.kind == 0) horizontal = phx_uniform(&s) <= hor;
    else horizontal = (double)(j + 1) <= d.hor_ver_ratio;
    ev[0] = horizontal ? 0.0 : 1.0;
    ev[1] = horizontal ? 1.0 : 0.0;
    ev[2] = 0.0;
  }
  double cosalfa = d_dot(ev, dirv);
  if (fabs(cosalfa) > 1.0) {
    atomicAdd(d_ctr(P, ENG_CNT_ERRORS), 1ull);
      atomicAdd(d_ctr(P, ENG_CNT_ERR_SOURCE), 1ull);
    return 0;
  }
  double c_alfa = acos(cosalfa);
  double c_ae = 1.0 / sin(c_alfa);
  double c_be = -c_ae * cosalfa;
  for (int k = 0; k < 3; k++) ev[k] = c_ae * ev[k] + c_be * dirv[k];

  /* xmi_photon_shift_first_layer */
  int layer = 0;
  if (coords[2] >= P.layer_zb[0]) {
    for (int l = 0; l < P.n_layers; l++)
      if (coords[2] < P.layer_ze[l]) {
        layer = l + 1;
        break;
      }
    if (layer == 0) {
      atomicAdd(d_ctr(P, ENG_CNT_ERRORS), 1ull);
      atomicAdd(d_ctr(P, ENG_CNT_ERR_SOURCE), 1ull);
      return 0;
    }
  } else {
    double pp[3] = {0.0, 0.0, P.layer_zb[0]}, ip[3];
    if (!d_intersect(pp, P.n_sample, coords, dirv, ip)) {
      atomicAdd(d_ctr(P, ENG_CNT_ERRORS), 1ull);
      atomicAdd(d_ctr(P, ENG_CNT_ERR_SOURCE), 1ull);
      return 0;
    }
    coords[0] = ip[0];
    coords[1] = ip[1];
    coords[2] = ip[2];
    layer = 1;
  }
  S.x[i] = coords[0];
  S.y[i] = coords[1];
  S.z[i] = coords[2];
  S.dx[i] = dirv[0];
  S.dy[i] = dirv[1];
  S.dz[i] = dirv[2];
  S.ex[i] = ev[0];
  S.ey[i] = ev[1];
  S.ez[i] = ev[2];
  S.theta[i] = acos(dirv[2]);
  S.phi[i] = atan2(dirv[1], dirv[0]);
  S.energy[i] = E;
  S.weight[i] = w;
  S.layer[i] = layer;
  return 1;
}

/* the batch's primaries: positions batch_start .. batch_start + count - 1
 * of the run order, or, with a batch queue Q, those of the batch that
 * k_batch_next popped; then slots count .. limit - 1 (limit: the launch's
 * count argument) are marked empty */
__global__ void __launch_bounds__(ENG_BLOCK) k_source(EngineParams P, PhotonSoA S, SourceTable T,
                                                       uint64_t batch_start, int count, const BatchQueue *Q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, limit = count;
  if (Q) {
    batch_start = Q->start;
    count = Q->count;
  }
  int alive = 0;
  if (i < count) {
    alive = d_source_one(P, S, T, batch_start, i);
  } else if (i < limit) {
    S.layer[i] = 0;
    if (P.off_half) S.layer[i + P.off_half] = 0;
  }
  d_wave_count(d_ctr(P, ENG_CNT_HISTORIES), alive);
}

/* the device batch queue: k_batch_init before a run's first replay, then
 * k_batch_next, the first node of every replay, pops the next batch of at
 * most cap positions (one thread each) */
__global__ void k_batch_init(BatchQueue *Q, unsigned long long total) {
  Q->next = 0ull;
  Q->total = total;
  Q->start = 0ull;
  Q->count = 0;
}

__global__ void k_batch_next(BatchQueue *Q, int cap) {
  const unsigned long long s = Q->next, left = Q->total > s ? Q->total - s : 0ull;
  const int c = left < (unsigned long long)cap ? (int)left : cap;
  Q->start = s;
  Q->count = c;
  Q->next = s + (unsigned long long)c;
}

/* ------------------------------------------------------------------------ */
/* queues */

__global__ void k_reset_queues(EngineParams P) {
  if (threadIdx.x < ENG_NQUEUES * ENG_SPREAD) P.qcount[threadIdx.x] = 0ull;
  if (threadIdx.x == 0 && P.ocount_out) *P.ocount_out = 0u;
}

/* wavefront-aggregated append of the lanes with pred to queue t, in this
 * wave's sub-queue: one atomic per wave, spread over ENG_SPREAD counters.  A
 * sub-queue receives the photons of every ENG_SPREAD-th wave of a launch with
 * one thread per photon slot, so qreg = ceil(waves / ENG_SPREAD) * 64 entries
 * cannot overflow.  Call from converged code. */
__device__ __forceinline__ void d_append(const EngineParams &P, int t, int pred, int idx) {
  const unsigned int j = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) & (ENG_SPREAD - 1);
  unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  int lane = __lane_id();
  int leader = __ffsll((long long)mask) - 1;
  unsigned int base = 0;
  if (lane == leader)
    base = (unsigned int)atomicAdd(&P.qcount[(size_t)t * ENG_SPREAD + j], (unsigned long long)__popcll(mask));
  base = __shfl(base, leader);
  if (pred) {
    unsigned long long below = mask & ((1ull << lane) - 1ull);
    P.queue[((size_t)t * ENG_SPREAD + j) * P.qreg + base + __popcll(below)] = idx;
  }
}

/* the interaction-type queues (Rayleigh, Compton, photo) */
__device__ __forceinline__ void d_enqueue(const EngineParams &P, int type, int idx) {
  for (int t = 0; t < 3; t++) d_append(P, t, type == t, idx);
}

/* photon of queue entry q (0 <= q < ENG_SPREAD * qreg) of type t, or -1 */
__device__ __forceinline__ int d_dequeue(const EngineParams &P, int t, unsigned int q) {
  const unsigned int j = q / (unsigned int)P.qreg, r = q - j * (unsigned int)P.qreg;
  if (j >= ENG_SPREAD || r >= P.qcount[(size_t)t * ENG_SPREAD + j]) return -1;
  return P.queue[((size_t)t * ENG_SPREAD + j) * P.qreg + r];
}

/* layer lists: k_step stores each wave's number of events per layer (0 = no
 * event), k_layer_scan turns them into offsets, k_vr_list writes the slots.
 * Both launches run one thread per slot with the same workgroups, so a slot
 * has the same wave and lane in both.  Call from converged code. */
__device__ __forceinline__ void d_layer_count(const EngineParams &P, int layer) {
  const unsigned int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  for (int L = 1; L <= P.n_layers; L++) {
    const unsigned long long m = __ballot(layer == L);
    if (__lane_id() == 0) P.lay_woff[(size_t)(L - 1) * P.n_waves + w] = __popcll(m);
  }
}

__device__ __forceinline__ void d_layer_list(const EngineParams &P, int layer, int idx) {
  const unsigned int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const unsigned long long below = (1ull << __lane_id()) - 1ull;
  for (int L = 1; L <= P.n_layers; L++) {
    const unsigned long long m = __ballot(layer == L);
    if (layer == L) {
      const int pos = P.lay_woff[(size_t)(L - 1) * P.n_waves + w] +
                      P.lay_segbase[(L - 1) * ENG_SCAN_NSEG + (int)(w / ENG_SCAN_SEG)] + __popcll(m & below);
      if ((unsigned int)pos < (unsigned int)P.cap) P.lay_list[(size_t)(L - 1) * P.cap + pos] = idx;
      else atomicAdd(d_ctr(P, ENG_CNT_ERRORS), 1ull); /* counts of k_step and k_vr_list disagree */
    }
  }
}

/* ------------------------------------------------------------------------ */
/* forced interaction + atom and interaction-type selection */

/* CS_Total_Kissel of element slot j of the layers (xrl_cs_total_c, the same
 * operations on the same values), from its cell record r: the photo sum of the
 * cell's edge count where no edge of the element lies in the cell, else
 * counted from the atom's edges */
__device__ __forceinline__ double d_elem_cs_total(const EngineParams &P, int j, const LayCell &r,
                                                 const xrl_ectx *c) {
  if (!(c->E > 0.0)) return 0.0;
  double pb = r.pb;
  if (!r.edge_free) {
    const xrl_atom *a = &P.atoms[P.el_u[j]];
    pb = a->photo_Bsum[xrl_edges_above(a, c->E)];
  }
  return c->pe * pb + (r.T0 + c->cf * r.dT);
}

/* atom selection and interaction type of the photon in slot i (energy E,
 * now in layer my), from its transport stream after the step's number:
 * S.elem set, the type returned (0 Rayleigh, 1 Compton, 2 photo-absorption) */
__device__ __forceinline__ int d_select(const EngineParams &P, const PhotonSoA &S, int i, double E, int my,
                                        phx_stream *sp) {
  phx_stream &s = *sp;
  const size_t cap = (size_t)P.cap;
  /* atom selection, src/xmi_main.F90:1558-1572 */
  double R2 = phx_uniform(&s);
  int off = P.layer_off[my - 1], nel = P.layer_nel[my - 1];
  double mu_cur = S.mus[(size_t)(my - 1) * cap + i];
  double thr = 0.0;
  int sel = nel - 1;
  xrl_ectx ec = xrl_energy(E);
  {
    /* the next element's record is loaded before this one's test */
    const size_t ci = (size_t)ec.ci;
    LayCell rn = P.el_cell[(size_t)off * XRL_NCS + ci];
    double wn = P.el_w[off];
    for (int k = 0; k < nel; k++) {
      const LayCell r = rn;
      const double w = wn;
      if (k + 1 < nel) {
        rn = P.el_cell[(size_t)(off + k + 1) * XRL_NCS + ci];
        wn = P.el_w[off + k + 1];
      }
      thr += w * d_elem_cs_total(P, off + k, r, &ec) / mu_cur;
      if (R2 < thr) {
        sel = k;
        break;
      }
    }
  }
  S.elem[i] = sel;
  /* interaction type, src/xmi_main.F90:1591-1652 */
  double R3 = phx_uniform(&s);
  int u = P.el_u[off + sel];
  const double *ie = P.ip_E + P.ip_off[u];
  int len = P.ip_off[u + 1] - P.ip_off[u];
  int pos;
  if (fabs(E - ie[0]) < 1e-10) pos = 0;
  else {
    int cell = (int)floor((E - P.ip_acc_E0) * P.ip_acc_inv);
    if (cell < 0) cell = 0;
    if (cell > ENG_IP_ACC - 1) cell = ENG_IP_ACC - 1;
    pos = P.ip_acc[(size_t)u * ENG_IP_ACC + cell];
    while (pos + 2 < len && ie[pos + 1] < E) pos++;
  }
  const double *ir = P.ip_r + P.ip_off[u], *irc = P.ip_rc + P.ip_off[u];
  double pr = d_interp(ie[pos], ir[pos], ie[pos + 1], ir[pos + 1], E);
  double pc = d_interp(ie[pos], irc[pos], ie[pos + 1], irc[pos + 1], E);
  return (R3 < pr) ? 0 : ((R3 < pc) ? 1 : 2);
}

#define VR_ACTIVE 1
#define VR_UP 2
#define VR_NC_SHIFT 8

/* the geometry of the detection-weight estimator for the photon of slot i,
 * at its interaction point coords (lab frame) with direction dirv_l in layer
 * cur (src/xmi_variance_reduction.F90:29-330): a point of the detector
 * window, the scattering angle and polarisation azimuth towards it, the path
 * through the layers and its escape probability, the solid angle from the
 * grid (or the slot queued for a cone estimate: *need_sa).  Writes the
 * photon's estimator record; returns 1 when there is an event (*info set),
 * 0 for a point behind the window plane. */
template <int NL>
__device__ __forceinline__ int d_vr_geom(const EngineParams &P, const PhotonSoA &S, const VrSoA &V, int n, int i,
                                         const double *coords, const double *dirv_l, int cur, int *need_sa,
                                         int *info) {
  const size_t cap = (size_t)P.cap;
  phx_stream s;
  phx_stream_init(&s, P.seed, P.sample, (uint32_t)n, PHX_SUB_VARRED, S.hid[i]);
  double radius = sqrt(phx_uniform(&s)) * P.det_radius;
  double th = 2.0 * ENG_PI * phx_uniform(&s);
  double dp[3] = {0.0, cos(th) * radius, sin(th) * radius};
  double rel[3] = {coords[0] - P.p_det[0], coords[1] - P.p_det[1], coords[2] - P.p_det[2]};
  double lpnt[3], dirv[3], ld[3];
  d_matvec(P.det_inv, rel, lpnt);
  d_matvec(P.det_inv, dirv_l, dirv);
  for (int k = 0; k < 3; k++) ld[k] = dp[k] - lpnt[k];
  if (ld[0] >= 0.0) {
    return 0;
  } else {
    double total_distance = d_distance(dp, lpnt);
    d_normalize(dirv);
    d_normalize(ld);
    double nd[3];
    d_matvec(P.det_new, ld, nd);
    double c = d_dot(dirv, ld);
    c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
    const double theta = acos(c);
    double proj[3], ndd = d_dot(nd, dirv_l);
    for (int k = 0; k < 3; k++) proj[k] = nd[k] - ndd * dirv_l[k];
    d_normalize(proj);
    double ev[3] = {S.ex[i], S.ey[i], S.ez[i]};
    double en = d_norm(ev);
    double evn[3] = {ev[0] / en, ev[1] / en, ev[2] / en};
    c = d_dot(proj, evn);
    c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
    const double phi = acos(c);
    const int up = d_dot(nd, P.n_sample) > 0.0;
    const int dir = up ? 1 : -1;
    const int ncross = up ? P.n_layers - cur + 1 : cur;
    /* path to the detector through the layers, in crossing order; the
     * layer holding the detector point ends the walk */
    double tc[3] = {coords[0], coords[1], coords[2]};
    int done = 0, nc = 0;
    double murhod = 0.0;
#pragma unroll
    for (int k = 0; k < NL; k++) {
      double dd = 0.0;
      if (k < ncross && !done) {
        const int l0 = cur - 1 + k * dir;
        double pp[3] = {0.0, 0.0, up ? P.layer_ze[l0] : P.layer_zb[l0]}, ip[3];
        d_intersect(pp, P.n_sample, tc, nd, ip);
        dd = d_distance(tc, ip);
        if (dd > total_distance) {
          dd = total_distance;
          done = 1;
        } else {
          tc[0] = ip[0];
          tc[1] = ip[1];
          tc[2] = ip[2];
          total_distance -= dd;
        }
        murhod += S.mus[(size_t)l0 * cap + i] * P.layer_density[l0] * dd;
        nc = k + 1;
      }
      V.dist[(size_t)k * cap + i] = dd;
    }
    /* solid angle, src/xmi_solid_angle_f.F90:712-801 */
    double omega = 0.0;
    double r = d_distance(P.p_det, coords);
    double dv[3] = {coords[0] - P.p_det[0], coords[1] - P.p_det[1], coords[2] - P.p_det[2]};
    d_normalize(dv);
    double tt = acos(d_dot(dv, P.n_det));
    if (tt > ENG_PI / 2.0) tt = ENG_PI - tt;
    double sth = ENG_PI / 2.0 - tt;
    if (sth >= P.sa_t[0]) {
      int p1 = d_findpos_uniform(P.sa_r, P.sa_nr, P.sa_r0, P.sa_rinv, r);
      int p2 = d_findpos_uniform(P.sa_t, P.sa_nt, P.sa_t0, P.sa_tinv, sth);
      if (p1 < 0 || p2 < 0) {
        *need_sa = 1;
        V.sar[i] = r;
        V.sat[i] = sth;
      } else {
        const double *a1 = P.sa_r, *a2 = P.sa_t;
        double denom = (a1[p1 + 1] - a1[p1]) * (a2[p2 + 1] - a2[p2]);
        double c1 = (a1[p1 + 1] - r) * (a2[p2 + 1] - sth) / denom;
        double c2 = (r - a1[p1]) * (a2[p2 + 1] - sth) / denom;
        double c3 = (a1[p1 + 1] - r) * (sth - a2[p2]) / denom;
        double c4 = (r - a1[p1]) * (sth - a2[p2]) / denom;
        const double *g0 = P.sa_grid + (size_t)p2 * P.sa_nr + p1;
        const double *g1 = P.sa_grid + (size_t)(p2 + 1) * P.sa_nr + p1;
        omega = c1 * g0[0] + c2 * g0[1] + c3 * g1[0] + c4 * g1[1];
      }
    }
    V.theta[i] = theta;
    V.phi[i] = phi;
    V.omega[i] = omega;
    V.pesc[i] = exp(-murhod);
    V.blk[i] = s.ctr[0]; /* two uniforms drawn: block 1, nothing buffered */
    *info = VR_ACTIVE | (up ? VR_UP : 0) | (nc << VR_NC_SHIFT);
  }
  return 1;
}

/* the layer lists of an order's estimator events: one thread per slot, as
 * the step kernel, whose per-wave counts k_layer_scan has turned into
 * offsets (a photon of layer l with energy above the threshold is listed,
 * with or without an event) */
__global__ void __launch_bounds__(ENG_BLOCK) k_vr_list(EngineParams P, PhotonSoA S, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int ev_layer = (i < count && S.layer[i] > 0 && S.energy[i] > 1.0) ? S.layer[i] : 0;
  d_layer_list(P, ev_layer, i);
}

/* NL bounds the layer count; path lengths and mu*rho are kept in crossing
 * order (k = 0: the photon's layer) so that register arrays are only indexed
 * by compile-time indices */
/* occupancy: six waves per SIMD up to four layers (80 VGPRs, two spilled;
 * measured 2 % faster than five waves); more layers keep their natural budget */
/* EST: the estimator's geometry of the new interaction point follows in the
 * same thread (d_vr_geom), while the photon's record is in registers */
template <int NL, bool EST>
__global__ void __launch_bounds__(ENG_BLOCK) __attribute__((amdgpu_waves_per_eu(NL <= 4 ? 6 : 1)))
k_step(EngineParams P, PhotonSoA S, VrSoA V, int n, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int type = -1, ev_layer = 0, need_sa = 0, info = 0, vr_active = 0;
  const size_t cap = (size_t)P.cap;
  if (i < count && S.layer[i] > 0) {
    double E = S.energy[i];
    if (E < 1.0) {
      S.layer[i] = 0;
    } else {
      phx_stream s;
      phx_stream_init(&s, P.seed, P.sample, (uint32_t)n, PHX_SUB_TRANSPORT, S.hid[i]);
      double R = phx_uniform(&s);
      double dirv[3] = {S.dx[i], S.dy[i], S.dz[i]};
      double coords[3] = {S.x[i], S.y[i], S.z[i]};
      double w = S.weight[i];
      const int cur = S.layer[i];
      const int up = d_dot(dirv, P.n_sample) > 0.0;
      const int dir = up ? 1 : -1;
      const int nc = up ? P.n_layers - cur + 1 : cur; /* layers up to the stack's edge */
      double dist[NL], md[NL];
      double lp[3] = {coords[0], coords[1], coords[2]};
#pragma unroll
      for (int k = 0; k < NL; k++) {
        dist[k] = 0.0;
        md[k] = 0.0;
        if (k < nc) {
          const int l0 = cur - 1 + k * dir;
          double pp[3] = {0.0, 0.0, up ? P.layer_ze[l0] : P.layer_zb[l0]}, ip[3];
          d_intersect(pp, P.n_sample, lp, dirv, ip);
          dist[k] = d_distance(lp, ip);
          lp[0] = ip[0];
          lp[1] = ip[1];
          lp[2] = ip[2];
          md[k] = S.mus[(size_t)l0 * cap + i] * P.layer_density[l0];
        }
      }
      double Pabs = 0.0;
#pragma unroll
      for (int k = 0; k < NL; k++)
        if (k < nc) Pabs += md[k] * dist[k];
      double Pabs2 = -1.0 * expm1(-1.0 * Pabs);
      w *= Pabs2;
      double negln = -1.0 * log1p(-1.0 * R * Pabs2);
      /* layer of the interaction: first where the optical depth exceeds negln */
      int kmy = 0, found = 0;
      double sum = 0.0;
#pragma unroll
      for (int k = 0; k < NL; k++) {
        if (k < nc && !found) {
          sum += md[k] * dist[k];
          if (sum > negln) {
            kmy = k;
            found = 1;
          }
        }
      }
      double mrm = md[0];
#pragma unroll
      for (int k = 1; k < NL; k++)
        if (k == kmy) mrm = md[k];
      double tsum = 0.0;
#pragma unroll
      for (int k = 0; k < NL; k++)
        if (k <= kmy) tsum += (1.0 - (md[k] / mrm)) * dist[k];
      tsum = tsum - 1.0 * log1p(-1.0 * R * Pabs2) / mrm;
      const int my = cur + kmy * dir;
      const double pt[3] = {coords[0] + tsum * dirv[0], coords[1] + tsum * dirv[1], coords[2] + tsum * dirv[2]};
      S.x[i] = pt[0];
      S.y[i] = pt[1];
      S.z[i] = pt[2];
      S.weight[i] = w;
      S.layer[i] = my;
      /* an estimator event of this order */
      ev_layer = E > 1.0 ? my : 0;

      /* atom and type serve the interaction, which the last order of the
       * estimator path does not launch */
      if (!EST || n < P.n_int) type = d_select(P, S, i, E, my, &s);
      if (EST && ev_layer) vr_active = d_vr_geom<NL>(P, S, V, n, i, pt, dirv, my, &need_sa, &info);
    }
  }
  d_enqueue(P, type, i);
  d_layer_count(P, ev_layer);
  if (EST) {
    if (i < count) V.info[i] = info;
    d_append(P, ENG_CNT_Q_SA, need_sa, i);
    d_wave_count(d_ctr(P, ENG_CNT_FALLBACKS), need_sa);
    d_wave_count(d_ctr(P, ENG_CNT_VR), vr_active);
    for (int L = 1; L <= P.n_layers; L++) d_wave_count(d_ctr(P, ENG_CNT_LAYER0 + L - 1), vr_active && ev_layer == L);
  }
}

/* exclusive scan of the per-wave counts in segments of ENG_SCAN_SEG waves:
 * workgroup (segment, layer) scans its segment in place (coalesced, one
 * count per thread) and stores the segment's total; k_layer_base then turns
 * the totals into segment offsets and each layer's total.  nw = waves of the
 * k_step launch, at most ENG_SCAN_SEG * ENG_SCAN_NSEG. */
__global__ void __launch_bounds__(ENG_SCAN_SEG) k_layer_scan(EngineParams P, int nw) {
  __shared__ int wsum[ENG_SCAN_SEG / 64];
  const int seg = blockIdx.x, L0 = blockIdx.y;
  int *row = P.lay_woff + (size_t)L0 * P.n_waves;
  const int k = seg * ENG_SCAN_SEG + threadIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int v = k < nw ? row[k] : 0;
  int incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int base = 0;
  for (int j = 0; j < w; j++) base += wsum[j];
  if (k < nw) row[k] = base + incl - v;
  if (threadIdx.x == ENG_SCAN_SEG - 1) P.lay_segbase[L0 * ENG_SCAN_NSEG + seg] = base + incl;
}

/* segment totals -> segment offsets, and each layer's total (one lane per
 * layer) */
__global__ void k_layer_base(EngineParams P, int nseg) {
  const int L0 = threadIdx.x;
  if (L0 >= P.n_layers) return;
  int run = 0;
  for (int s = 0; s < nseg; s++) {
    const int t = P.lay_segbase[L0 * ENG_SCAN_NSEG + s];
    P.lay_segbase[L0 * ENG_SCAN_NSEG + s] = run;
    run += t;
  }
  P.lay_total[L0] = run;
}

/* ------------------------------------------------------------------------ */
/* detection-weight estimator of every photon at order n
 * (src/xmi_variance_reduction.F90:29-726, xmi_compton_varred2 :949-1101),
 * after its geometry in the step kernel (d_vr_geom: detector point, angles,
 * path through the layers to the detector, escape probability, solid angle
 * from the grid; off-grid points go to a queue), as kernels with their own
 * register and LDS budgets:
 *
 *   k_vr_list     one lane per slot: the per-layer event lists
 *   k_vr_cone     one wavefront per queued photon: the cone estimate of
 *                 src/xmi_solid_angle_f.F90:432-710 with the trials split over
 *                 the lanes (same draws, same hit count as the serial loop)
 *   k_vr_scatter  one lane per photon: Rayleigh and Compton terms of every
 *                 element of the photon's layer; the Compton Doppler draws
 *                 continue the photon's estimator stream in element order
 *   k_vr_lines    a few photons per lane; the wavefront walks the layer's
 *                 (element, fluorescence line) entries with the line data in
 *                 scalar registers and sums each line's terms over its photons
 *
 * Path lengths are kept in crossing order (k = 0: the photon's own layer) so
 * that register arrays have compile-time indices only (NL bounds the layer
 * count).  Tallies are 128-bit fixed point, privatised in LDS when they fit;
 * integer adds commute, so totals do not depend on the order of the adds. */

__device__ __forceinline__ double d_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ void d_tally_wave(unsigned long long *bin, double v, int shift) {
  double s = d_wave_sum(v);
  if (__lane_id() == 0 && s > 0.0) fx_add_global(bin, s, shift);
}

__device__ __forceinline__ double d_readlane(double v, int lane) {
  long long b = __double_as_longlong(v);
  unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

/* read-only tables through the constant address space: wave-uniform reads
 * become scalar loads (the tables are not written while a kernel runs) */
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *d_const(const T *p) {
  return (const __attribute__((address_space(4))) T *)p;
}

/* one DPP lane move of a double (two 32-bit moves); lanes the row mask
 * leaves out, or whose source lane lies outside the row, read 0 */
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double d_dpp0(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo));
}

/* sum of v over the wavefront, in a fixed order (row prefix sums, then the
 * row totals), returned wave-uniform: DPP moves keep the reduction in the
 * VALU, without the LDS round trips of a shuffle butterfly */
__device__ __forceinline__ double d_wave_total(double v) {
  v += d_dpp0<0x111, 0xf>(v); /* row_shr:1 */
  v += d_dpp0<0x112, 0xf>(v); /* row_shr:2 */
  v += d_dpp0<0x114, 0xf>(v); /* row_shr:4 */
  v += d_dpp0<0x118, 0xf>(v); /* row_shr:8: lane 15 of a row holds its sum */
  v += d_dpp0<0x142, 0xa>(v); /* row_bcast:15 into rows 1 and 3 */
  v += d_dpp0<0x143, 0xc>(v); /* row_bcast:31 into rows 2 and 3: lane 63 holds the total */
  return d_readlane(v, 63);
}

/* e^x for x <= 0, table-driven (P. T. P. Tang, "Table-driven implementation
 * of the exponential function in IEEE floating-point arithmetic", ACM TOMS
 * 15(2), 1989): x = (64 m + j) ln2/64 + r with |r| <= ln2/128, so
 * e^x = 2^m * 2^(j/64) * e^r; e^r - 1 is its Taylor polynomial of degree 5
 * (truncation below 4e-17 relative).  tab[j] = 2^(j/64), j = 0..63.
 * Arguments below -746 give 0, as e^x underflows there. */
__device__ __forceinline__ double d_exp_neg(double x, const double *tab) {
  x = fmax(x, -746.0);
  const double n = rint(x * 92.33248261689366);          /* 64 / ln 2 */
  double r = fma(-n, 0.010830424696249145, x);           /* ln2/64, leading part */
  r = fma(-n, 3.623510646634843e-19, r);                 /* ln2/64, trailing part */
  const int ni = (int)n;
  double q = fma(r, 1.0 / 120.0, 1.0 / 24.0);
  q = fma(r, q, 1.0 / 6.0);
  q = fma(r, q, 0.5);
  const double p = fma(r * r, q, r);
  const double T = tab[ni & 63];
  return ldexp(fma(T, p, T), ni >> 6);
}


/* persistent: the off-grid queue's photons (entries flattened over the
 * ENG_SPREAD sub-queues) times ENG_CONE_SEGS(hits) segments of
 * ENG_CONE_SEG trials are the work items; a wavefront takes items w, w + W,
 * ...  Trials 2 b and 2 b + 1 of a photon use block ctr0 + b of its
 * estimator stream (one 32-bit word per uniform), as the serial loop does,
 * and each segment adds its integer hit count and one
 * completion to the photon's packed accumulator (completions in bits 40 and
 * up); the wavefront that completes the photon writes its estimate and clears
 * the accumulator for the next launch.  Integer sums: the estimate equals the
 * serial one whatever the order of the segments. */
#define ENG_CONE_SEG 1024
#define ENG_CONE_DONE_SHIFT 40
__global__ void __launch_bounds__(ENG_BLOCK) k_vr_cone(EngineParams P, PhotonSoA S, VrSoA V, int n) {
  const int lane = __lane_id();
  const unsigned int wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const unsigned int waves = (gridDim.x * blockDim.x) >> 6;
  /* sub-queue sizes, one per lane, and their inclusive prefix sums */
  const unsigned long long sz = P.qcount[(size_t)ENG_CNT_Q_SA * ENG_SPREAD + lane];
  unsigned long long incl = sz;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long v = __shfl_up(incl, off);
    if (lane >= off) incl += v;
  }
  const unsigned long long total = __shfl(incl, 63);
  const long hits = P.hits_per_single;
  const unsigned long long nseg = hits > 0 ? (unsigned long long)((hits + ENG_CONE_SEG - 1) / ENG_CONE_SEG) : 1ull;
  const SaGeom g = d_sa_geom(P);
  for (unsigned long long item = wave; item < total * nseg; item += waves) {
    const unsigned long long e = item / nseg, sg = item - e * nseg;
    /* sub-queue holding entry e: first lane whose inclusive prefix exceeds e */
    const int j = __ffsll((long long)__ballot(incl > e)) - 1;
    const unsigned long long before = __shfl(incl - sz, j);
    const int i = P.queue[((size_t)ENG_CNT_Q_SA * ENG_SPREAD + j) * P.qreg + (e - before)];
    phx_stream s;
    phx_stream_init(&s, P.seed, P.sample, (uint32_t)n, PHX_SUB_VARRED, S.hid[i]);
    const uint32_t ctr0 = V.blk[i];
    SaCone cone;
    const int open = d_sa_setup(g, V.sar[i], V.sat[i], &cone);
    unsigned long long nhit = 0;
    if (open) {
      /* trial pairs (2 b, 2 b + 1) from block ctr0 + b; segments start at
       * even trials (ENG_CONE_SEG is even) */
      const long t1 = min((long)((sg + 1) * ENG_CONE_SEG), hits);
      const uint32_t cs[4] = {ctr0, s.ctr[1], s.ctr[2], s.ctr[3]};
      for (long t = (long)(sg * ENG_CONE_SEG) + 2 * lane; t < t1; t += 128)
        nhit += d_sa_trial_pair(cone, cs, s.key, (uint32_t)(t >> 1), t + 1 < t1);
    }
    for (int off = 32; off > 0; off >>= 1) nhit += __shfl_xor(nhit, off);
    if (lane == 0) {
      unsigned long long *acc = P.cone_acc + e;
      const unsigned long long old = atomicAdd(acc, (1ull << ENG_CONE_DONE_SHIFT) + nhit);
      if ((old >> ENG_CONE_DONE_SHIFT) + 1ull == nseg) {
        const unsigned long long all = (old & ((1ull << ENG_CONE_DONE_SHIFT) - 1ull)) + nhit;
        V.omega[i] = open ? cone.cone_sa * (double)all / (double)hits : 0.0;
        if (open) V.blk[i] = ctr0 + (uint32_t)((hits + 1) / 2);
        *acc = 0ull;
      }
    }
  }
}

/* walk over the layer lists in chunks of 64 entries: chunk c of the
 * concatenation of every layer's chunks (layer 1 first); a wavefront takes
 * chunks in increasing order, so the layer only moves forward (wave-uniform) */
struct LayerChunks {
  int L, tot;
  long long c_begin, c_end;
};

__device__ __forceinline__ bool d_chunk_next(const EngineParams &P, LayerChunks &w, long long c, int csz = 64) {
  while (c >= w.c_end && w.L < P.n_layers) {
    w.L++;
    w.tot = P.lay_total[w.L - 1];
    w.c_begin = w.c_end;
    w.c_end += (w.tot + csz - 1) / csz;
  }
  return c < w.c_end;
}

/* 32-bit words of a lane's Philox stream, two blocks deep: the current
 * block (a, na words left) and the next one (b, when hasb).  Refills are
 * wave-collective: when an active lane is about to need a block it has not
 * computed, every active lane without a next block computes its next one.
 * Lanes empty their blocks at different trials, so a refill per lane would
 * run the Philox rounds at nearly every trial of a wavefront; this way a
 * wavefront runs them about once per four trials.  The words come out in
 * the stream's order (blocks ctr, ctr + 1, ...; each block's words first to
 * last), as phx_word gives them. */
struct WaveWords {
  uint32_t a[4], b[4];
  uint32_t ctr[4], key[2];
  int na, hasb;
};

__device__ __forceinline__ uint32_t d_ww_word(WaveWords &w) {
  if (__ballot(w.na == 0 && !w.hasb)) {
    if (!w.hasb) {
      phx_4x32_10(w.ctr, w.key, w.b);
      w.ctr[0]++;
      w.hasb = 1;
    }
  }
  if (w.na == 0) {
    w.a[0] = w.b[0];
    w.a[1] = w.b[1];
    w.a[2] = w.b[2];
    w.a[3] = w.b[3];
    w.na = 4;
    w.hasb = 0;
  }
  const uint32_t x = w.a[0];
  w.a[0] = w.a[1];
  w.a[1] = w.a[2];
  w.a[2] = w.a[3];
  w.na--;
  return x;
}

/* the Compton profile's CDF at momentum pz from its grid (the
 * xmi_profile_cdf_grid table of the element; gs grid points per unit
 * momentum), clamped to the grid's last point, 0 for pz <= 0; in two parts,
 * so that the grid values of one element can be loaded while the previous
 * element is scored: the two grid values the lookup interpolates between,
 * then the linear interpolation */
__device__ __forceinline__ void d_profile_cdf_load(const double *g, double gs, double pz, double v[2]) {
  v[0] = v[1] = 0.0;
  if (!(pz > 0.0)) return;
  const double x = pz * gs;
  if (x >= (double)(XMI_CP_GRID - 1)) {
    v[0] = v[1] = g[XMI_CP_GRID - 1];
    return;
  }
  const int j = (int)x;
  v[0] = g[j];
  v[1] = g[j + 1];
}

__device__ __forceinline__ double d_profile_cdf_eval(const double v[2], double gs, double pz) {
  if (!(pz > 0.0)) return 0.0;
  const double x = pz * gs;
  if (x >= (double)(XMI_CP_GRID - 1)) return v[0];
  return v[0] + (v[1] - v[0]) * (x - (double)(int)x);
}

/* ---- the advanced Compton model (use_advanced_compton) ------------------
 * src/xmi_main.F90:4785-4983 (transport) and src/xmi_variance_reduction.F90:
 * 752-947 (estimator) with src/xmi_aux_f.F90:1951-2057 (kinematics) */

/* the largest projected momentum an electron of binding energy Ii can carry
 * for scattering by theta (xmi_get_qimax) */
__device__ __forceinline__ double d_qimax(double energy, double Ii, double theta) {
  if (Ii != 0.0 && energy < Ii) return 0.0;
  const double EminIi = energy - Ii, ct = cos(theta);
  const double Q = 137.0 * (EminIi * energy * (1.0 - ct) / ENG_MEC2 - Ii);
  return Q / sqrt(EminIi * EminIi + energy * energy - 2.0 * EminIi * energy * ct);
}

__device__ __forceinline__ double d_q_from_energy(double E0, double E1, double theta) {
  const double ct = cos(theta);
  const double Q = 137.0 * (E1 - E0 + (1.0 - ct) * E0 * E1 / ENG_MEC2);
  return Q / sqrt(E1 * E1 + E0 * E0 - 2.0 * E0 * E1 * ct);
}

/* xmi_get_energy_from_q: the root of the quadratic whose momentum has Q's
 * sign; 0 in the reference's exceptions, -1 when no root has it (the
 * reference stops there) */
__device__ double d_energy_from_q(double E0, double Q, double theta) {
  const double c = cos(theta);
  if (fabs(c - 1.0) < 1e-8) return 0.0;
  if (fabs(Q) < 1e-4) return E0 / (1.0 + E0 * (1.0 - c) / ENG_MEC2);
  const double d = 1.0 + E0 / ENG_MEC2 - E0 * c / ENG_MEC2;
  const double aq = 137.0 * 137.0 * d * d - Q * Q;
  const double bq = -2.0 * 137.0 * 137.0 * E0 * d + 2.0 * E0 * c * Q * Q;
  const double cq = 137.0 * 137.0 * E0 * E0 - E0 * E0 * Q * Q;
  double E1, E2;
  if (aq == 0.0) {
    if (bq == 0.0) return 0.0;
    E1 = E2 = -1.0 * cq / bq;
  } else {
    const double delta = bq * bq - 4.0 * aq * cq;
    if (delta < 0.0) return 0.0;
    if (delta == 0.0) {
      E1 = E2 = -bq / 2.0 / aq;
    } else {
      const double sd = sqrt(delta);
      const double t1 = (-bq + sd) / 2.0 / aq, t2 = (-bq - sd) / 2.0 / aq;
      E1 = t1 < t2 ? t1 : t2;
      E2 = t1 < t2 ? t2 : t1;
    }
  }
  const double Q1 = d_q_from_energy(E0, E1, theta), Q2 = d_q_from_energy(E0, E2, theta);
  if (Q * Q1 > 0.0) return E1;
  if (Q * Q2 > 0.0) return E2;
  if (fabs(E1 - E2) < 1e-10 || fabs(Q1 - Q2) < 1e-10) return E1;
  return -1.0;
}

/* subshell g's two-sided partial-profile CDF at its Qimax */
__device__ __forceinline__ double d_adv_cdf(const EngineParams &P, const xrl_atom *a, int g, double energy,
                                            double theta) {
  const double q = d_qimax(energy, a->edge[P.cp_sh_idx[g]], theta);
  if (q < -100.0) return 0.0;
  if (q > 100.0) return 1.0;
  const double qp = fabs(q);
  const double *Qs = P.cp_Qs, *pc = P.cp_pcdf + (size_t)g * P.cp_len;
  int pos = (int)(qp / (Qs[1] - Qs[0]));
  if (pos > P.cp_len - 2) pos = P.cp_len - 2;
  const double v = d_interp(Qs[pos], pc[pos], Qs[pos + 1], pc[pos + 1], qp);
  return q < 0.0 ? 1.0 - (0.5 + v) : 0.5 + v;
}

/* the projected momentum at a value cdf of subshell g's two-sided profile */
__device__ __forceinline__ double d_adv_q(const EngineParams &P, int g, double cdf) {
  const double *inv = P.cp_inv, *qi = P.cp_qinv + (size_t)g * P.cp_len;
  const double cp = cdf < 0.5 ? 0.5 - cdf : cdf - 0.5;
  int pos = (int)(cp / (inv[1] - inv[0]));
  if (pos > P.cp_len - 2) pos = P.cp_len - 2;
  const double Q = d_interp(inv[pos], qi[pos], inv[pos + 1], qi[pos + 1], cp);
  return cdf < 0.5 ? -Q : Q;
}

template <int NL, bool use_lds, bool ADV>
__global__ void __launch_bounds__(ENG_BLOCK) __attribute__((amdgpu_waves_per_eu(3))) k_vr_scatter(EngineParams P, PhotonSoA S, VrSoA V, int n) {
  /* LDS tallies (when they fit): channels [nch][2], Rayleigh/Compton rows
   * [nZ][2][2]; flushed once per workgroup */
  extern __shared__ unsigned long long lds[];
  __shared__ double exp_tab[64];
  const size_t n_ch2 = (size_t)P.nch * 2, n_rc2 = (size_t)P.nZ * 4;
  const size_t cap = (size_t)P.cap;
  unsigned long long *gch = P.t_ch + (size_t)(n - 1) * n_ch2;
  unsigned long long *grc = P.t_rc + (size_t)(n - 1) * n_rc2;
  unsigned long long *tch = gch, *trc = grc;
  if (threadIdx.x < 64) exp_tab[threadIdx.x] = exp2((double)threadIdx.x / 64.0);
  if (use_lds) {
    tch = lds;
    trc = lds + n_ch2;
    for (size_t k = threadIdx.x; k < n_ch2 + n_rc2; k += blockDim.x) lds[k] = 0ull;
  }
  __syncthreads();
  const int shift = P.tally_shift;
  const int lane = __lane_id();
  const double inv_gain = 1.0 / P.gain;
  /* the Compton profile's random-number grid is uniform */
  const double inv_drn = 1.0 / (P.cp_rn[1] - P.cp_rn[0]);
  /* single-precision literals of src/xmi_variance_reduction.F90:949-1008 */
  const double cc = (double)1.2399e-6f, c0 = (double)4.85e-12f, c1 = (double)1.456e-2f;
  /* wave index through readfirstlane: wave-uniform for the compiler too,
   * so that the chunk walk and the layer's data stay in scalar registers */
  const long long wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6),
                  waves = (gridDim.x * blockDim.x) >> 6;
  LayerChunks lc = {0, 0, 0, 0};
  for (long long c = wave; d_chunk_next(P, lc, c); c += waves) {
    const int L = lc.L;
    const int e = (int)((c - lc.c_begin) * 64) + lane;
    const int i = e < lc.tot ? P.lay_list[(size_t)(L - 1) * cap + e] : 0;
    const int info = e < lc.tot ? V.info[i] : 0;
    const int mine = info & VR_ACTIVE;
    /* per-photon factors, shared by every element of the layer */
    double E = 0.0, wo = 0.0, pesc = 0.0, inv_mu = 0.0, c_lamb0 = 1.0, A0 = 0.0, A1 = 0.0, th_e = 0.0;
    phx_stream sd; /* advanced model: uniforms from the same stream position */
    /* rho * path in the k-th layer crossed: in registers for up to four
     * layers, read from the records beyond (register budget) */
    constexpr int NR = NL <= 4 ? NL : 1;
    double rd[NR];
    xrl_dcsp_ctx dc = {0.0, 0.0, 0.0};
    int dir = 1, nc = 0, ch_r = -1;
    WaveWords s = {};
#pragma unroll
    for (int k = 0; k < NR; k++) rd[k] = 0.0;
    if (mine) {
      E = S.energy[i];
      const double theta = V.theta[i], phi = V.phi[i];
      wo = V.omega[i] * S.weight[i];
      pesc = V.pesc[i];
      inv_mu = 1.0 / S.mus[(size_t)(L - 1) * cap + i];
      dir = (info & VR_UP) ? 1 : -1;
      nc = info >> VR_NC_SHIFT;
      if (NL <= 4) {
#pragma unroll
        for (int k = 0; k < NR; k++)
          if (k < nc) rd[k] = P.layer_density[L - 1 + k * dir] * V.dist[(size_t)k * cap + i];
      }
      /* the estimator stream from the block after the geometry and cone
       * draws, as words */
      {
        phx_stream s0;
        phx_stream_init(&s0, P.seed, P.sample, (uint32_t)n, PHX_SUB_VARRED, S.hid[i]);
        s.key[0] = s0.key[0];
        s.key[1] = s0.key[1];
        s.ctr[0] = V.blk[i];
        s.ctr[1] = s0.ctr[1];
        s.ctr[2] = s0.ctr[2];
        s.ctr[3] = s0.ctr[3];
        s.na = 0;
        s.hasb = 0;
        if (ADV) {
          sd = s0;
          sd.ctr[0] = V.blk[i];
        }
      }
      th_e = theta;
      dc = xrl_dcsp_setup(E, theta, phi);
      const double sth2 = sin(theta / 2.0);
      c_lamb0 = cc / (E * 1000.0);
      A0 = c0 * sth2 * sth2;
      A1 = c1 * c_lamb0 * sth2;
      ch_r = E >= 1.0 ? (int)((E - P.zero) * inv_gain) : -1;
    }
    const int off = __builtin_amdgcn_readfirstlane(P.layer_off[L - 1]);
    const int nel = __builtin_amdgcn_readfirstlane(P.layer_nel[L - 1]);
    double twr_sum = 0.0; /* Rayleigh keeps the energy: one channel per photon */
    double rsum = 0.0;    /* lane 2j + kind: row total of element k0 + j of this group */
    /* the accepted momenta's bounds of the Doppler draw (per photon), and
     * the profile-CDF values at them of the next element, loaded one
     * element ahead */
    const double pz_max = A1 > 0.0 ? A0 / A1 : 0.0, pz_neg = A1 > 0.0 ? (c_lamb0 + A0) / A1 : 0.0;
    double cdf_max[2] = {0.0, 0.0}, cdf_neg[2] = {0.0, 0.0};
    if (!ADV && mine && A1 > 0.0 && nel > 0) {
      const int u0 = d_const(P.el_u)[off];
      d_profile_cdf_load(P.cp_cdfg + (size_t)u0 * XMI_CP_GRID, P.cp_gscale[u0], pz_max, cdf_max);
      d_profile_cdf_load(P.cp_cdfg + (size_t)u0 * XMI_CP_GRID, P.cp_gscale[u0], pz_neg, cdf_neg);
    }
    for (int k = 0; k < nel; k++) {
      const int u = d_const(P.el_u)[off + k];
      const xrl_atom *a = &P.atoms[u];
      const double Pconv = d_const(P.el_w)[off + k] * inv_mu;
      double f_max = 0.0, f_neg = 0.0;
      if (!ADV && mine && A1 > 0.0) {
        const double gs = P.cp_gscale[u];
        f_max = d_profile_cdf_eval(cdf_max, gs, pz_max);
        f_neg = d_profile_cdf_eval(cdf_neg, gs, pz_neg);
        if (k + 1 < nel) {
          const int u1 = d_const(P.el_u)[off + k + 1];
          d_profile_cdf_load(P.cp_cdfg + (size_t)u1 * XMI_CP_GRID, P.cp_gscale[u1], pz_max, cdf_max);
          d_profile_cdf_load(P.cp_cdfg + (size_t)u1 * XMI_CP_GRID, P.cp_gscale[u1], pz_neg, cdf_neg);
        }
      }
      /* Rayleigh */
      double twr = 0.0;
      if (mine) {
        twr = Pconv * (xrl_dcsp_rayl_c(a, &dc) * wo) * pesc;
        twr_sum += twr;
      }
      /* Compton, xmi_compton_varred2: Doppler-broadened energy from the
       * element's Compton profile, continuing the photon's estimator stream */
      double twc = 0.0;
      if (ADV && mine) {
        /* every occupied subshell with weight count * CDF(Qimax) / sum, one
         * uniform each (xmi_compton_varred) */
        const int g0 = P.cp_sh_off[u], g1 = P.cp_sh_off[u + 1];
        double csum = 0.0;
        for (int g = g0; g < g1; g++) csum += a->occupancy[P.cp_sh_idx[g]] * d_adv_cdf(P, a, g, E, th_e);
        const double dcw = Pconv * (xrl_dcsp_compt_c(a, &dc) * wo);
        for (int g = g0; csum != 0.0 && g < g1; g++) {
          const double cdf = d_adv_cdf(P, a, g, E, th_e);
          const double sw = a->occupancy[P.cp_sh_idx[g]] * cdf / csum;
          if (sw == 0.0) continue;
          const double energy = d_energy_from_q(E, d_adv_q(P, g, phx_uniform(&sd) * cdf), th_e);
          if (!(energy > 0.0)) continue;
          const xrl_ectx ce = xrl_energy(energy);
          double m2 = 0.0;
#pragma unroll 1
          for (int kk = 0; kk < nc; kk++) {
            const int l0 = L - 1 + kk * dir;
            m2 += d_layer_mu_tab(P, l0, &ce) * (P.layer_density[l0] * V.dist[(size_t)kk * cap + i]);
          }
          const double tw = dcw * d_exp_neg(-m2, exp_tab) * sw;
          twc += tw;
          const int ch = energy >= 1.0 ? (int)((energy - P.zero) * inv_gain) : -1;
          if (ch >= 0 && ch <= P.nch - 1 && tw > 0.0) fx_add_global(&tch[2 * ch], tw, shift);
        }
      }
      if (!ADV && mine) {
        /* xmi_compton_varred2 (src/xmi_variance_reduction.F90:949-1008)
         * redraws the momentum until E' <= E; the accepted momenta are pz <=
         * A0/A1 and, where E' turns negative, pz > (c_lamb0 + A0)/A1.  They
         * are drawn here directly from that part of the profile, one word
         * each (the same distribution, and no wavefront waits on its lanes'
         * longest rejection run); r ranges over [0, r_end), the inverse
         * table without its last interval */
        const double *rn = P.cp_rn;
        const double *icdf = P.cp_icdf + (size_t)u * P.cp_len;
        const double r_end = rn[P.cp_len - 2];
        if (!(A1 > 0.0)) f_max = f_neg = r_end; /* the profile CDFs at A0 / A1 were loaded ahead */
        const double m_pos = f_max + (r_end - f_neg);
        const uint32_t wd = d_ww_word(s);
        const double uu = (double)wd * 0x1.0p-32 * (r_end + m_pos);
        double r, sgn;
        if (uu < r_end) {
          r = uu;
          sgn = -1.0;
        } else {
          const double v = uu - r_end;
          r = v < f_max ? v : v - f_max + f_neg;
          sgn = 1.0;
        }
        int pos = (int)(r * inv_drn);
        if (pos > P.cp_len - 3) pos = P.cp_len - 3;
        const double pz = sgn * (icdf[pos] + (icdf[pos + 1] - icdf[pos]) * ((r - rn[pos]) * inv_drn));
        const double energy = (cc / 1000.0) / (c_lamb0 + (A0 - A1 * pz));
        {
          const xrl_ectx ce = xrl_energy(energy);
          double m2 = 0.0;
          if (NL <= 4) {
#pragma unroll
            for (int kk = 0; kk < NR; kk++)
              if (kk < nc) m2 += d_layer_mu_tab(P, L - 1 + kk * dir, &ce) * rd[kk];
          } else {
#pragma unroll 1
            for (int kk = 0; kk < nc; kk++) {
              const int l0 = L - 1 + kk * dir;
              m2 += d_layer_mu_tab(P, l0, &ce) * (P.layer_density[l0] * V.dist[(size_t)kk * cap + i]);
            }
          }
          twc = Pconv * (xrl_dcsp_compt_c(a, &dc) * wo) * d_exp_neg(-m2, exp_tab);
          const int ch = energy >= 1.0 ? (int)((energy - P.zero) * inv_gain) : -1;
          if (ch >= 0 && ch <= P.nch - 1 && twc > 0.0) fx_add_global(&tch[2 * ch], twc, shift);
        }
      }
      /* history rows: wave totals collected in lanes, 32 elements a group */
      const double tr = d_wave_total(twr), tc = d_wave_total(twc);
      const int j = k & 31;
      if (lane == 2 * j) rsum = tr;
      if (lane == 2 * j + 1) rsum = tc;
      if (j == 31 || k == nel - 1) {
        const int kl = k - j + (lane >> 1);
        if (kl <= k && rsum > 0.0) {
          const int ul = d_const(P.el_u)[off + kl];
          fx_add_global(&trc[((size_t)ul * 2 + (lane & 1)) * 2], rsum, shift);
        }
        rsum = 0.0;
      }
    }
    if (ch_r >= 0 && ch_r <= P.nch - 1 && twr_sum > 0.0) fx_add_global(&tch[2 * ch_r], twr_sum, shift);
  }
  if (use_lds) {
    __syncthreads();
    for (size_t k = threadIdx.x; k < n_ch2 / 2; k += blockDim.x)
      if (tch[2 * k] | tch[2 * k + 1]) fx_merge_global(&gch[2 * k], tch[2 * k], tch[2 * k + 1]);
    for (size_t k = threadIdx.x; k < n_rc2 / 2; k += blockDim.x)
      if (trc[2 * k] | trc[2 * k + 1]) fx_merge_global(&grc[2 * k], trc[2 * k], trc[2 * k + 1]);
  }
}

/* lines phase: KP photons per lane, a chunk of 64 KP photons (all in layer
 * L); the wavefront walks the layer's (element, line) entries.  Per (line t,
 * photon) term
 *   wf_t * P_shell(t, photon) * g * exp(-sum_l mu_rho(t, l) * d_l)
 * with g = Omega / 4 pi * weight / mu(E) and d_l the photon's path in layer l
 * on its way to the detector.  The line's data are wave-uniform (scalar
 * loads); P_shell is formed once per (element, shell) of the walk: a first
 * interaction (or a photon not coming from fluorescence) uses
 * E^-p * shell_C[cascade][edges above E][shell], a fluorescence photon the
 * precalculated cascade table (src/xmi_variance_reduction.F90:540-580).  The
 * terms of a line are summed over each lane's photons, then over the lanes in
 * a fixed order (deterministic), and added once to the line's fixed-point
 * bin: one wavefront sum per line and 64 KP photons.  The lane sums go
 * through LDS eight lines at a time: the lanes store their terms of eight
 * lines, then eight lanes per line add its 64 terms (eight each, then three
 * DPP steps), instead of a six-step DPP reduction per line. */
#define ENG_LRED_STRIDE 72 /* doubles per line row of the staging: 8 mod 32, conflict-free reads */
template <int NL, int KP, bool use_lds>
__global__ void __launch_bounds__(ENG_BLOCK) k_vr_lines(EngineParams P, PhotonSoA S, VrSoA V, int n) {
  /* LDS tallies (when they fit): compact line bins [n_lbins][2] */
  extern __shared__ unsigned long long lds[];
  __shared__ double exp_tab[64];
  /* per wavefront: the lanes' terms of 8 lines, row j = line j of the group */
  __shared__ double lred[ENG_BLOCK / 64][8 * ENG_LRED_STRIDE];
  const size_t n_line2 = (size_t)P.n_lbins * 2;
  const size_t cap = (size_t)P.cap;
  unsigned long long *gline = P.t_line + (size_t)(n - 1) * P.nZ * ENG_MAX_ACT * 2;
  unsigned long long *tline = gline;
  if (threadIdx.x < 64) exp_tab[threadIdx.x] = exp2((double)threadIdx.x / 64.0);
  if (use_lds) {
    tline = lds;
    for (size_t k = threadIdx.x; k < n_line2; k += blockDim.x) lds[k] = 0ull;
  }
  __syncthreads();
  const int shift = P.tally_shift;
  const int smax = P.use_M_lines ? 8 : 3;
  /* wave index through readfirstlane: wave-uniform for the compiler too,
   * so that the chunk walk and the layer's data stay in scalar registers */
  const long long wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6),
                  waves = (gridDim.x * blockDim.x) >> 6;
  LayerChunks lc = {0, 0, 0, 0};
  for (long long c = wave; d_chunk_next(P, lc, c, 64 * KP); c += waves) {
    const int L = lc.L;
    /* the lane's KP photons: list entries e0, e0 + 64, ... of layer L */
    const int e0 = (int)((c - lc.c_begin) * (64 * KP)) + __lane_id();
    double E[KP], pe[KP], g[KP], x[KP];
    /* path length per layer, by layer index (0 where not crossed) */
    double dl[KP][NL];
    int act[KP], fo[KP], pl[KP], pu[KP], kE[KP];
    int any_act = 0;
#pragma unroll
    for (int k = 0; k < KP; k++) {
      const int e = e0 + 64 * k;
      const int i = e < lc.tot ? P.lay_list[(size_t)(L - 1) * cap + e] : 0;
      const int info = e < lc.tot ? V.info[i] : 0;
      act[k] = info & VR_ACTIVE;
      any_act |= act[k];
      E[k] = 2.0;
      pe[k] = g[k] = x[k] = 0.0;
      fo[k] = 1;
      pl[k] = pu[k] = kE[k] = 0;
#pragma unroll
      for (int l = 0; l < NL; l++) dl[k][l] = 0.0;
      if (act[k]) {
        E[k] = S.energy[i];
        pe[k] = xrl_energy(E[k]).pe;
        g[k] = V.omega[i] / 4.0 / ENG_PI * S.weight[i] / S.mus[(size_t)(L - 1) * cap + i];
        pl[k] = S.prev_line[i];
        pu[k] = S.prev_u[i];
        fo[k] = !(n > 1 && pl[k] > 0);
        const int nc = info >> VR_NC_SHIFT, dir = (info & VR_UP) ? 1 : -1;
#pragma unroll
        for (int q = 0; q < NL; q++) {
          const double d = q < nc ? V.dist[(size_t)q * cap + i] : 0.0;
          const int l0 = L - 1 + q * dir;
#pragma unroll
          for (int l = 0; l < NL; l++)
            if (q < nc && l == l0) dl[k][l] = d;
        }
      }
    }
    if (__ballot(any_act) == 0ull) continue;
    const int t0 = __builtin_amdgcn_readfirstlane(P.fl_off[L - 1]), t1 = __builtin_amdgcn_readfirstlane(P.fl_off[L]);
    int u_cur = -1, z_cur = -1, any = 0;
    double *wred = lred[threadIdx.x >> 6];
    const int lane = __lane_id();
    const auto *fl_info = d_const((const int *)P.fl_info);
    const auto *fl_wf = d_const(P.fl_wf);
    const auto *fl_mr = d_const(P.fl_mr);
    /* the next entry's scalars are loaded one iteration ahead (the tables
     * end with a zero entry), so that their latency overlaps this entry's
     * arithmetic */
    int nx_u = fl_info[4 * t0], nx_z = fl_info[4 * t0 + 2];
    double nx_wf = fl_wf[t0], nx_mr[NL];
#pragma unroll
    for (int l = 0; l < NL; l++) nx_mr[l] = fl_mr[(size_t)t0 * NL + l];
    for (int t = t0; t < t1; t++) {
      /* {element, active line, shell, compact bin} */
      const int2 it = make_int2(nx_u, nx_z);
      const double wf = nx_wf;
      double mr[NL];
#pragma unroll
      for (int l = 0; l < NL; l++) mr[l] = nx_mr[l];
      nx_u = fl_info[4 * (t + 1)];
      nx_z = fl_info[4 * (t + 1) + 2];
      nx_wf = fl_wf[t + 1];
#pragma unroll
      for (int l = 0; l < NL; l++) nx_mr[l] = fl_mr[(size_t)(t + 1) * NL + l];
      if (it.x != u_cur) {
        /* edges of the element above each photon's energy */
        const auto *a = d_const(&P.atoms[it.x]);
#pragma unroll
        for (int k = 0; k < KP; k++) {
          kE[k] = 0;
#pragma unroll
          for (int q = 0; q < 9; q++) kE[k] += a->edge[q] > E[k] ? 1 : 0;
        }
        u_cur = it.x;
        z_cur = -1;
      }
      if (it.y != z_cur) {
        z_cur = it.y;
        int nz = 0;
#pragma unroll
        for (int k = 0; k < KP; k++) {
          double v;
          if (fo[k]) {
            v = (!P.use_M_lines && z_cur >= 4) ? 0.0 : pe[k] * P.atoms[u_cur].shell_C[P.cascade - 1][kE[k]][z_cur];
          } else {
            v = z_cur <= smax ? P.precalc_xrf[(((size_t)u_cur * 9 + z_cur) * P.nZ + pu[k]) * XMI_NLINES_VR_DEV +
                                              (pl[k] - 1)]
                              : 0.0;
          }
          /* shell factor times the photon's weight factor; 0 for lanes
           * without a photon, so that their terms vanish */
          x[k] = act[k] ? v * g[k] : 0.0;
          nz |= x[k] != 0.0;
        }
        any = __ballot(nz) != 0ull;
      }
      const int j = (t - t0) & 7;
      double part = 0.0;
      if (any) {
        /* attenuation rows are padded to NL entries (zeros past the stack);
         * the lane's photons are summed before the wavefront sum */
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < KP; k++) {
          double m2 = 0.0;
#pragma unroll
          for (int l = 0; l < NL; l++) m2 += mr[l] * dl[k][l];
          acc = fma(x[k], d_exp_neg(-m2, exp_tab), acc);
        }
        part = acc * wf;
      }
      wred[j * ENG_LRED_STRIDE + lane] = part;
      if (j == 7 || t == t1 - 1) {
        /* line m of the group (m <= j): lane 8m + q adds the terms of lanes
         * 8r + q (r = 0..7), then the eight lanes of m are summed (xor 1,
         * xor 2 in quads, then the half-row mirror joins the two quads); the
         * staging of the next group is written after these reads (one
         * wavefront's LDS accesses complete in order) */
        __builtin_amdgcn_wave_barrier();
        const int m = lane >> 3, q = lane & 7;
        double sum = 0.0;
#pragma unroll
        for (int r = 0; r < 8; r++) sum += wred[m * ENG_LRED_STRIDE + 8 * r + q];
        sum += d_dpp0<0xB1, 0xf>(sum);  /* quad_perm [1, 0, 3, 2] */
        sum += d_dpp0<0x4E, 0xf>(sum);  /* quad_perm [2, 3, 0, 1] */
        sum += d_dpp0<0x141, 0xf>(sum); /* row_half_mirror */
        __builtin_amdgcn_wave_barrier();
        const int tl = t - j + m;
        if (q == 0 && tl <= t && sum > 0.0) {
          const int4 il = P.fl_info[tl];
          unsigned long long *bin =
              use_lds ? &tline[(size_t)il.w * 2] : &tline[((size_t)il.x * ENG_MAX_ACT + il.y) * 2];
          fx_add_global(bin, sum, shift);
        }
      }
    }
  }
  if (use_lds) {
    __syncthreads();
    for (size_t k = threadIdx.x; k < n_line2 / 2; k += blockDim.x)
      if (tline[2 * k] | tline[2 * k + 1])
        fx_merge_global(&gline[(size_t)P.lbin_g[k] * 2], tline[2 * k], tline[2 * k + 1]);
  }
}

/* ------------------------------------------------------------------------ */
/* interactions */

/* transport substream of slot i: a cascade's offspring photon has its own */
__device__ __forceinline__ uint32_t d_tr_sub(const EngineParams &P, int i) {
  return (P.off_half && i >= P.off_half) ? PHX_SUB_OFFSPRING : PHX_SUB_TRANSPORT;
}

/* the interaction routines are compiled twice: BRUTE (brute-force runs)
 * adds the history rows of Rayleigh and Compton, the analogue yield and the
 * cascades, which the variance-reduction path does without */

template <bool BRUTE>
__device__ __forceinline__ void d_rayleigh_one(const EngineParams &P, const PhotonSoA &S, int n, int i) {
  const size_t cap = (size_t)P.cap;
  int L = S.layer[i];
  int u = P.el_u[P.layer_off[L - 1] + S.elem[i]];
  phx_stream s;
  phx_stream_init(&s, P.seed, P.sample, (uint32_t)n, BRUTE ? d_tr_sub(P, i) : PHX_SUB_TRANSPORT, S.hid[i]);
  d_stream_seek(&s, 3u);
  double E = S.energy[i];
  double r = phx_uniform(&s);
  double theta_i = d_bilinear(P.ray_icdf + (size_t)u * P.n_e * P.n_r, P.tab_E, P.n_e, P.tab_E0, P.tab_Einv, P.tab_R,
                              P.n_r, P.tab_R0, P.tab_Rinv, E, r);
  double tt = sin(theta_i);
  tt = tt * tt;
  tt = tt / (4.0 - 2.0 * tt);
  double r2 = phx_uniform(&s);
  double phi_i = d_bilinear(P.phi_icdf, P.phi_T, P.phi_nt, P.phi_T0, P.phi_Tinv, P.phi_R, P.phi_nr, P.phi_R0,
                            P.phi_Rinv, tt, r2);
  double th = S.theta[i], ph = S.phi[i];
  double ev[3] = {S.ex[i], S.ey[i], S.ez[i]};
  double phi0 = d_phi0(th, ph, ev);
  double dirv[3];
  d_update_dirv(dirv, &th, &ph, theta_i, phi0 + phi_i);
  d_update_elecv(dirv, ev);
  S.dx[i] = dirv[0];
  S.dy[i] = dirv[1];
  S.dz[i] = dirv[2];
  S.ex[i] = ev[0];
  S.ey[i] = ev[1];
  S.ez[i] = ev[2];
  S.theta[i] = th;
  S.phi[i] = ph;
  /* brute force: the history row of the last interaction, and its element */
  S.prev_line[i] = BRUTE ? 384 : 0;
  if (BRUTE) S.prev_u[i] = u;
  (void)cap;
}

template <bool BRUTE, bool ADV>
__device__ __forceinline__ void d_compton_one(const EngineParams &P, const PhotonSoA &S, int n, int i) {
  const size_t cap = (size_t)P.cap;
  int L = S.layer[i];
  int u = P.el_u[P.layer_off[L - 1] + S.elem[i]];
  phx_stream s;
  phx_stream_init(&s, P.seed, P.sample, (uint32_t)n, BRUTE ? d_tr_sub(P, i) : PHX_SUB_TRANSPORT, S.hid[i]);
  d_stream_seek(&s, 3u);
  double E = S.energy[i];
  double theta_i = d_bilinear(P.comp_icdf + (size_t)u * P.n_e * P.n_r, P.tab_E, P.n_e, P.tab_E0, P.tab_Einv, P.tab_R,
                              P.n_r, P.tab_R0, P.tab_Rinv, E, phx_uniform(&s));
  double K0K = 1.0 + E * (1.0 - cos(theta_i)) / ENG_MEC2;
  double tt = sin(theta_i);
  tt = tt * tt;
  tt = tt / (K0K + (1.0 / K0K) - tt) / 2.0;
  double phi_i = d_bilinear(P.phi_icdf, P.phi_T, P.phi_nt, P.phi_T0, P.phi_Tinv, P.phi_R, P.phi_nr, P.phi_R0,
                            P.phi_Rinv, tt, phx_uniform(&s));
  double th = S.theta[i], ph = S.phi[i];
  double ev[3] = {S.ex[i], S.ey[i], S.ez[i]};
  double phi0 = d_phi0(th, ph, ev);
  if (ADV) {
    /* xmi_update_photon_energy_compton: the subshell by count * CDF(Qimax),
     * Q from its profile below Qimax, the energy from Q; 0 ends the photon */
    const xrl_atom *a = &P.atoms[u];
    const int g0 = P.cp_sh_off[u], g1 = P.cp_sh_off[u + 1];
    double csum = 0.0;
    for (int g = g0; g < g1; g++) csum += a->occupancy[P.cp_sh_idx[g]] * d_adv_cdf(P, a, g, E, theta_i);
    if (csum == 0.0) {
      S.energy[i] = 0.0;
      return;
    }
    const double r = phx_uniform(&s);
    double ts = 0.0, cdf_sel = 0.0;
    int sel = -1, last = g0;
    for (int g = g0; g < g1; g++) {
      const double w = a->occupancy[P.cp_sh_idx[g]] * d_adv_cdf(P, a, g, E, theta_i);
      if (w > 0.0) last = g;
      ts += w / csum;
      if (r <= ts) {
        sel = g;
        break;
      }
    }
    if (sel < 0) sel = last; /* rounding left the running sum just below r */
    cdf_sel = d_adv_cdf(P, a, sel, E, theta_i);
    const double e = d_energy_from_q(E, d_adv_q(P, sel, phx_uniform(&s) * cdf_sel), theta_i);
    if (!(e > 0.0)) {
      S.energy[i] = 0.0;
      return;
    }
    E = e;
    S.energy[i] = E;
    d_mu_all(P, E, S.mus, cap, i);
  } else {
  /* xmi_update_photon_energy_compton2 */
  const double cc = (double)1.2399e-6f, c0 = (double)4.85e-12f, c1 = (double)1.456e-2f;
  double energy = E * 1000.0;
  double c_lamb0 = cc / energy;
  double sth2 = sin(theta_i / 2.0);
  const double *rn = P.cp_rn;
  const double *icdf = P.cp_icdf + (size_t)u * P.cp_len;
  const double drn = rn[1] - rn[0];
  /* the reference loops until acceptance; every draw with pz <= 0 is accepted,
   * so the bound below is never reached by valid data and only turns a
   * corrupted table into a counted error instead of a hung wave */
  int accepted = 0;
  for (int it = 0; it < 4096; it++) {
    double r = phx_uniform(&s);
    int pos = (int)(r / drn);
    double pz = d_interp(rn[pos], icdf[pos], rn[pos + 1], icdf[pos + 1], r);
    if (phx_uniform(&s) < 0.5) pz = -pz;
    double dlamb = c0 * sth2 * sth2 - c1 * c_lamb0 * sth2 * pz;
    energy = cc / (c_lamb0 + dlamb) / 1000.0;
    if (energy <= E) {
      accepted = 1;
      break;
    }
  }
  if (!accepted) {
    atomicAdd(d_ctr(P, ENG_CNT_ERRORS), 1ull);
    atomicAdd(d_ctr(P, ENG_CNT_ERR_DOPPLER), 1ull);
    S.energy[i] = 0.0;
    return;
  }
  E = energy;
  S.energy[i] = E;
  d_mu_all(P, E, S.mus, cap, i);
  if (E == 0.0) return;
  }
  double dirv[3];
  d_update_dirv(dirv, &th, &ph, theta_i, phi_i + phi0);
  d_update_elecv(dirv, ev);
  double ct = cos(theta_i), cph = cos(phi_i), sph = sin(phi_i);
  double pp = 2.0 * ((ct * cph) * (ct * cph) + sph * sph);
  double rat = 1.0 / (1.0 + (1.0 - ct) * E / (double)510.998910f);
  double rk = rat - 2.0 + 1.0 / rat;
  pp = pp / (rk + pp);
  double r = phx_uniform(&s);
  double w_h = (1.0 + pp) / 2.0;
  if (r > w_h) {
    double c0x = dirv[1] * ev[2] - dirv[2] * ev[1];
    double c1x = dirv[2] * ev[0] - dirv[0] * ev[2];
    double c2x = dirv[0] * ev[1] - dirv[1] * ev[0];
    ev[0] = c0x;
    ev[1] = c1x;
    ev[2] = c2x;
  }
  S.dx[i] = dirv[0];
  S.dy[i] = dirv[1];
  S.dz[i] = dirv[2];
  S.ex[i] = ev[0];
  S.ey[i] = ev[1];
  S.ez[i] = ev[2];
  S.theta[i] = th;
  S.phi[i] = ph;
  S.prev_line[i] = BRUTE ? 385 : 0;
  if (BRUTE) S.prev_u[i] = u;
}

/* Coster-Kronig vacancy shift, src/xmi_main.F90:5184-5323 */
__device__ int d_coster_kronig(const xrl_atom *a, phx_stream *s, int shell) {
  const double *f = a->ck;
  if (shell == 1 || shell == 2) {
    for (;;) {
      double r = phx_uniform(s);
      if (shell == 2) {
        if (r < f[XRL_FL23]) shell = 3;
        return shell;
      }
      if (r < f[XRL_FL12]) {
        shell = 2;
        continue;
      }
      if (r < f[XRL_FL12] + f[XRL_FL13]) shell = 3;
      return shell;
    }
  } else if (shell >= 4 && shell <= 7) {
    for (;;) {
      double r = phx_uniform(s);
      if (shell == 7) {
        if (r < f[XRL_FM45]) shell = 8;
        return shell;
      } else if (shell == 6) {
        if (r < f[XRL_FM34]) {
          shell = 7;
          continue;
        }
        if (r < f[XRL_FM34] + f[XRL_FM35]) shell = 8;
        return shell;
      } else if (shell == 5) {
        double sum = f[XRL_FM23];
        if (r < sum) {
          shell = 6;
          continue;
        }
        sum += f[XRL_FM24];
        if (r < sum) {
          shell = 7;
          continue;
        }
        sum += f[XRL_FM25];
        if (r < sum) shell = 8;
        return shell;
      } else {
        double sum = f[XRL_FM12];
        if (r < sum) {
          shell = 5;
          continue;
        }
        sum += f[XRL_FM13];
        if (r < sum) {
          shell = 6;
          continue;
        }
        sum += f[XRL_FM14];
        if (r < sum) {
          shell = 7;
          continue;
        }
        sum += f[XRL_FM15];
        if (r < sum) shell = 8;
        return shell;
      }
    }
  }
  return shell;
}

/* ------------------------------------------------------------------------ */
/* cascades of brute-force runs, src/xmi_main.F90:2413-4783 */

/* xmi_fluorescence_yield_check (src/xmi_main.F90:5325-5350): 1 when the
 * vacancy fluoresces.  Shells beyond M5 never do (the reference reads past
 * the end of its nine corrected yields there); the number is drawn all the
 * same. */
__device__ __forceinline__ int d_yield_check(const EngineParams &P, int u, int shell, phx_stream *s) {
  const double r = phx_uniform(s);
  const double y = (shell >= 0 && shell <= 8) ? P.fycorr[(size_t)u * 9 + shell] : 0.0;
  return !(r > y);
}

/* xmi_fluorescence_line_check (src/xmi_main.F90:5352-5437): the line of a
 * vacancy by the radiative rates, its energy, or 0 and line 0 */
__device__ double d_line_check(const EngineParams &P, const xrl_atom *a, int u, int shell, phx_stream *s,
                               int *line) {
  const double r = phx_uniform(s);
  *line = 0;
  if (shell < 0 || shell > 8) return 0.0;
  int first, last;
  xrl_shell_lines(shell, &first, &last);
  double sumz = 0.0;
  for (int l = first; l <= last; l++) {
    sumz += a->rad_rate[l];
    if (r < sumz) {
      *line = l;
      return P.line_E[(size_t)u * XMI_NLINES_VR_DEV + (l - 1)];
    }
  }
  return 0.0;
}

/* a cascade's photon in slot j: energy and attenuation, isotropic direction,
 * random polarisation made orthogonal to it, history row of its line, no
 * cascades left (src/xmi_main.F90:4330-4370,4720-4765) */
__device__ void d_cascade_emit(const EngineParams &P, const PhotonSoA &S, int j, double E, int line, int u,
                               phx_stream *s) {
  S.energy[j] = E;
  d_mu_all(P, E, S.mus, (size_t)P.cap, j);
  S.cflags[j] = 0;
  const double th = acos(2.0 * phx_uniform(s) - 1.0);
  const double ph = 2.0 * ENG_PI * phx_uniform(s);
  const double dirv[3] = {sin(th) * cos(ph), sin(th) * sin(ph), cos(th)};
  const double r = 2.0 * ENG_PI * phx_uniform(s);
  double ev[3] = {cos(r), sin(r), 0.0};
  const double cosalfa = d_dot(ev, dirv);
  const double c_alfa = acos(cosalfa);
  const double c_ae = 1.0 / sin(c_alfa), c_be = -c_ae * cosalfa;
  for (int k = 0; k < 3; k++) ev[k] = c_ae * ev[k] + c_be * dirv[k];
  S.dx[j] = dirv[0];
  S.dy[j] = dirv[1];
  S.dz[j] = dirv[2];
  S.ex[j] = ev[0];
  S.ey[j] = ev[1];
  S.ez[j] = ev[2];
  S.theta[j] = th;
  S.phi[j] = ph;
  S.prev_line[j] = line;
  S.prev_u[j] = u;
}

/* the offspring of slot i starts at the parent's interaction point, in its
 * layer, with its weight and history id */
__device__ __forceinline__ int d_offspring(const EngineParams &P, const PhotonSoA &S, int i) {
  const int j = i + P.off_half;
  S.x[j] = S.x[i];
  S.y[j] = S.y[i];
  S.z[j] = S.z[i];
  S.layer[j] = S.layer[i];
  S.elem[j] = S.elem[i];
  S.weight[j] = S.weight[i];
  S.hid[j] = S.hid[i];
  return j;
}

/* xmi_simulate_photon_cascade_auger (src/xmi_main.F90:2413-4594): the
 * vacancy's Auger transition (the provider's rates) leaves two vacancies;
 * the first may fluoresce as the parent photon, the second as the
 * offspring.  M vacancies count with use_M_lines only. */
__device__ int d_cascade_auger(const EngineParams &P, const PhotonSoA &S, int i, int u, const xrl_atom *a,
                               int shell, phx_stream *s) {
  int first, last, spawned = 0;
  if (!xrl_auger_range(shell, &first, &last)) return 0;
  const double r = phx_uniform(s);
  double sumz = 0.0;
  int t = -1;
  for (int k = first; k <= last; k++) {
    sumz += xrl_auger_rate(a, k);
    if (r < sumz) {
      t = k;
      break;
    }
  }
  if (t < 0) return 0;
  int sh[2];
  xrl_auger_shells(t, &sh[0], &sh[1]);
  for (int k = 0; k < 2; k++) {
    const int m_shell = sh[k] >= XRL_M1_SHELL && sh[k] <= XRL_M5_SHELL;
    if ((m_shell && !P.use_M_lines) || !d_yield_check(P, u, sh[k], s)) continue;
    const int v = d_coster_kronig(a, s, sh[k]);
    int line;
    const double e = d_line_check(P, a, u, v, s, &line);
    if (e <= 1.0) continue;
    d_cascade_emit(P, S, k == 0 ? i : d_offspring(P, S, i), e, line, u, s);
    spawned = k == 1;
  }
  return spawned;
}

/* xmi_simulate_photon_cascade_radiative (src/xmi_main.F90:4596-4783): the
 * line the parent emitted leaves a vacancy in its other shell (K -> L or M;
 * L -> M with use_M_lines), which may fluoresce as the offspring; then the
 * parent has no cascades left either */
__device__ int d_cascade_radiative(const EngineParams &P, const PhotonSoA &S, int i, int u, const xrl_atom *a,
                                   int shell, int line, phx_stream *s) {
  int sn = -1;
  if (shell == XRL_K_SHELL) {
    if (line >= xrl_line_KL(0) && line <= xrl_line_KM(4)) sn = line; /* KL1..KL3, KM1..KM5 -> L1..M5 */
  } else if (shell >= XRL_L1_SHELL && shell <= XRL_L3_SHELL && P.use_M_lines) {
    const int base = xrl_line_LM(shell - 1, 0);
    if (line >= base && line <= base + 4) sn = XRL_M1_SHELL + (line - base);
  } else {
    return 0;
  }
  if (sn < 0) return 0;
  if (sn >= XRL_M1_SHELL && sn <= XRL_M5_SHELL && !P.use_M_lines) return 0;
  if (!d_yield_check(P, u, sn, s)) return 0;
  sn = d_coster_kronig(a, s, sn);
  int ln;
  const double e = d_line_check(P, a, u, sn, s, &ln);
  if (e <= 1.0) return 0;
  S.cflags[i] = 0;
  d_cascade_emit(P, S, d_offspring(P, S, i), e, ln, u, s);
  return 1;
}

/* returns 1 when a cascade started an offspring photon (brute force) */
template <bool BRUTE>
__device__ __forceinline__ int d_photo_one(const EngineParams &P, const PhotonSoA &S, int n, int i) {
  const size_t cap = (size_t)P.cap;
  int L = S.layer[i];
  int u = P.el_u[P.layer_off[L - 1] + S.elem[i]];
  const xrl_atom *a = &P.atoms[u];
  phx_stream s;
  phx_stream_init(&s, P.seed, P.sample, (uint32_t)n, BRUTE ? d_tr_sub(P, i) : PHX_SUB_TRANSPORT, S.hid[i]);
  d_stream_seek(&s, 3u);
  double E = S.energy[i];
  xrl_ectx ec = xrl_energy(E);
  double photo_total = xrl_cs_photo_total_c(a, &ec);
  double sumz = 0.0;
  double r = phx_uniform(&s);
  int max_shell = P.use_M_lines ? 8 : 3;
  int shell = -1;
  for (int sh = 0; sh <= max_shell; sh++) {
    sumz += xrl_cs_photo_partial_c(a, sh, &ec) / photo_total;
    if (r < sumz) {
      shell = sh;
      break;
    }
  }
  S.prev_line[i] = 0;
  if (shell < 0) {
    S.energy[i] = 0.0;
    return 0;
  }
  const double ry = phx_uniform(&s); /* unused by the reference's yield weighting */
  if (BRUTE) {
    /* the analogue yield, perhaps an Auger cascade (src/xmi_main.F90:2298-2315) */
    if (ry > P.fycorr[(size_t)u * 9 + shell]) {
      S.energy[i] = 0.0;
      return (S.cflags[i] & 1) ? d_cascade_auger(P, S, i, u, a, shell, &s) : 0;
    }
  } else {
    S.weight[i] *= P.fycorr[(size_t)u * 9 + shell];
  }
  shell = d_coster_kronig(a, &s, shell);
  r = phx_uniform(&s);
  sumz = 0.0;
  int first, lastl, line = 0;
  xrl_shell_lines(shell, &first, &lastl);
  for (int l = first; l <= lastl; l++) {
    sumz += a->rad_rate[l];
    if (r < sumz) {
      line = l;
      break;
    }
  }
  if (line == 0) {
    S.energy[i] = 0.0;
    return 0;
  }
  E = P.line_E[(size_t)u * XMI_NLINES_VR_DEV + (line - 1)];
  S.energy[i] = E;
  for (int l = 0; l < P.n_layers; l++)
    S.mus[(size_t)l * cap + i] = P.precalc_mu[((size_t)l * P.nZ + u) * XMI_NLINES_VR_DEV + (line - 1)];
  double theta_i = acos(-2.0 * phx_uniform(&s) + 1.0);
  double phi_i = 2.0 * ENG_PI * phx_uniform(&s);
  double th = S.theta[i], ph = S.phi[i];
  double ev[3] = {S.ex[i], S.ey[i], S.ez[i]};
  double dirv[3];
  d_update_dirv(dirv, &th, &ph, theta_i, phi_i);
  d_update_elecv(dirv, ev);
  S.dx[i] = dirv[0];
  S.dy[i] = dirv[1];
  S.dz[i] = dirv[2];
  S.ex[i] = ev[0];
  S.ey[i] = ev[1];
  S.ez[i] = ev[2];
  S.theta[i] = th;
  S.phi[i] = ph;
  S.prev_line[i] = line;
  S.prev_u[i] = u;
  if (BRUTE && (S.cflags[i] & 2)) return d_cascade_radiative(P, S, i, u, a, shell, line, &s);
  return 0;
}

/* exclusive prefix sums of the 3 x ENG_SPREAD sub-queue lengths of the
 * interaction queues into off[0..3 ENG_SPREAD], off[3 ENG_SPREAD] = total
 * (first wave of the workgroup; the caller synchronises) */
__device__ __forceinline__ void d_queue_offsets(const unsigned long long *qc, unsigned int *off) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    unsigned int base = 0;
    for (int t = 0; t < 3; t++) {
      const unsigned int c = (unsigned int)qc[(size_t)t * ENG_SPREAD + lane];
      unsigned int x = c;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const unsigned int y = __shfl_up(x, d);
        if (lane >= d) x += y;
      }
      off[t * ENG_SPREAD + lane] = base + x - c;
      base += __shfl(x, 63);
    }
    if (lane == 0) off[3 * ENG_SPREAD] = base;
  }
}

/* wavefront-aggregated append of the lanes with pred to a list (one atomic
 * per wave; for rare entries such as a cascade's offspring) */
__device__ __forceinline__ void d_list_append(int *list, unsigned int *cnt, int pred, int idx) {
  const unsigned long long m = __ballot(pred);
  if (!m) return;
  const int lane = __lane_id(), leader = __ffsll((long long)m) - 1;
  unsigned int base = 0;
  if (lane == leader) base = atomicAdd(cnt, (unsigned int)__popcll(m));
  base = __shfl(base, leader);
  if (pred) list[base + __popcll(m & ((1ull << lane) - 1ull))] = idx;
}

/* entry q of the concatenated interaction queues (Rayleigh, Compton, photo;
 * each queue's ENG_SPREAD sub-queues in order; off: their prefix sums):
 * the photon's interaction.  Returns 1 when a cascade started an offspring */
template <bool BRUTE, bool ADV>
__device__ __forceinline__ int d_interact_entry(const EngineParams &P, const PhotonSoA &S, int n,
                                                const unsigned int *off, unsigned int q, int *slot) {
  constexpr int NS = 3 * ENG_SPREAD;
  int lo = 0, hi = NS;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= q) lo = mid;
    else hi = mid;
  }
  const int t = lo / ENG_SPREAD;
  const int i = P.queue[(size_t)lo * P.qreg + (q - off[lo])];
  *slot = i;
  if (t == 0) d_rayleigh_one<BRUTE>(P, S, n, i);
  else if (t == 1) d_compton_one<BRUTE, ADV>(P, S, n, i);
  else return d_photo_one<BRUTE>(P, S, n, i);
  return 0;
}

/* one launch serves the three interaction queues: thread q takes entry q of
 * their concatenation, so the grid's empty waves are only those beyond the
 * photons that interact at this order, and a wavefront straddles two queues
 * at two places at most.  Each photon is one thread's, with its own random
 * stream, so results do not depend on the assignment. */
/* occupancy: five waves per SIMD (96 VGPRs, a few spilled; 6 % faster than four).
 * ADV: the advanced Compton model, compiled apart so that the default
 * instance spills less (10 VGPRs instead of 28) */
template <bool ADV>
__global__ void __launch_bounds__(ENG_BLOCK) __attribute__((amdgpu_waves_per_eu(5)))
k_interact(EngineParams P, PhotonSoA S, int n) {
  __shared__ unsigned int off[3 * ENG_SPREAD + 1];
  d_queue_offsets(P.qcount, off);
  __syncthreads();
  const unsigned int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= off[3 * ENG_SPREAD]) return;
  int i;
  (void)d_interact_entry<false, ADV>(P, S, n, off, q, &i);
}

/* the brute-force interactions (analogue yields, cascades), one thread per
 * queue entry as in k_interact; the offspring that cascades start join the
 * offspring list.  (A fixed grid striding over the entries would keep the
 * kernel's many table pointers live across the loop: at five waves per SIMD
 * that spills 200 VGPRs instead of about 50.)  Four waves per SIMD run 3 %
 * faster than five here. */
template <bool ADV>
__global__ void __launch_bounds__(ENG_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_interact_brute(EngineParams P, PhotonSoA S, int n) {
  __shared__ unsigned int off[3 * ENG_SPREAD + 1];
  d_queue_offsets(P.qcount, off);
  __syncthreads();
  const unsigned int total = off[3 * ENG_SPREAD], q = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x * blockDim.x >= total) return;
  int i = -1, spawned = 0;
  if (q < total) spawned = d_interact_entry<true, ADV>(P, S, n, off, q, &i);
  d_list_append(P.olist_out, P.ocount_out, spawned, i + P.off_half);
}

/* ------------------------------------------------------------------------ */
/* escape-ratio runs (src/xmi_main.F90:5473-5801 with the analogue step of
 * :1274-1407): after its forced first interaction a photon leaves the crystal
 * when the analogue step's random number (the first draw of the next order's
 * transport stream) exceeds the probability of interacting on the way out;
 * escaped Compton photons are tallied by energy, fluorescence photons by
 * element and line (weight times fluorescence yield), and the weight after
 * the forced step of every interacted photon normalises its input energy.
 * One thread per queue entry of the three interaction queues. */
/* exact wave-aggregated fixed-point add: every lane with key >= 0 adds
 * value to bin key of tally; the lanes of one key are summed as integers
 * (three 42-bit limbs, which 64 lanes cannot overflow) and added once, so
 * that thousands of photons of one bin do not serialise on its address and
 * the sum does not depend on which photons share a wavefront.  Call from
 * converged code. */
__device__ __forceinline__ void d_fx_add_keyed(unsigned long long *tally, long key, double value, int shift) {
  unsigned long long lo = 0ull, hi = 0ull;
  if (key >= 0) fx_split(value, shift, &lo, &hi);
  const unsigned long long M42 = (1ull << 42) - 1ull;
  const unsigned long long l0 = lo & M42, l1 = (lo >> 42) | ((hi & ((1ull << 20) - 1ull)) << 22), l2 = hi >> 20;
  unsigned long long pending = __ballot(key >= 0);
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const long k = __shfl(key, leader);
    const int mine = key == k;
    unsigned long long a0 = mine ? l0 : 0ull, a1 = mine ? l1 : 0ull, a2 = mine ? l2 : 0ull;
    for (int off = 32; off > 0; off >>= 1) {
      a0 += __shfl_xor(a0, off);
      a1 += __shfl_xor(a1, off);
      a2 += __shfl_xor(a2, off);
    }
    if (__lane_id() == leader) {
      const unsigned __int128 t = (unsigned __int128)a0 + ((unsigned __int128)a1 << 42) + ((unsigned __int128)a2 << 84);
      fx_merge_global(&tally[(size_t)k * 2], (unsigned long long)t, (unsigned long long)(t >> 64));
    }
    pending &= ~__ballot(mine);
  }
}

__global__ void __launch_bounds__(ENG_BLOCK) k_escape(EngineParams P, PhotonSoA S, VrSoA V, EscParams Q) {
  const unsigned int per = (unsigned int)ENG_SPREAD * (unsigned int)P.qreg;
  const unsigned int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int type = (int)(q / per);
  const int i = type <= 2 ? d_dequeue(P, type, q - (unsigned int)type * per) : -1;
  const int shift = P.tally_shift;
  long key_n = -1, key_c = -1, key_f = -1;
  double w_n = 0.0, w = 0.0;
  if (i >= 0) {
    const int iE = (int)(S.hid[i] / (uint64_t)Q.n_photons);
    if (iE >= 0 && iE < Q.nE) {
      /* the photon's weight after the forced step (V.pesc) normalises;
       * the photo-absorption step has multiplied S.weight by the
       * fluorescence yield, which an
       * escape-ratio run applies to the escape tally only
       * (src/xmi_variance_reduction.F90:742-746) */
      key_n = iE;
      w_n = V.pesc[i];
      w = S.weight[i];
      const double E = S.energy[i];
      const int cur = S.layer[i];
      int escaped = E >= 1.0 && cur >= 1; /* below the energy threshold: stays inside */
      if (escaped) {
        phx_stream s;
        phx_stream_init(&s, P.seed, P.sample, 2u, PHX_SUB_TRANSPORT, S.hid[i]);
        const double R = phx_uniform(&s);
        const double dirv[3] = {S.dx[i], S.dy[i], S.dz[i]};
        const int up = d_dot(dirv, P.n_sample) > 0.0;
        double lp[3] = {S.x[i], S.y[i], S.z[i]};
        double max_r = 0.0, blbs = 1.0;
        const size_t cap = (size_t)P.cap;
        for (int l = cur - 1; up ? l < P.n_layers : l >= 0; l += up ? 1 : -1) {
          double pp[3] = {0.0, 0.0, up ? P.layer_ze[l] : P.layer_zb[l]}, ip[3];
          d_intersect(pp, P.n_sample, lp, dirv, ip);
          const double dist = d_distance(lp, ip);
          const double temp_prod = -1.0 * dist * P.layer_density[l] * S.mus[(size_t)l * cap + i];
          const double tempexp = exp(temp_prod);
          max_r = max_r - blbs * expm1(temp_prod);
          if (R <= max_r) { /* interacts on the way out: inside */
            escaped = 0;
            break;
          }
          lp[0] = ip[0];
          lp[1] = ip[1];
          lp[2] = ip[2];
          blbs = blbs * tempexp;
        }
      }
      if (escaped && type == 1) {
        const int idx = (int)((E - Q.out0) / Q.dOut) + 1;
        if (idx >= 1 && idx <= Q.nOut) key_c = (long)(idx - 1) * Q.nE + iE;
      } else if (escaped && type == 2) {
        const int line = S.prev_line[i], u = S.prev_u[i];
        if (line >= 1 && line <= ENG_ESC_LINES && u >= 0) key_f = ((long)iE * ENG_ESC_LINES + (line - 1)) * P.nZ + u;
      }
    }
  }
  d_fx_add_keyed(Q.interacted, key_n, w_n, shift);
  d_fx_add_keyed(Q.compton, key_c, w, shift);
  d_fx_add_keyed(Q.fluo, key_f, w, shift);
}

/* ------------------------------------------------------------------------ */
/* host-side launch wrappers */

static size_t lds_if_fits(size_t words) {
  size_t b = words * sizeof(unsigned long long);
  return b <= ENG_LDS_BYTES ? b : 0;
}
static size_t scatter_lds_bytes(const EngineParams *P) { return lds_if_fits((size_t)P->nch * 2 + (size_t)P->nZ * 4); }
static size_t lines_lds_bytes(const EngineParams *P) { return lds_if_fits((size_t)P->n_lbins * 2); }

/* Workgroups of a grid-stride kernel: as many as can be resident at once,
 * since every workgroup ends with one flush of its LDS tallies to global
 * memory.  per_cu > 0 (the run's launch configuration) overrides the
 * resident count per CU. */
static int resident_blocks(const void *kernel, size_t lds, int n_cu, int blocks, int per_cu = 0) {
  if (per_cu <= 0 &&
      (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, ENG_BLOCK, lds) != hipSuccess || per_cu < 1))
    per_cu = 1;
  int vb = n_cu * per_cu;
  if (vb > blocks) vb = blocks;
  return vb < 1 ? 1 : vb;
}

/* ------------------------------------------------------------------------ */
/* brute-force runs (use_variance_reduction = 0): the analogue step of
 * src/xmi_main.F90:1274-1407 with the detector tests of every path segment,
 * one thread per photon slot; a photon that reaches the detector is tallied
 * by its energy and its last interaction (src/xmi_main.F90:445-535) and its
 * history ends, as it does when it leaves the sample elsewhere or hits the
 * collimator */

/* lab -> detector frame: x along the detector axis towards the sample, the
 * window at x = 0 */
__device__ __forceinline__ void d_det_frame(const EngineParams &P, const double *x, double *o) {
  const double d[3] = {x[0] - P.p_det[0], x[1] - P.p_det[1], x[2] - P.p_det[2]};
  d_matvec(P.det_inv, d, o);
}

/* xmi_check_detector_intersection (src/xmi_aux_f.F90:1622-1833) of the path
 * segment b -> e: 0 none, 1 collimator, 2 window, 3 window from behind.
 * Without a collimator the reference's line direction is norm(e - b), a
 * scalar assigned to all three components (src/xmi_aux_f.F90:1668): the line
 * (L, L, L) crosses x = 0 at (0, e_y - e_x, e_z - e_x), always as a crossing
 * from behind.  Kept as computed there.  A line parallel to the window,
 * where the reference stops the program, counts as no crossing. */
__device__ int d_detector_segment(const EngineParams &P, const double *b, const double *e) {
  double bd[3], ed[3], ip[3];
  d_det_frame(P, b, bd);
  d_det_frame(P, e, ed);
  const double o[3] = {0.0, 0.0, 0.0}, nx[3] = {1.0, 0.0, 0.0};
  if (!P.coll_present) {
    if (bd[0] * ed[0] > 0.0) return 0;
    const double L = d_distance(ed, bd);
    const double ld[3] = {L, L, L};
    if (!d_intersect(o, nx, ed, ld, ip)) return 0;
    ip[0] = 0.0;
    if (d_norm(ip) <= P.det_radius) return ld[0] >= 0.0 ? 3 : 2;
    return 0;
  }
  const double ld[3] = {ed[0] - bd[0], ed[1] - bd[1], ed[2] - bd[2]};
  const double t_begin = (bd[0] - ed[0]) / ld[0], t_end = (ed[0] - ed[0]) / ld[0];
  const double t_hi = t_begin > t_end ? t_begin : t_end, t_lo = t_begin < t_end ? t_begin : t_end;
  const double c2t = P.det_cos2;
  const double m[3] = {1.0 - c2t, -c2t, -c2t};
  const double delta[3] = {ed[0] - P.det_vertex[0], ed[1] - P.det_vertex[1], ed[2] - P.det_vertex[2]};
  const double dM[3] = {ld[0] * m[0], ld[1] * m[1], ld[2] * m[2]};
  const double deltaM[3] = {delta[0] * m[0], delta[1] * m[1], delta[2] * m[2]};
  const double c2 = d_dot(dM, ld), c1 = d_dot(dM, delta), c0 = d_dot(deltaM, delta);
  const double disc = c1 * c1 - c0 * c2;
  if (disc < 0.0) return 0;
  const double t1 = (-c1 + sqrt(disc)) / c2, t2 = (-c1 - sqrt(disc)) / c2;
  const double x1 = ed[0] + t1 * ld[0], x2 = ed[0] + t2 * ld[0];
  const int v1 = -(x1 - P.det_vertex[0]) >= 0.0, v2 = -(x2 - P.det_vertex[0]) >= 0.0;
  const double h = P.coll_height;
  if (!v1 && !v2) return 0;
  if (v1 && v2) {
    if (t1 <= t_hi && t1 >= t_lo && x1 <= h) return 1;
    if (t2 <= t_hi && t2 >= t_lo && x2 <= h) return 1;
    return 0;
  }
  const double t = v1 ? t1 : t2, xx = v1 ? x1 : x2;
  if (t <= t_hi && t >= t_lo && xx <= h) return 1;
  if (!d_intersect(o, nx, ed, ld, ip)) return 0;
  ip[0] = 0.0;
  if (d_norm(ip) <= P.det_radius && d_distance(bd, ip) <= d_distance(bd, ed)) return ld[0] >= 0.0 ? 3 : 2;
  return 0;
}

/* xmi_check_photon_detector_hit (src/xmi_main.F90:1920-1984): a photon that
 * left the sample reaches the detector when it moves against the detector
 * normal and crosses the window (and the collimator's top) within their
 * radii */
__device__ int d_escape_hits_detector(const EngineParams &P, const double *coords, const double *dirv) {
  if (d_dot(dirv, P.n_det) >= 0.0) return 0;
  double dd[3], cd[3], ip[3];
  d_matvec(P.det_inv, dirv, dd);
  d_det_frame(P, coords, cd);
  const double o[3] = {0.0, 0.0, 0.0}, nx[3] = {1.0, 0.0, 0.0};
  if (!d_intersect(o, nx, cd, dd, ip)) return 0;
  if (d_norm(ip) > P.det_radius) return 0;
  if (!P.coll_present) return 1;
  const double pc[3] = {P.coll_height, 0.0, 0.0};
  if (!d_intersect(pc, nx, cd, dd, ip)) return 0;
  ip[0] = 0.0;
  return d_norm(ip) <= P.coll_radius;
}

/* the analogue step of photon i at order n: a photon alive after n - 1
 * interactions takes its step; one that stays in the sample with fewer than
 * n_int interactions gets its atom and interaction type (*type), the others
 * end, those that reach the detector tallied (*key_ch, *key_h, *w_hit) */
__device__ __forceinline__ void d_brute_step(const EngineParams &P, const PhotonSoA &S, int n, int i, int *type,
                                             int *stepped, int *reached, long *key_ch, long *key_h, double *w_hit) {
  const size_t cap = (size_t)P.cap;
  const double E = S.energy[i];
  if (E < 1.0) {
    S.layer[i] = 0;
    return;
  }
  *stepped = 1;
  phx_stream s;
  phx_stream_init(&s, P.seed, P.sample, (uint32_t)n, d_tr_sub(P, i), S.hid[i]);
  const double R = phx_uniform(&s);
  const double dirv[3] = {S.dx[i], S.dy[i], S.dz[i]};
  double c[3] = {S.x[i], S.y[i], S.z[i]};
  const int cur = S.layer[i];
  const int up = d_dot(dirv, P.n_sample) > 0.0;
  const int last = up ? P.n_layers : 1, dir = up ? 1 : -1;
  double max_r = 0.0, min_r = 0.0, blbs = 1.0;
  int inside = 0, hit = 0, ended = 0;
  for (int l = cur; up ? l <= last : l >= last; l += dir) {
    const int l0 = l - 1;
    const double pp[3] = {0.0, 0.0, up ? P.layer_ze[l0] : P.layer_zb[l0]};
    double ip[3];
    d_intersect(pp, P.n_sample, c, dirv, ip);
    const double mu = S.mus[(size_t)l0 * cap + i];
    double dist = d_distance(c, ip);
    const double temp_prod = -1.0 * dist * P.layer_density[l0] * mu;
    const double tempexp = exp(temp_prod);
    min_r = max_r;
    max_r = max_r - blbs * expm1(temp_prod);
    if (R <= max_r) {
      dist = -1.0 * log1p(-1.0 * (R - min_r) / blbs) / mu / P.layer_density[l0];
      const double to[3] = {c[0] + dist * dirv[0], c[1] + dist * dirv[1], c[2] + dist * dirv[2]};
      const int rc = d_detector_segment(P, c, to);
      c[0] = to[0];
      c[1] = to[1];
      c[2] = to[2];
      if (rc) {
        ended = 1;
        hit = rc == 2;
      } else {
        inside = l;
      }
      break;
    }
    const int rc = d_detector_segment(P, c, ip);
    if (rc) {
      ended = 1;
      hit = rc == 2;
      break;
    }
    c[0] = ip[0];
    c[1] = ip[1];
    c[2] = ip[2];
    blbs = blbs * tempexp;
  }
  if (!inside && !ended) hit = d_escape_hits_detector(P, c, dirv);
  if (hit) {
    /* channels(n_interactions:, channel) and the history row of the last
     * interaction; the host sums the orders into the cumulative rows */
    *reached = 1;
    *w_hit = S.weight[i];
    const int ch = (int)((E - P.zero) / P.gain);
    if (ch >= 0 && ch < P.nch) *key_ch = (long)(n - 1) * P.nch + ch;
    if (n >= 2 && S.prev_line[i] > 0 && S.prev_u[i] >= 0)
      *key_h = ((long)(n - 2) * P.nZ + S.prev_u[i]) * 385 + (S.prev_line[i] - 1);
  }
  if (!inside || n > P.n_int) {
    S.layer[i] = 0;
  } else {
    S.x[i] = c[0];
    S.y[i] = c[1];
    S.z[i] = c[2];
    S.layer[i] = inside;
    *type = d_select(P, S, i, E, inside, &s);
  }
}

/* order n of a brute-force run.  The photons alive at its start are the
 * batch's primaries at n = 1; afterwards the entries of order n - 1's
 * interaction queues (those that interacted) followed by the offspring that
 * its cascades started.  A fixed grid strides over them, every lane of a
 * wave taking the same trips, so that the queue appends, counters and
 * tallies see converged waves.  Three waves per SIMD (166 VGPRs, 8 spilled)
 * run 17 % faster than the two that 183 unconstrained VGPRs allow. */
__global__ void __launch_bounds__(ENG_BLOCK) __attribute__((amdgpu_waves_per_eu(3)))
k_step_brute(EngineParams P, PhotonSoA S, int n, int count) {
  constexpr int NS = 3 * ENG_SPREAD;
  __shared__ unsigned int off[NS + 1];
  unsigned int nq = 0, total = (unsigned int)count;
  if (n >= 2) {
    d_queue_offsets(P.pqcount, off);
    __syncthreads();
    nq = off[NS];
    total = nq + (P.off_half ? *P.ocount_in : 0u);
  }
  const unsigned int stride = gridDim.x * blockDim.x;
  for (unsigned int base = blockIdx.x * blockDim.x; base < total; base += stride) {
    const unsigned int q = base + threadIdx.x;
    int type = -1, stepped = 0, reached = 0, i = -1;
    long key_ch = -1, key_h = -1;
    double w_hit = 0.0;
    if (q < total) {
      if (n == 1) {
        i = (int)q;
      } else if (q < nq) {
        int lo = 0, hi = NS;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (off[mid] <= q) lo = mid;
          else hi = mid;
        }
        i = P.pqueue[(size_t)lo * P.qreg + (q - off[lo])];
      } else {
        i = P.olist_in[q - nq];
      }
      if (S.layer[i] > 0) d_brute_step(P, S, n, i, &type, &stepped, &reached, &key_ch, &key_h, &w_hit);
    }
    d_enqueue(P, type, i);
    d_wave_count(d_ctr(P, ENG_CNT_STEPS), stepped);
    d_wave_count(d_ctr(P, ENG_CNT_HITS), reached);
    d_fx_add_keyed(P.t_bch, key_ch, w_hit, P.tally_shift);
    d_fx_add_keyed(P.t_bhist, key_h, w_hit, P.tally_shift);
  }
}

static void launch_interact(const EngineParams *P, const PhotonSoA *S, int n, int blocks, hipStream_t st) {
  if (P->adv_compton) hipLaunchKernelGGL(k_interact<true>, dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, n);
  else hipLaunchKernelGGL(k_interact<false>, dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, n);
}

static void launch_interact_brute(const EngineParams *P, const PhotonSoA *S, int n, int blocks, hipStream_t st) {
  if (P->adv_compton) hipLaunchKernelGGL(k_interact_brute<true>, dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, n);
  else hipLaunchKernelGGL(k_interact_brute<false>, dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, n);
}

template <int NL, int KP>
static void launch_lines_kp(const EngineParams *P, const PhotonSoA *S, const VrSoA *V, int n, int n_cu, int blocks,
                            hipStream_t st) {
  const size_t ll = lines_lds_bytes(P);
  if (ll) {
    int vb = resident_blocks((const void *)k_vr_lines<NL, KP, true>, ll, n_cu, blocks, P->wg_per_cu);
    hipLaunchKernelGGL((k_vr_lines<NL, KP, true>), dim3(vb), dim3(ENG_BLOCK), ll, st, *P, *S, *V, n);
  } else {
    int vb = resident_blocks((const void *)k_vr_lines<NL, KP, false>, 0, n_cu, blocks, P->wg_per_cu);
    hipLaunchKernelGGL((k_vr_lines<NL, KP, false>), dim3(vb), dim3(ENG_BLOCK), 0, st, *P, *S, *V, n);
  }
}

/* photons per lane of k_vr_lines (P->lines_kp, 0: the default): each
 * line's wavefront sum then serves 64 KP photons; fewer where the
 * per-photon path lengths take more registers.  Three rather than four up
 * to four layers by default: 126 VGPRs give four waves per SIMD instead of
 * three (3 % faster) */
template <int NL>
static void launch_lines(const EngineParams *P, const PhotonSoA *S, const VrSoA *V, int n, int n_cu, int blocks,
                         hipStream_t st) {
  const int kp = P->lines_kp;
  if constexpr (NL <= 4) {
    if (kp == 2) launch_lines_kp<NL, 2>(P, S, V, n, n_cu, blocks, st);
    else if (kp == 4) launch_lines_kp<NL, 4>(P, S, V, n, n_cu, blocks, st);
    else launch_lines_kp<NL, 3>(P, S, V, n, n_cu, blocks, st);
  } else if constexpr (NL <= 8) {
    if (kp == 1) launch_lines_kp<NL, 1>(P, S, V, n, n_cu, blocks, st);
    else launch_lines_kp<NL, 2>(P, S, V, n, n_cu, blocks, st);
  } else {
    launch_lines_kp<NL, 1>(P, S, V, n, n_cu, blocks, st);
  }
}

template <int NL, bool ADV>
static void launch_scatter(const EngineParams *P, const PhotonSoA *S, const VrSoA *V, int n, int n_cu, int blocks,
                           size_t ls, hipStream_t st) {
  if (ls) {
    int vb = resident_blocks((const void *)k_vr_scatter<NL, true, ADV>, ls, n_cu, blocks, P->wg_per_cu);
    hipLaunchKernelGGL((k_vr_scatter<NL, true, ADV>), dim3(vb), dim3(ENG_BLOCK), ls, st, *P, *S, *V, n);
  } else {
    int vb = resident_blocks((const void *)k_vr_scatter<NL, false, ADV>, 0, n_cu, blocks, P->wg_per_cu);
    hipLaunchKernelGGL((k_vr_scatter<NL, false, ADV>), dim3(vb), dim3(ENG_BLOCK), 0, st, *P, *S, *V, n);
  }
}

static void launch_cone(const EngineParams *P, const PhotonSoA *S, const VrSoA *V, int n, int n_cu, hipStream_t st) {
  /* one resident round: the trial segments are of equal size, so a second,
   * partial round of workgroups would only add a tail */
  const int cb = resident_blocks((const void *)k_vr_cone, 0, n_cu, n_cu * 4);
  hipLaunchKernelGGL(k_vr_cone, dim3(cb), dim3(ENG_BLOCK), 0, st, *P, *S, *V, n);
}

/* the estimator of order n after the step and the layer scan, each kernel
 * timed with ev.  (The line kernel on a second stream beside the scatter
 * kernel, and the off-grid solid angles beside the layer lists, measured no
 * faster: 9.50 and 9.21 s per step against 9.20 on one stream.) */
template <int NL>
static void launch_estimator_nl(const EngineParams *P, const PhotonSoA *S, const VrSoA *V, int n, int count,
                                int n_cu, int blocks, hipStream_t st, hipEvent_t *ev) {
  const size_t ls = scatter_lds_bytes(P);
  hipLaunchKernelGGL(k_vr_list, dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, count);
  if (ev) (void)hipEventRecord(ev[0], st);
  launch_cone(P, S, V, n, n_cu, st);
  if (ev) (void)hipEventRecord(ev[1], st);
  if (P->adv_compton)
    launch_scatter<NL, true>(P, S, V, n, n_cu, blocks, ls, st);
  else
    launch_scatter<NL, false>(P, S, V, n, n_cu, blocks, ls, st);
  if (ev) (void)hipEventRecord(ev[2], st);
  launch_lines<NL>(P, S, V, n, n_cu, blocks, st);
  if (ev) (void)hipEventRecord(ev[3], st);
}

/* NL, the layer-count bound of the estimator and the step: the smallest of
 * 1, 2, 4, 8, 16 that holds n_layers */
static void launch_estimator(const EngineParams *P, const PhotonSoA *S, const VrSoA *V, int n, int count, int n_cu,
                             int blocks, hipStream_t st, hipEvent_t *ev) {
  const int nl = P->n_layers;
  if (nl <= 1) launch_estimator_nl<1>(P, S, V, n, count, n_cu, blocks, st, ev);
  else if (nl <= 2) launch_estimator_nl<2>(P, S, V, n, count, n_cu, blocks, st, ev);
  else if (nl <= 4) launch_estimator_nl<4>(P, S, V, n, count, n_cu, blocks, st, ev);
  else if (nl <= 8) launch_estimator_nl<8>(P, S, V, n, count, n_cu, blocks, st, ev);
  else launch_estimator_nl<16>(P, S, V, n, count, n_cu, blocks, st, ev);
}

extern "C" {

hipError_t eng_launch_grid(const EngineParams *P, double *grid, const double *rv, const double *tv, int n_r, int n_t,
                           long hits, uint64_t seed, hipStream_t st) {
  long total = (long)n_r * n_t;
  int blocks = (int)((total + ENG_BLOCK - 1) / ENG_BLOCK);
  hipLaunchKernelGGL(k_solid_angle_grid, dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, grid, rv, tv, n_r, n_t, hits, seed);
  return hipGetLastError();
}

hipError_t eng_launch_source(const EngineParams *P, const PhotonSoA *S, const SourceTable *T, uint64_t start,
                             int count, hipStream_t st) {
  int blocks = (count + ENG_BLOCK - 1) / ENG_BLOCK;
  hipLaunchKernelGGL(k_source, dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *T, start, count,
                     (const BatchQueue *)nullptr);
  return hipGetLastError();
}

hipError_t eng_launch_batch_init(BatchQueue *Q, unsigned long long total, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_init, dim3(1), dim3(1), 0, st, Q, total);
  return hipGetLastError();
}

hipError_t eng_launch_source_queued(const EngineParams *P, const PhotonSoA *S, const SourceTable *T, BatchQueue *Q,
                                    int cap, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_next, dim3(1), dim3(1), 0, st, Q, cap);
  const int blocks = (cap + ENG_BLOCK - 1) / ENG_BLOCK;
  hipLaunchKernelGGL(k_source, dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *T, (uint64_t)0, cap,
                     (const BatchQueue *)Q);
  return hipGetLastError();
}

/* one escape-ratio batch: the forced first interaction, the weight after it
 * kept in V.pesc (the weight of the escape tallies), the interaction, the
 * escape test and tallies */
hipError_t eng_launch_escape(const EngineParams *P, const PhotonSoA *S, const VrSoA *V, const EscParams *Q,
                             int count, hipStream_t st) {
  int blocks = (count + ENG_BLOCK - 1) / ENG_BLOCK;
  hipLaunchKernelGGL(k_reset_queues, dim3(1), dim3(ENG_NQUEUES * ENG_SPREAD), 0, st, *P);
  {
    const int nl = P->n_layers;
    if (nl <= 1) hipLaunchKernelGGL((k_step<1, false>), dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, 1, count);
    else if (nl <= 2) hipLaunchKernelGGL((k_step<2, false>), dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, 1, count);
    else if (nl <= 4) hipLaunchKernelGGL((k_step<4, false>), dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, 1, count);
    else if (nl <= 8) hipLaunchKernelGGL((k_step<8, false>), dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, 1, count);
    else hipLaunchKernelGGL((k_step<16, false>), dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, 1, count);
  }
  hipError_t e = hipMemcpyAsync(V->pesc, S->weight, sizeof(double) * (size_t)count, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return e;
  launch_interact(P, S, 1, (count + ENG_BLOCK - 1) / ENG_BLOCK, st);
  const int eblocks = (int)((3 * (size_t)ENG_SPREAD * P->qreg + ENG_BLOCK - 1) / ENG_BLOCK);
  hipLaunchKernelGGL(k_escape, dim3(eblocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, *Q);
  return hipGetLastError();
}

hipError_t eng_launch_order_brute(const EngineParams *P, const PhotonSoA *S, int n, int count, int blocks,
                                  hipStream_t st, hipEvent_t *ev) {
  hipLaunchKernelGGL(k_reset_queues, dim3(1), dim3(ENG_NQUEUES * ENG_SPREAD), 0, st, *P);
  if (ev) (void)hipEventRecord(ev[0], st);
  hipLaunchKernelGGL(k_step_brute, dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, n, count);
  if (ev) (void)hipEventRecord(ev[1], st);
  /* the interactions: one thread per photon that may be alive */
  const int alive = (n == 1 || !P->off_half) ? count : 2 * count;
  if (n <= P->n_int) launch_interact_brute(P, S, n, (alive + ENG_BLOCK - 1) / ENG_BLOCK, st);
  if (ev) (void)hipEventRecord(ev[2], st);
  return hipGetLastError();
}

int eng_brute_blocks(int n_cu) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_step_brute, ENG_BLOCK, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return n_cu * per_cu;
}

hipError_t eng_launch_order(const EngineParams *P, const PhotonSoA *S, const VrSoA *V, int n, int count, int n_cu,
                            hipStream_t st, hipEvent_t *ev /* ENG_ORDER_EVENTS events or NULL */) {
  int blocks = (count + ENG_BLOCK - 1) / ENG_BLOCK;
  if (ev) (void)hipEventRecord(ev[0], st);
  hipLaunchKernelGGL(k_reset_queues, dim3(1), dim3(ENG_NQUEUES * ENG_SPREAD), 0, st, *P);
  {
    const int nl = P->n_layers;
    if (nl <= 1) hipLaunchKernelGGL((k_step<1, true>), dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, n, count);
    else if (nl <= 2) hipLaunchKernelGGL((k_step<2, true>), dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, n, count);
    else if (nl <= 4) hipLaunchKernelGGL((k_step<4, true>), dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, n, count);
    else if (nl <= 8) hipLaunchKernelGGL((k_step<8, true>), dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, n, count);
    else hipLaunchKernelGGL((k_step<16, true>), dim3(blocks), dim3(ENG_BLOCK), 0, st, *P, *S, *V, n, count);
  }
  {
    const int nw = blocks * (ENG_BLOCK / 64), nseg = (nw + ENG_SCAN_SEG - 1) / ENG_SCAN_SEG;
    hipLaunchKernelGGL(k_layer_scan, dim3(nseg, P->n_layers), dim3(ENG_SCAN_SEG), 0, st, *P, nw);
    hipLaunchKernelGGL(k_layer_base, dim3(1), dim3(64), 0, st, *P, nseg);
  }
  if (ev) (void)hipEventRecord(ev[1], st);
  launch_estimator(P, S, V, n, count, n_cu, blocks, st, ev ? ev + 2 : nullptr);
  /* the interactions of the last order change only photon states that no
   * later step reads (the estimator has scored this order): not launched */
  if (n < P->n_int) launch_interact(P, S, n, (count + ENG_BLOCK - 1) / ENG_BLOCK, st);
  if (ev) (void)hipEventRecord(ev[6], st);
  return hipGetLastError();
}

} /* extern "C" */
