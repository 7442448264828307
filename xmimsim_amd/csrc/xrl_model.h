/*
 * xrl_model.h -- self-contained atomic-physics provider for the photon-history
 * engine.
 *
 * The reference engine evaluates xraylib at run time (CS_Total_Kissel,
 * CS_Photo_Partial, DCSP_Rayl, DCSP_Compt, FluorYield, RadRate, LineEnergy,
 * EdgeEnergy, CosKronTransProb and the P{L,M}x_*_cascade_kissel family; call
 * sites listed in src/xmi_variance_reduction.F90:344-649 and
 * src/xmi_main.F90:1564,2257,2272,5200-5291,5406).  xraylib and its data files
 * are not available offline, so this provider defines a closed, smooth,
 * physically shaped model of the same quantities:
 *
 *   - binding (edge) energies: power laws a*(Z-z0)^p per subshell, fitted to a
 *     handful of tabulated edges; a subshell exists when Z > z0;
 *   - line energy = E_edge(vacancy shell) - E_edge(donor shell);
 *   - photo-ionisation: every subshell s has tau_s(E) = B_s * E^-2.75 above its
 *     edge, with the jump ratios J_s fixing the split (nested construction, so
 *     that the total is continuous in log-log between edges);
 *   - Rayleigh / Compton: Thomson and Klein-Nishina differential cross sections
 *     times a model form factor F(x,Z) and incoherent scattering function
 *     S(x,Z); the integrated cross sections are tabulated per element on a
 *     uniform ln(E) grid and interpolated log-log;
 *   - fluorescence yields, Coster-Kronig probabilities, radiative rates
 *     (dipole rule weighted by donor occupancy), Auger vacancy transfer;
 *   - Compton profiles: sum over occupied subshells of hydrogenic 1s-shaped
 *     profiles with the subshell's characteristic momentum.
 *
 * All per-element parameters live in one POD struct (xrl_atom) so that the host
 * (oracle, table generator) and the device kernels evaluate exactly the same
 * formulas on exactly the same numbers.  The functions below are the only
 * place where the model is evaluated.
 */
#ifndef XMIMSIM_AMD_XRL_MODEL_H
#define XMIMSIM_AMD_XRL_MODEL_H

#include <math.h>

#if defined(__HIPCC__)
#define XRL_FN __host__ __device__ static inline
#else
#define XRL_FN static inline
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* xraylib shell numbering: K=0, L1..L3=1..3, M1..M5=4..8, N1..N7=9..15,
 * O1..O7=16..22, P1..P5=23..27, Q1..Q3=28..30 */
#define XRL_NSHELLS 31
#define XRL_K_SHELL 0
#define XRL_L1_SHELL 1
#define XRL_L2_SHELL 2
#define XRL_L3_SHELL 3
#define XRL_M1_SHELL 4
#define XRL_M5_SHELL 8

/* xraylib line numbering (absolute values of the *_LINE macros):
 * KL1=1 ... KP5=29, L1L2=30 ... L3Q1=113, M1M2=114 ... M5P5=219 ... P3P5=383 */
#define XRL_NLINES 383
#define XRL_KL1 1
#define XRL_KP5 29
#define XRL_L1_FIRST 30
#define XRL_L1_LAST 58
#define XRL_L2_FIRST 59
#define XRL_L2_LAST 85
#define XRL_L3_FIRST 86
#define XRL_L3_LAST 113
#define XRL_M1_FIRST 114
#define XRL_M1_LAST 136
#define XRL_M2_FIRST 137
#define XRL_M2_LAST 158
#define XRL_M3_FIRST 159
#define XRL_M3_LAST 180
#define XRL_M4_FIRST 181
#define XRL_M4_LAST 200
#define XRL_M5_FIRST 201
#define XRL_M5_LAST 219
#define XRL_L3Q1 113
#define XRL_M5P5 219

/* Coster-Kronig transitions, xraylib order: FL12, FL13, FLP13, FL23, FM12,
 * FM13, FM14, FM15, FM23, FM24, FM25, FM34, FM35, FM45 */
#define XRL_NCK 14
enum { XRL_FL12 = 0, XRL_FL13, XRL_FLP13, XRL_FL23, XRL_FM12, XRL_FM13, XRL_FM14,
       XRL_FM15, XRL_FM23, XRL_FM24, XRL_FM25, XRL_FM34, XRL_FM35, XRL_FM45 };

/* cascade types, numbering of src/xmi_aux_f.F90:662-665 */
#define XRL_CASCADE_NONE 1
#define XRL_CASCADE_NONRADIATIVE 2
#define XRL_CASCADE_RADIATIVE 3
#define XRL_CASCADE_FULL 4

/* integrated Rayleigh/Compton cross sections: uniform grid in ln(E) */
#define XRL_NCS 1024
#define XRL_CS_LNE0 (-4.605170185988091)  /* ln(0.01 keV) */
#define XRL_CS_LNE1 (6.907755278982137)   /* ln(1000 keV) */

#define XRL_PHOTO_EXPONENT 2.75
#define XRL_MEC2 510.998910              /* keV, value used in xmi_main.F90:2198 */
#define XRL_RE2 0.079407877              /* classical electron radius^2, barn */
#define XRL_AVOG_BARN 0.602214129        /* N_A * 1e-24 */
#define XRL_KEV_ANGSTROM 12.39842        /* hc in keV*Angstrom */

typedef struct xrl_atom {
  int Z;
  int n_occupied;                        /* number of occupied subshells */
  double A;                              /* g/mol */
  double edge[XRL_NSHELLS];              /* keV, 0 when absent */
  double occupancy[XRL_NSHELLS];         /* electrons */
  double fluor_yield[9];                 /* K..M5 */
  double ck[XRL_NCK];
  double rad_rate[XRL_NLINES + 1];       /* index |line|, [0] unused */
  double line_energy[XRL_NLINES + 1];    /* keV, 0 when the line is absent */
  double photo_B[9];                     /* tau_s(E) = photo_B[s]*E^-p, E>=edge */
  double photo_B_outer;                  /* all shells beyond M5 */
  double auger_KL[3];                    /* vacancies produced in L1..L3 per K Auger decay */
  double auger_KM[5];                    /* ... in M1..M5 per K Auger decay */
  double auger_LM[3][5];                 /* ... in M1..M5 per L1..L3 Auger decay */
  double ff_x0;                          /* Rayleigh form factor scale, 1/Angstrom */
  double sf_x0;                          /* incoherent scattering function scale */
  double profile_p[XRL_NSHELLS];         /* Compton profile momentum, atomic units */
  /* derived, for evaluation: with k = number of K..M5 edges above E,
   *   photo total  = E^-p * photo_Bsum[k]
   *   P_shell[s]   = E^-p * shell_C[cascade-1][k][s]   (P_K .. P_M5) */
  double photo_Bsum[10];
  double shell_C[4][10][9];
  double cs_rayl[XRL_NCS];               /* cm2/g on the ln(E) grid, linear interpolation */
  double cs_compt[XRL_NCS];
  double cs_rc[XRL_NCS];                 /* cs_rayl + cs_compt */
  double inv_ff_x0, inv_sf_x0;           /* 1 / ff_x0, 1 / sf_x0 */
  double barn_per_A;                     /* N_A * 1e-24 / A: barn/atom to cm2/g */
} xrl_atom;

/* per-energy quantities shared by every element evaluated at that energy */
typedef struct xrl_ectx {
  double E;
  double pe;   /* E^-p */
  double cf;   /* fraction within the ln(E) cell */
  int ci;      /* ln(E) cell */
} xrl_ectx;

/* --------------------------------------------------------------------------
 * Evaluation (host and device)
 * ------------------------------------------------------------------------*/

XRL_FN xrl_ectx xrl_energy(double E) {
  xrl_ectx c;
  c.E = E;
  if (!(E > 0.0)) { /* xraylib returns 0 for non-positive energies */
    c.pe = 0.0;
    c.cf = 0.0;
    c.ci = 0;
    return c;
  }
  /* one logarithm serves the grid position and E^-p = exp(-p ln E) */
  const double inv_h = (double)(XRL_NCS - 1) / (XRL_CS_LNE1 - XRL_CS_LNE0);
  const double lnE = log(E);
  double x = (lnE - XRL_CS_LNE0) * inv_h;
  int i = (int)floor(x);
  if (i < 0) i = 0;
  if (i > XRL_NCS - 2) i = XRL_NCS - 2;
  c.ci = i;
  c.cf = x - (double)i;
  c.pe = exp(-XRL_PHOTO_EXPONENT * lnE);
  return c;
}

/* number of K..M5 edges above E (edges decrease along K..M5; absent shells are 0) */
XRL_FN int xrl_edges_above(const xrl_atom *a, double E) {
  int k = 0;
  for (int s = 0; s < 9; s++) k += (a->edge[s] > E) ? 1 : 0;
  return k;
}

XRL_FN double xrl_table(const double *tab, const xrl_ectx *c) {
  if (!(c->E > 0.0)) return 0.0;
  return tab[c->ci] + c->cf * (tab[c->ci + 1] - tab[c->ci]);
}

XRL_FN double xrl_cs_rayl_c(const xrl_atom *a, const xrl_ectx *c) { return xrl_table(a->cs_rayl, c); }
XRL_FN double xrl_cs_compt_c(const xrl_atom *a, const xrl_ectx *c) { return xrl_table(a->cs_compt, c); }

/* total photo-ionisation cross section, cm2/g */
XRL_FN double xrl_cs_photo_total_c(const xrl_atom *a, const xrl_ectx *c) {
  if (!(c->E > 0.0)) return 0.0;
  return c->pe * a->photo_Bsum[xrl_edges_above(a, c->E)];
}

/* CS_Total_Kissel equivalent: photo + Rayleigh + Compton, cm2/g */
XRL_FN double xrl_cs_total_c(const xrl_atom *a, const xrl_ectx *c) {
  if (!(c->E > 0.0)) return 0.0;
  return c->pe * a->photo_Bsum[xrl_edges_above(a, c->E)] + xrl_table(a->cs_rc, c);
}

/* photo-ionisation partial cross section of shell s (0..8), cm2/g */
XRL_FN double xrl_cs_photo_partial_c(const xrl_atom *a, int s, const xrl_ectx *c) {
  if (s < 0 || s > 8) return 0.0;
  double e = a->edge[s];
  if (e <= 0.0 || !(c->E >= e)) return 0.0;
  return a->photo_B[s] * c->pe;
}

XRL_FN double xrl_cs_rayl(const xrl_atom *a, double E) {
  xrl_ectx c = xrl_energy(E);
  return xrl_cs_rayl_c(a, &c);
}
XRL_FN double xrl_cs_compt(const xrl_atom *a, double E) {
  xrl_ectx c = xrl_energy(E);
  return xrl_cs_compt_c(a, &c);
}
XRL_FN double xrl_cs_photo_total(const xrl_atom *a, double E) {
  xrl_ectx c = xrl_energy(E);
  return xrl_cs_photo_total_c(a, &c);
}
XRL_FN double xrl_cs_total(const xrl_atom *a, double E) {
  xrl_ectx c = xrl_energy(E);
  return xrl_cs_total_c(a, &c);
}
XRL_FN double xrl_cs_photo_partial(const xrl_atom *a, int s, double E) {
  xrl_ectx c = xrl_energy(E);
  return xrl_cs_photo_partial_c(a, s, &c);
}

/* momentum transfer x = sin(theta/2)/lambda, 1/Angstrom */
XRL_FN double xrl_moment_transf(double E, double theta) {
  return E / XRL_KEV_ANGSTROM * sin(theta * 0.5);
}

XRL_FN double xrl_ff_rayl(const xrl_atom *a, double x) {
  double u = x * a->inv_ff_x0;
  double d = 1.0 + 3.0 * u * u;
  return (double)a->Z / (d * d);
}

XRL_FN double xrl_sf_compt(const xrl_atom *a, double x) {
  double u = x * a->inv_sf_x0;
  double d = 1.0 + u * u;
  return (double)a->Z * (1.0 - 1.0 / (d * d));
}

/* unpolarised differential cross sections, cm2/g/sr */
XRL_FN double xrl_dcs_rayl(const xrl_atom *a, double E, double theta) {
  double c = cos(theta);
  double F = xrl_ff_rayl(a, xrl_moment_transf(E, theta));
  return XRL_RE2 * 0.5 * (1.0 + c * c) * F * F * a->barn_per_A;
}

XRL_FN double xrl_dcs_compt(const xrl_atom *a, double E, double theta) {
  double c = cos(theta);
  double k = 1.0 / (1.0 + E / XRL_MEC2 * (1.0 - c));
  double kn = XRL_RE2 * 0.5 * k * k * (k + 1.0 / k - (1.0 - c * c));
  return kn * xrl_sf_compt(a, xrl_moment_transf(E, theta)) * a->barn_per_A;
}

/* linearly polarised differential cross sections (DCSP_*), cm2/g/sr.  The
 * factors that depend on (E, theta, phi) only are formed once in an
 * xrl_dcsp_ctx and shared by every element scattering at that geometry. */
typedef struct xrl_dcsp_ctx {
  double x;       /* momentum transfer */
  double thomson; /* r_e^2 (1 - sin^2 theta cos^2 phi) */
  double kn;      /* polarised Klein-Nishina */
} xrl_dcsp_ctx;

XRL_FN xrl_dcsp_ctx xrl_dcsp_setup(double E, double theta, double phi) {
  xrl_dcsp_ctx d;
  double c = cos(theta), st = sin(theta), cp = cos(phi);
  double k = 1.0 / (1.0 + E / XRL_MEC2 * (1.0 - c));
  d.x = xrl_moment_transf(E, theta);
  d.thomson = XRL_RE2 * (1.0 - st * st * cp * cp);
  d.kn = XRL_RE2 * 0.5 * k * k * (k + 1.0 / k - 2.0 * st * st * cp * cp);
  return d;
}

XRL_FN double xrl_dcsp_rayl_c(const xrl_atom *a, const xrl_dcsp_ctx *d) {
  double F = xrl_ff_rayl(a, d->x);
  return d->thomson * F * F * a->barn_per_A;
}

XRL_FN double xrl_dcsp_compt_c(const xrl_atom *a, const xrl_dcsp_ctx *d) {
  return d->kn * xrl_sf_compt(a, d->x) * a->barn_per_A;
}

XRL_FN double xrl_dcsp_rayl(const xrl_atom *a, double E, double theta, double phi) {
  xrl_dcsp_ctx d = xrl_dcsp_setup(E, theta, phi);
  return xrl_dcsp_rayl_c(a, &d);
}

XRL_FN double xrl_dcsp_compt(const xrl_atom *a, double E, double theta, double phi) {
  xrl_dcsp_ctx d = xrl_dcsp_setup(E, theta, phi);
  return xrl_dcsp_compt_c(a, &d);
}

/* total Compton profile (electrons per atomic unit of momentum) */
XRL_FN double xrl_compton_profile_partial(const xrl_atom *a, int s, double pz) {
  if (a->occupancy[s] <= 0.0) return 0.0;
  double p = a->profile_p[s];
  double u = pz / p;
  double d = 1.0 + u * u;
  return a->occupancy[s] * 8.0 / (3.0 * 3.14159265358979323846 * p) / (d * d * d);
}

XRL_FN double xrl_compton_profile(const xrl_atom *a, double pz) {
  double sum = 0.0;
  for (int s = 0; s < XRL_NSHELLS; s++) sum += xrl_compton_profile_partial(a, s, pz);
  return sum;
}

/* points of the Compton-profile CDF grid that the detection-weight estimator
 * draws truncated profiles from (xmi_profile_cdf_grid) */
#define XMI_CP_GRID 4096

/* CDF of the hydrogenic profile shape: int_0^u (1 + t^2)^-3 dt */
XRL_FN double xrl_profile_G3(double u) {
  double d = 1.0 + u * u;
  return u / (4.0 * d * d) + 3.0 * u / (8.0 * d) + 0.375 * atan(u);
}

/* unnormalised CDF of the total Compton profile on [0, pz] */
XRL_FN double xrl_total_profile_cdf(const xrl_atom *a, double pz) {
  double s = 0.0;
  for (int k = 0; k < XRL_NSHELLS; k++)
    if (a->occupancy[k] > 0.0) s += a->occupancy[k] * xrl_profile_G3(pz / a->profile_p[k]);
  return s;
}

/* q in [lo, hi] with G3(q / p) = target, by 100 bisection steps */
XRL_FN double xrl_bisect_G3(double target, double p, double lo, double hi) {
  for (int it = 0; it < 100; it++) {
    double mid = 0.5 * (lo + hi);
    if (xrl_profile_G3(mid / p) < target) lo = mid;
    else hi = mid;
  }
  return 0.5 * (lo + hi);
}

/* the 100-step bisection of the total profile's inverse CDF at target */
XRL_FN double xrl_bisect_total_profile(const xrl_atom *a, double target, double lo, double hi) {
  for (int it = 0; it < 100; it++) {
    double mid = 0.5 * (lo + hi);
    if (xrl_total_profile_cdf(a, mid) < target) lo = mid;
    else hi = mid;
  }
  return 0.5 * (lo + hi);
}

/* azimuth inverse CDF of CDF(phi) = (phi - t sin 2 phi) / 2 pi at target * 2 pi,
 * by 100 bisection steps on [0, 2 pi] */
XRL_FN double xrl_bisect_phi(double t, double target) {
  double lo = 0.0, hi = 2.0 * 3.14159265358979323846;
  for (int it = 0; it < 100; it++) {
    double mid = 0.5 * (lo + hi);
    if (mid - t * sin(2.0 * mid) < target) lo = mid;
    else hi = mid;
  }
  return 0.5 * (lo + hi);
}

/* the angle inverse CDF of one (element, energy): cumulative trapezoid sum
 * of DCS(theta) sin(theta) on n_int points of [0, pi], inverted linearly at
 * rs[0..n_r-1]; computed in two passes (total, then the walk), so that no
 * CDF array is stored; row[0] = 0 and row[n_r-1] = pi */
XRL_FN void xrl_theta_icdf_row(const xrl_atom *a, int compton, double E, const double *rs, int n_r, int n_int,
                               double *row, long row_stride) {
  const double h = 3.14159265358979323846 / (n_int - 1);
  double total = 0.0, prev = 0.0;
  for (int k = 1; k < n_int; k++) {
    double th = k * h;
    double f = (compton ? xrl_dcs_compt(a, E, th) : xrl_dcs_rayl(a, E, th)) * sin(th);
    total = total + 0.5 * (f + prev) * h;
    prev = f;
  }
  /* walk: c0 = cdf[k - 1], c1 = cdf[k], advancing k while cdf[k] < target */
  int k = 1;
  double c0 = 0.0, f0 = 0.0;
  double f1 = (compton ? xrl_dcs_compt(a, E, h) : xrl_dcs_rayl(a, E, h)) * sin(h);
  double c1 = c0 + 0.5 * (f1 + f0) * h;
  for (int m = 0; m < n_r; m++) {
    double target = rs[m] * total;
    while (k < n_int - 1 && c1 < target) {
      k++;
      c0 = c1;
      f0 = f1;
      double th = k * h;
      f1 = (compton ? xrl_dcs_compt(a, E, th) : xrl_dcs_rayl(a, E, th)) * sin(th);
      c1 = c0 + 0.5 * (f1 + f0) * h;
    }
    double fr = (c1 > c0) ? (target - c0) / (c1 - c0) : 0.0;
    if (fr < 0.0) fr = 0.0;
    if (fr > 1.0) fr = 1.0;
    row[(long)m * row_stride] = ((k - 1) + fr) * h;
  }
  row[0] = 0.0;
  row[(long)(n_r - 1) * row_stride] = 3.14159265358979323846;
}

/* vacancy-shell -> first/last line, lines are 1-based |xraylib line| */
XRL_FN void xrl_shell_lines(int shell, int *first, int *last) {
  const int f[9] = {XRL_KL1, XRL_L1_FIRST, XRL_L2_FIRST, XRL_L3_FIRST, XRL_M1_FIRST,
                    XRL_M2_FIRST, XRL_M3_FIRST, XRL_M4_FIRST, XRL_M5_FIRST};
  const int l[9] = {XRL_KP5, XRL_L1_LAST, XRL_L2_LAST, XRL_L3_LAST, XRL_M1_LAST,
                    XRL_M2_LAST, XRL_M3_LAST, XRL_M4_LAST, XRL_M5_LAST};
  if (shell < 0 || shell > 8) { *first = 1; *last = 0; return; }
  *first = f[shell];
  *last = l[shell];
}

/* vacancy shell (0..8) of a line 1..219, -1 otherwise */
XRL_FN int xrl_line_shell(int line) {
  if (line >= XRL_KL1 && line <= XRL_KP5) return 0;
  if (line >= XRL_L1_FIRST && line <= XRL_L1_LAST) return 1;
  if (line >= XRL_L2_FIRST && line <= XRL_L2_LAST) return 2;
  if (line >= XRL_L3_FIRST && line <= XRL_L3_LAST) return 3;
  if (line >= XRL_M1_FIRST && line <= XRL_M1_LAST) return 4;
  if (line >= XRL_M2_FIRST && line <= XRL_M2_LAST) return 5;
  if (line >= XRL_M3_FIRST && line <= XRL_M3_LAST) return 6;
  if (line >= XRL_M4_FIRST && line <= XRL_M4_LAST) return 7;
  if (line >= XRL_M5_FIRST && line <= XRL_M5_LAST) return 8;
  return -1;
}

/* xraylib line numbers of the K->L_i, K->M_j, L_i->M_j radiative transitions */
XRL_FN int xrl_line_KL(int i) { return 1 + i; }          /* KL1..KL3 */
XRL_FN int xrl_line_KM(int j) { return 4 + j; }          /* KM1..KM5 */
XRL_FN int xrl_line_LM(int i, int j) {                   /* L1M1.., L2M1.., L3M1.. */
  const int base[3] = {32, 60, 86};
  return base[i] + j;
}

/*
 * Auger transitions of a K or L_i vacancy, the equivalent of xraylib's
 * AugerRate(Z, trans) over the transition list K_L1L1_AUGER .. L3_M5Q3_AUGER
 * that src/xmi_main.F90:2413-4594 walks.  Each transition t leaves two
 * vacancies (xrl_auger_shells); the model fills only the shells whose
 * vacancy transfer it defines (auger_KL, auger_KM, auger_LM), so its list is
 *   K   : t = 0 .. 63,            pairs (X, Y), X and Y in L1 L2 L3 M1 .. M5
 *   L_i : t = 64 + 30 i .. +29,   pairs (X, Y), X in M1 .. M5, Y in M1 .. M5
 *                                 or N1 (every shell beyond M5)
 * and its rates are those pairs' probabilities under the model's transfer:
 * a K decay puts each vacancy independently into L or M (the fractions of
 * auger_KL and auger_KM, split by occupancy); an L decay puts the first into
 * M and the second into M (0.6) or beyond (0.4).  The expected vacancies of
 * every shell per decay are then exactly auger_KL, auger_KM and auger_LM, so
 * explicit cascades reproduce the cascade-corrected cross sections.
 */
#define XRL_N1_SHELL 9
#define XRL_AUGER_N 154

/* first and last transition of a vacancy in shell; 0 when it has none */
XRL_FN int xrl_auger_range(int shell, int *first, int *last) {
  if (shell == XRL_K_SHELL) {
    *first = 0;
    *last = 63;
    return 1;
  }
  if (shell >= XRL_L1_SHELL && shell <= XRL_L3_SHELL) {
    *first = 64 + 30 * (shell - 1);
    *last = *first + 29;
    return 1;
  }
  return 0;
}

XRL_FN void xrl_auger_shells(int t, int *s1, int *s2) {
  if (t < 64) {
    *s1 = 1 + t / 8;
    *s2 = 1 + t % 8;
  } else {
    const int k = (t - 64) % 30;
    *s1 = XRL_M1_SHELL + k / 6;
    *s2 = k % 6 < 5 ? XRL_M1_SHELL + k % 6 : XRL_N1_SHELL;
  }
}

XRL_FN double xrl_auger_rate(const xrl_atom *a, int t) {
  if (t < 0 || t >= XRL_AUGER_N) return 0.0;
  double nL = a->occupancy[1] + a->occupancy[2] + a->occupancy[3], nM = 0.0;
  for (int j = 0; j < 5; j++) nM += a->occupancy[4 + j];
  int s1, s2;
  xrl_auger_shells(t, &s1, &s2);
  if (t < 64) {
    /* q(shell): the chance that one vacancy of a K decay lands there */
    const double fL = nM > 0.0 ? 0.8 : 1.0, fM = nL > 0.0 ? 0.2 : 1.0;
    double q[2];
    const int s[2] = {s1, s2};
    for (int k = 0; k < 2; k++) {
      if (s[k] <= XRL_L3_SHELL) q[k] = nL > 0.0 ? fL * a->occupancy[s[k]] / nL : 0.0;
      else q[k] = nM > 0.0 ? fM * a->occupancy[s[k]] / nM : 0.0;
    }
    return q[0] * q[1];
  }
  const int i = (t - 64) / 30;
  if (!(a->edge[1 + i] > 0.0) || !(nM > 0.0)) return 0.0;
  const double p1 = a->occupancy[s1] / nM;
  return s2 == XRL_N1_SHELL ? p1 * 0.4 : p1 * 0.6 * a->occupancy[s2] / nM;
}

/*
 * Cascade-corrected shell ionisation cross sections (P_K .. P_M5) at energy E,
 * the equivalent of the xraylib family P{L1..M5}_{pure,auger_cascade,
 * rad_cascade,full_cascade}_kissel used in src/xmi_variance_reduction.F90:443-572
 * and src/xmi_data_f.F90:1311-1474, as a linear map of the photo-ionisation
 * cross sections tau[K..M5].  Used once per element to fill shell_C.
 */
XRL_FN void xrl_shell_cross_sections_build(const xrl_atom *a, int cascade_type, const double tau[9], double P[9]) {
  const int rad = (cascade_type == XRL_CASCADE_RADIATIVE || cascade_type == XRL_CASCADE_FULL);
  const int aug = (cascade_type == XRL_CASCADE_NONRADIATIVE || cascade_type == XRL_CASCADE_FULL);
  double PK = tau[0];
  P[0] = PK;
  /* L shells: feeding from K vacancies, then L Coster-Kronig */
  double fromK[3];
  for (int i = 0; i < 3; i++) {
    double t = 0.0;
    if (rad) t += a->fluor_yield[0] * a->rad_rate[xrl_line_KL(i)];
    if (aug) t += (1.0 - a->fluor_yield[0]) * a->auger_KL[i];
    fromK[i] = PK * t;
  }
  double PL1 = tau[1] + fromK[0];
  double PL2 = tau[2] + fromK[1] + a->ck[XRL_FL12] * PL1;
  double PL3 = tau[3] + fromK[2] + a->ck[XRL_FL23] * PL2 + (a->ck[XRL_FL13] + a->ck[XRL_FLP13]) * PL1;
  P[1] = PL1;
  P[2] = PL2;
  P[3] = PL3;
  double PL[3] = {PL1, PL2, PL3};
  double feed[5];
  for (int j = 0; j < 5; j++) {
    double t = 0.0;
    if (rad) {
      t += PK * a->fluor_yield[0] * a->rad_rate[xrl_line_KM(j)];
      for (int i = 0; i < 3; i++) t += PL[i] * a->fluor_yield[1 + i] * a->rad_rate[xrl_line_LM(i, j)];
    }
    if (aug) {
      t += PK * (1.0 - a->fluor_yield[0]) * a->auger_KM[j];
      for (int i = 0; i < 3; i++) t += PL[i] * (1.0 - a->fluor_yield[1 + i]) * a->auger_LM[i][j];
    }
    feed[j] = t;
  }
  double PM1 = tau[4] + feed[0];
  double PM2 = tau[5] + feed[1] + a->ck[XRL_FM12] * PM1;
  double PM3 = tau[6] + feed[2] + a->ck[XRL_FM23] * PM2 + a->ck[XRL_FM13] * PM1;
  double PM4 = tau[7] + feed[3] + a->ck[XRL_FM34] * PM3 + a->ck[XRL_FM24] * PM2 + a->ck[XRL_FM14] * PM1;
  double PM5 = tau[8] + feed[4] + a->ck[XRL_FM45] * PM4 + a->ck[XRL_FM35] * PM3 + a->ck[XRL_FM25] * PM2 +
               a->ck[XRL_FM15] * PM1;
  P[4] = PM1;
  P[5] = PM2;
  P[6] = PM3;
  P[7] = PM4;
  P[8] = PM5;
}

/* P_K .. P_M5 at energy E for a cascade type (1..4): E^-p * shell_C[k] */
XRL_FN void xrl_shell_cross_sections_c(const xrl_atom *a, int cascade_type, const xrl_ectx *c, double P[9]) {
  if (!(c->E > 0.0)) {
    for (int s = 0; s < 9; s++) P[s] = 0.0;
    return;
  }
  const double *C = a->shell_C[cascade_type - 1][xrl_edges_above(a, c->E)];
  for (int s = 0; s < 9; s++) P[s] = c->pe * C[s];
}

XRL_FN void xrl_shell_cross_sections(const xrl_atom *a, int cascade_type, double E, double P[9]) {
  xrl_ectx c = xrl_energy(E);
  xrl_shell_cross_sections_c(a, cascade_type, &c, P);
}

/* --------------------------------------------------------------------------
 * Construction (host only, xrl_model.c)
 * ------------------------------------------------------------------------*/
/* fill *a for element Z (1..94); returns 1 on success */
int xrl_atom_init(xrl_atom *a, int Z);
/* process-wide cache of all 94 elements; never freed */
const xrl_atom *xrl_atom_get(int Z);
void xrl_atoms_preload(const int *Zs, int n);
/* xraylib line names, index 1..383 (0 = ""), 384 = "Rayleigh", 385 = "Compton" */
const char *xrl_line_name(int line);
/* donor/vacancy shells of a line (-1 for composite lines such as KO, L1N67) */
int xrl_line_shells(int line, int *vacancy, int *donor);
/* fluorescence yields corrected for Coster-Kronig transfer (src/xmi_data_f.F90:1244-1305) */
void xrl_fluor_yields_corr(const xrl_atom *a, double out[9]);

#ifdef __cplusplus
}
#endif
#endif
