/*
 * xrl_model.c -- construction of the per-element parameters of the analytic
 * atomic-physics provider (see xrl_model.h for the model).  Host only.
 *
 * Everything here is deterministic and depends on Z only, so the same numbers
 * are produced on every machine; the table generator (xmi_db.c), the C oracle
 * and the device kernels all read the resulting xrl_atom structs.
 */
#include "xrl_model.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* quantum numbers of the 31 xraylib subshells: n, l, 2j */
static const int shell_n[XRL_NSHELLS] = {1, 2, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4, 4, 4, 4, 4,
                                         5, 5, 5, 5, 5, 5, 5, 6, 6, 6, 6, 6, 7, 7, 7};
static const int shell_l[XRL_NSHELLS] = {0, 0, 1, 1, 0, 1, 1, 2, 2, 0, 1, 1, 2, 2, 3, 3,
                                         0, 1, 1, 2, 2, 3, 3, 0, 1, 1, 2, 2, 0, 1, 1};
static const int shell_2j[XRL_NSHELLS] = {1, 1, 1, 3, 1, 1, 3, 3, 5, 1, 1, 3, 3, 5, 5, 7,
                                          1, 1, 3, 3, 5, 5, 7, 1, 1, 3, 3, 5, 1, 1, 3};
static const char *shell_names[XRL_NSHELLS] = {
    "K",  "L1", "L2", "L3", "M1", "M2", "M3", "M4", "M5", "N1", "N2",
    "N3", "N4", "N5", "N6", "N7", "O1", "O2", "O3", "O4", "O5", "O6",
    "O7", "P1", "P2", "P3", "P4", "P5", "Q1", "Q2", "Q3"};

/* binding energy model a*(Z-z0)^p (keV), fitted to tabulated K..O5 edges of
 * representative elements; the shells beyond O5 use small ad-hoc values */
static const double edge_a[XRL_NSHELLS] = {
    6.608549e-03, 4.355522e-04, 4.460712e-04, 6.034148e-04, 8.869090e-05, 4.759895e-05,
    1.210806e-04, 2.073429e-04, 2.233069e-04, 1.113278e-05, 2.149831e-04, 3.341717e-04,
    4.704660e-04, 4.908485e-04, 4.471002e-06, 3.395162e-06, 1.257023e-04, 1.190179e-03,
    9.170210e-04, 3.349160e-05, 2.695161e-05, 1.0e-06,      1.0e-06,      5.0e-05,
    5.0e-05,      5.0e-05,      5.0e-05,      5.0e-05,      1.0e-05,      1.0e-05,
    1.0e-05};
static const double edge_p[XRL_NSHELLS] = {
    2.1561, 2.3981, 2.3984, 2.2971, 2.5029, 2.6531, 2.3926, 2.2924, 2.2647, 2.7255, 2.0784,
    1.9232, 1.8143, 1.7889, 3.2221, 3.2943, 1.8572, 1.2789, 1.2850, 2.6233, 2.6608, 3.0,
    3.0,    2.0,    1.8,    1.8,    1.8,    1.8,    1.5,    1.5,    1.5};

/* ------------------------------------------------------------------------ */
/* line table: xraylib line numbering 1..383 built from the vacancy shell and
 * the list of donor shells; composite lines (KO, KP, L1N67, ...) get donor -1 */
typedef struct {
  char name[10];
  int vacancy, donor;
} line_def;

static line_def g_lines[XRL_NLINES + 3];
static int g_lines_ready = 0;

static int shell_index(const char *s) {
  for (int i = 0; i < XRL_NSHELLS; i++)
    if (strcmp(shell_names[i], s) == 0) return i;
  return -1;
}

static void add_line(int *n, int vacancy, const char *donor) {
  line_def *d = &g_lines[*n];
  snprintf(d->name, sizeof(d->name), "%s%s", shell_names[vacancy], donor);
  d->vacancy = vacancy;
  d->donor = shell_index(donor);
  (*n)++;
}

static void add_range(int *n, int vacancy, char shell, int from, int to) {
  char buf[16];
  for (int i = from; i <= to; i++) {
    snprintf(buf, sizeof(buf), "%c%d", shell, i);
    add_line(n, vacancy, buf);
  }
}

static void build_lines(void) {
  int n = 1;
  strcpy(g_lines[0].name, "");
  g_lines[0].vacancy = g_lines[0].donor = -1;
  /* K */
  add_range(&n, 0, 'L', 1, 3);
  add_range(&n, 0, 'M', 1, 5);
  add_range(&n, 0, 'N', 1, 7);
  add_line(&n, 0, "O");
  add_range(&n, 0, 'O', 1, 7);
  add_line(&n, 0, "P");
  add_range(&n, 0, 'P', 1, 5);
  /* L1 */
  add_range(&n, 1, 'L', 2, 3);
  add_range(&n, 1, 'M', 1, 5);
  add_range(&n, 1, 'N', 1, 6);
  add_line(&n, 1, "N67");
  add_line(&n, 1, "N7");
  add_range(&n, 1, 'O', 1, 4);
  add_line(&n, 1, "O45");
  add_range(&n, 1, 'O', 5, 7);
  add_range(&n, 1, 'P', 1, 2);
  add_line(&n, 1, "P23");
  add_range(&n, 1, 'P', 3, 5);
  /* L2 */
  add_line(&n, 2, "L3");
  add_range(&n, 2, 'M', 1, 5);
  add_range(&n, 2, 'N', 1, 7);
  add_range(&n, 2, 'O', 1, 7);
  add_range(&n, 2, 'P', 1, 2);
  add_line(&n, 2, "P23");
  add_range(&n, 2, 'P', 3, 5);
  add_line(&n, 2, "Q1");
  /* L3 */
  add_range(&n, 3, 'M', 1, 5);
  add_range(&n, 3, 'N', 1, 7);
  add_range(&n, 3, 'O', 1, 4);
  add_line(&n, 3, "O45");
  add_range(&n, 3, 'O', 5, 7);
  add_range(&n, 3, 'P', 1, 2);
  add_line(&n, 3, "P23");
  add_range(&n, 3, 'P', 3, 4);
  add_line(&n, 3, "P45");
  add_line(&n, 3, "P5");
  add_line(&n, 3, "Q1");
  /* M1..M5 */
  add_range(&n, 4, 'M', 2, 5);
  add_range(&n, 4, 'N', 1, 7);
  add_range(&n, 4, 'O', 1, 7);
  add_range(&n, 4, 'P', 1, 5);
  add_range(&n, 5, 'M', 3, 5);
  add_range(&n, 5, 'N', 1, 7);
  add_range(&n, 5, 'O', 1, 7);
  add_range(&n, 5, 'P', 1, 5);
  add_range(&n, 6, 'M', 4, 5);
  add_range(&n, 6, 'N', 1, 7);
  add_range(&n, 6, 'O', 1, 7);
  add_range(&n, 6, 'P', 1, 5);
  add_line(&n, 6, "Q1");
  add_line(&n, 7, "M5");
  add_range(&n, 7, 'N', 1, 7);
  add_range(&n, 7, 'O', 1, 7);
  add_range(&n, 7, 'P', 1, 5);
  add_range(&n, 8, 'N', 1, 7);
  add_range(&n, 8, 'O', 1, 7);
  add_range(&n, 8, 'P', 1, 5);
  /* N1..N7 */
  for (int v = 9; v <= 15; v++) {
    add_range(&n, v, 'N', v - 9 + 2, 7);
    add_range(&n, v, 'O', 1, 7);
    add_range(&n, v, 'P', 1, 5);
  }
  /* O1..O5 */
  for (int v = 16; v <= 20; v++) {
    add_range(&n, v, 'O', v - 16 + 2, 7);
    add_range(&n, v, 'P', 1, 5);
  }
  /* O6, O7 */
  add_line(&n, 21, "O7");
  add_range(&n, 21, 'P', 4, 5);
  add_range(&n, 22, 'P', 4, 5);
  /* P1..P3 */
  add_range(&n, 23, 'P', 2, 5);
  add_range(&n, 24, 'P', 3, 5);
  add_range(&n, 25, 'P', 4, 5);
  if (n != XRL_NLINES + 1) {
    fprintf(stderr, "xrl_model: line table has %d entries, expected %d\n", n - 1, XRL_NLINES);
    abort();
  }
  strcpy(g_lines[384].name, "Rayleigh");
  strcpy(g_lines[385].name, "Compton");
  g_lines[384].vacancy = g_lines[384].donor = -1;
  g_lines[385].vacancy = g_lines[385].donor = -1;
}

static pthread_once_t lines_once = PTHREAD_ONCE_INIT;
static void lines_init(void) {
  build_lines();
  g_lines_ready = 1;
}

const char *xrl_line_name(int line) {
  pthread_once(&lines_once, lines_init);
  if (line < 0 || line > 385) return NULL;
  return g_lines[line].name;
}

int xrl_line_shells(int line, int *vacancy, int *donor) {
  pthread_once(&lines_once, lines_init);
  if (line < 1 || line > XRL_NLINES) return 0;
  *vacancy = g_lines[line].vacancy;
  *donor = g_lines[line].donor;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* ground-state occupancy by Madelung filling of the (n,l,j) subshells */
static void fill_occupancy(int Z, double occ[XRL_NSHELLS]) {
  /* filling order by (n+l, n) then j: listed as shell indices */
  static const int order[] = {0,  1,  2,  3,  4,  5,  6,  9,  7,  8,  10, 11, 16,
                              12, 13, 17, 18, 23, 14, 15, 19, 20, 24, 25, 28, 21,
                              22, 26, 27, 29, 30};
  memset(occ, 0, sizeof(double) * XRL_NSHELLS);
  int left = Z;
  for (size_t k = 0; k < sizeof(order) / sizeof(order[0]) && left > 0; k++) {
    int s = order[k];
    int cap = shell_2j[s] + 1;
    int take = left < cap ? left : cap;
    occ[s] = take;
    left -= take;
  }
}

static double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

/* integrated cross section by composite Simpson rule in theta */
static double integrate_dcs(const xrl_atom *a, double E, int compton) {
  const int n = 2000; /* even */
  const double h = M_PI / n;
  double sum = 0.0;
  for (int i = 0; i <= n; i++) {
    double th = i * h;
    double f = (compton ? xrl_dcs_compt(a, E, th) : xrl_dcs_rayl(a, E, th)) * sin(th);
    double w = (i == 0 || i == n) ? 1.0 : ((i & 1) ? 4.0 : 2.0);
    sum += w * f;
  }
  return 2.0 * M_PI * sum * h / 3.0;
}

int xrl_atom_init(xrl_atom *a, int Z) {
  if (Z < 1 || Z > 94) return 0;
  pthread_once(&lines_once, lines_init);
  memset(a, 0, sizeof(*a));
  a->Z = Z;
  a->A = (Z == 1) ? 1.008 : Z * (2.0 + 0.0065 * Z);

  /* occupancy and binding energies */
  fill_occupancy(Z, a->occupancy);
  for (int s = 0; s < XRL_NSHELLS; s++) {
    if (a->occupancy[s] <= 0.0) continue;
    a->n_occupied++;
    /* threshold of the power law: the element before the first one that fills
     * the subshell, so the binding energy starts from zero */
    int zfirst = Z;
    for (int zz = 1; zz <= Z; zz++) {
      double occ[XRL_NSHELLS];
      fill_occupancy(zz, occ);
      if (occ[s] > 0.0) { zfirst = zz; break; }
    }
    double z0 = zfirst - 0.5;
    a->edge[s] = edge_a[s] * pow(Z - z0, edge_p[s]);
  }
  /* keep binding energies strictly decreasing along the xraylib shell order
   * within K..M5 so that line energies of the main series stay positive */
  for (int s = 1; s < 9; s++)
    if (a->edge[s] > 0.0 && a->edge[s - 1] > 0.0 && a->edge[s] >= a->edge[s - 1])
      a->edge[s] = 0.98 * a->edge[s - 1];

  /* fluorescence yields */
  double z4 = pow((double)Z, 4.0);
  double wK = z4 / (z4 + 1.0e6);
  double wL3 = z4 / (z4 + 8.9e7);
  double wM = z4 / (z4 + 1.2e9);
  double fy[9] = {wK, 0.3 * wL3, 0.95 * wL3, wL3, 0.5 * wM, 0.5 * wM, 0.6 * wM, wM, wM};
  for (int s = 0; s < 9; s++) a->fluor_yield[s] = (a->edge[s] > 0.0) ? fy[s] : 0.0;

  /* Coster-Kronig */
  double t = clamp01((Z - 20.0) / 70.0);
  double ck[XRL_NCK] = {0.05 + 0.15 * t, 0.25 + 0.35 * t, 0.0, 0.10 + 0.05 * t, 0.05, 0.10, 0.08,
                        0.10, 0.08, 0.05, 0.05, 0.06, 0.10, 0.06};
  /* transition (from,to) shells */
  const int ck_from[XRL_NCK] = {1, 1, 1, 2, 4, 4, 4, 4, 5, 5, 5, 6, 6, 7};
  const int ck_to[XRL_NCK] = {2, 3, 3, 3, 5, 6, 7, 8, 6, 7, 8, 7, 8, 8};
  for (int k = 0; k < XRL_NCK; k++)
    a->ck[k] = (a->edge[ck_from[k]] > 0.0 && a->edge[ck_to[k]] > 0.0) ? ck[k] : 0.0;

  /* line energies and radiative rates (dipole rule, donor occupancy weight) */
  for (int line = 1; line <= XRL_NLINES; line++) {
    int v = g_lines[line].vacancy, d = g_lines[line].donor;
    if (d < 0 || a->edge[v] <= 0.0 || a->edge[d] <= 0.0) continue;
    double e = a->edge[v] - a->edge[d];
    if (e <= 0.0) continue;
    a->line_energy[line] = e;
  }
  for (int v = 0; v < 9; v++) {
    int first, last;
    xrl_shell_lines(v, &first, &last);
    double sum = 0.0;
    for (int line = first; line <= last; line++) {
      int d = g_lines[line].donor;
      if (d < 0 || a->line_energy[line] <= 0.0) continue;
      int dl = shell_l[d] - shell_l[v];
      int dj = shell_2j[d] - shell_2j[v];
      if ((dl != 1 && dl != -1) || dj > 2 || dj < -2) continue;
      int dn = shell_n[d] - shell_n[v];
      static const double g[5] = {0.002, 1.0, 0.17, 0.04, 0.01};
      double w = a->occupancy[d] * g[dn < 0 ? 0 : (dn > 4 ? 4 : dn)];
      a->rad_rate[line] = w;
      sum += w;
    }
    if (sum > 0.0)
      for (int line = first; line <= last; line++) a->rad_rate[line] /= sum;
  }

  /* Auger vacancy transfer: two vacancies per non-radiative decay, 80% into the
   * next main shell and 20% into the one after, split by occupancy */
  {
    double nL = a->occupancy[1] + a->occupancy[2] + a->occupancy[3];
    double nM = 0.0;
    for (int j = 0; j < 5; j++) nM += a->occupancy[4 + j];
    double fracL = (nM > 0.0) ? 0.8 : 1.0;
    if (nL > 0.0)
      for (int i = 0; i < 3; i++) a->auger_KL[i] = 2.0 * fracL * a->occupancy[1 + i] / nL;
    if (nM > 0.0) {
      double fracM = (nL > 0.0) ? 0.2 : 1.0;
      for (int j = 0; j < 5; j++) a->auger_KM[j] = 2.0 * fracM * a->occupancy[4 + j] / nM;
      for (int i = 0; i < 3; i++)
        if (a->edge[1 + i] > 0.0)
          for (int j = 0; j < 5; j++) a->auger_LM[i][j] = 1.6 * a->occupancy[4 + j] / nM;
    }
  }

  /* photo-ionisation: nested jump-ratio construction */
  {
    const double p = XRL_PHOTO_EXPONENT;
    double jump[9] = {125.0 / Z + 3.5, 1.16, 1.41, 80.0 / Z + 1.9, 1.04, 1.05, 1.10, 1.30, 1.60};
    /* total photo cross section just above the K edge (cm2/g) */
    double eK = a->edge[0];
    double above = 3190.0 * pow(Z / 14.0, -3.41);
    double e_prev = eK;
    for (int s = 0; s < 9; s++) {
      double e = a->edge[s];
      if (e <= 0.0) continue;
      if (s > 0) above *= pow(e_prev / e, p);
      double part = above * (1.0 - 1.0 / jump[s]);
      a->photo_B[s] = part * pow(e, p);
      above /= jump[s];
      e_prev = e;
    }
    a->photo_B_outer = above * pow(e_prev, p);
  }

  /* form factor / incoherent scattering function scales */
  a->ff_x0 = 0.485 * cbrt((double)Z);
  a->sf_x0 = 0.35 * cbrt((double)Z);
  a->inv_ff_x0 = 1.0 / a->ff_x0;
  a->inv_sf_x0 = 1.0 / a->sf_x0;
  a->barn_per_A = XRL_AVOG_BARN / a->A;

  /* Compton profile momenta from binding energies (hydrogenic: E = p^2/2) */
  for (int s = 0; s < XRL_NSHELLS; s++) {
    if (a->occupancy[s] <= 0.0) continue;
    double eb = a->edge[s] > 0.0 ? a->edge[s] : 0.005;
    double hartree = eb / 0.0272113862;
    double pp = sqrt(2.0 * hartree);
    a->profile_p[s] = pp < 0.3 ? 0.3 : pp;
  }

  /* integrated Rayleigh/Compton cross sections */
  const double h = (XRL_CS_LNE1 - XRL_CS_LNE0) / (double)(XRL_NCS - 1);
  for (int i = 0; i < XRL_NCS; i++) {
    double E = exp(XRL_CS_LNE0 + i * h);
    a->cs_rayl[i] = integrate_dcs(a, E, 0);
    a->cs_compt[i] = integrate_dcs(a, E, 1);
    a->cs_rc[i] = a->cs_rayl[i] + a->cs_compt[i];
  }

  /* photo cross section and cascade-corrected shell cross sections as
   * functions of the number k of K..M5 edges above the energy */
  for (int k = 0; k <= 9; k++) {
    double tau[9], sum = a->photo_B_outer;
    for (int s2 = 0; s2 < 9; s2++) {
      tau[s2] = (s2 >= k && a->edge[s2] > 0.0) ? a->photo_B[s2] : 0.0;
      sum += tau[s2];
    }
    a->photo_Bsum[k] = sum;
    for (int c = 0; c < 4; c++) xrl_shell_cross_sections_build(a, c + 1, tau, a->shell_C[c][k]);
  }
  return 1;
}

static xrl_atom *g_atoms[95];
static pthread_mutex_t g_atoms_lock = PTHREAD_MUTEX_INITIALIZER;

const xrl_atom *xrl_atom_get(int Z) {
  if (Z < 1 || Z > 94) return NULL;
  pthread_mutex_lock(&g_atoms_lock);
  const xrl_atom *r = g_atoms[Z];
  pthread_mutex_unlock(&g_atoms_lock);
  if (r) return r;
  /* built outside the lock, so that different elements build in parallel;
   * the first record published wins */
  xrl_atom *a = (xrl_atom *)malloc(sizeof(xrl_atom));
  if (!a || !xrl_atom_init(a, Z)) {
    free(a);
    return NULL;
  }
  pthread_mutex_lock(&g_atoms_lock);
  if (!g_atoms[Z]) g_atoms[Z] = a;
  else free(a);
  r = g_atoms[Z];
  pthread_mutex_unlock(&g_atoms_lock);
  return r;
}

/* builds the records of Zs[0..n-1] on all host threads */
void xrl_atoms_preload(const int *Zs, int n) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < n; i++) (void)xrl_atom_get(Zs[i]);
}

/* corrected fluorescence yields: probability that a primary vacancy in shell s
 * ends, after Coster-Kronig transfer, in a radiative decay of some shell.
 * Same combinatorics as src/xmi_data_f.F90:1244-1305. */
void xrl_fluor_yields_corr(const xrl_atom *a, double out[9]) {
  const double *w = a->fluor_yield;
  const double *f = a->ck;
  out[0] = w[0];
  out[1] = w[1] + f[XRL_FL12] * w[2] + (f[XRL_FL13] + f[XRL_FL12] * f[XRL_FL23]) * w[3];
  out[2] = w[2] + f[XRL_FL23] * w[3];
  out[3] = w[3];
  double m12 = f[XRL_FM12], m13 = f[XRL_FM13], m14 = f[XRL_FM14], m15 = f[XRL_FM15];
  double m23 = f[XRL_FM23], m24 = f[XRL_FM24], m25 = f[XRL_FM25];
  double m34 = f[XRL_FM34], m35 = f[XRL_FM35], m45 = f[XRL_FM45];
  out[4] = w[4] + m12 * w[5] + (m13 + m12 * m23) * w[6] +
           (m14 + m13 * m34 + m12 * m24 + m12 * m23 * m34) * w[7] +
           (m15 + m14 * m45 + m13 * m35 + m12 * m25 + m13 * m34 * m45 + m12 * m24 * m45 +
            m12 * m23 * m35 + m12 * m23 * m34 * m45) * w[8];
  out[5] = w[5] + m23 * w[6] + (m24 + m23 * m34) * w[7] +
           (m25 + m24 * m45 + m23 * m35 + m23 * m34 * m45) * w[8];
  out[6] = w[6] + m34 * w[7] + (m35 + m34 * m45) * w[8];
  out[7] = w[7] + m45 * w[8];
  out[8] = w[8];
}
