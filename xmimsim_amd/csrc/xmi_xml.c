/*
 * xmi_xml.c -- the XMSI input reader/writer and the XMSO output writer/reader
 * on either side of the photon-history engine, with libxml2's tree API.
 *
 * Reference interfaces replaced (XMI-MSIM 8.1):
 *   xmi_input_read_from_xml_file      src/xmi_xml.c:1286 (sections :364-1284)
 *   xmi_input_read_from_xml_string    src/xmi_xml.c:1814
 *   xmi_input_write_to_xml_file       src/xmi_xml.c:1404 (body :1679-1812, layers :2553-2576)
 *   xmi_input_write_to_xml_string     src/xmi_xml.c:1363
 *   xmi_output_raw2struct             src/xmi_data_structs.c:1371-1519
 *   xmi_output_write_to_xml_file      src/xmi_xml.c:1638 (body :1453-1636, graphs :1880-2031)
 *   xmi_output_read_from_xml_file     src/xmi_xml.c:2033 (body :2158-2308, :165-362)
 *   xmi_input_validate                src/xmi_data_structs.c (rules per section)
 *   xmi_input_free, xmi_output_free   src/xmi_data_structs.c
 *
 * Element and attribute names, their order, number formats (%g, %i, %li) and
 * the semantics of the readers (weights normalised and sorted by Z, zero
 * weights dropped, duplicate discrete energies merged, orientations
 * normalised, nchannels defaulting to 2048) follow the file format of
 * xml/xmimsim-1.0.dtd.  Documents are not validated against the DTD (the
 * reference resolves it through an XML catalog); instead every section that
 * the DTD requires must be present, and the input passes xmi_input_validate.
 * Errors are reported through an xmi_error whose leading members have the
 * layout of GError (domain, code, message), and also as the thread's last
 * error message.
 */
#include <ctype.h>
#include <math.h>
#include <pwd.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <libxml/parser.h>
#include <libxml/tree.h>

#include "xmi_internal.h"

#define XMI_XML_VERSION "8.1"
#define XMI_DTD_URL "http://www.xmi.UGent.be/xml/xmimsim-1.0.dtd"
#define XMI_XML_ERROR_DOMAIN 0x584d4c31u /* "XML1" */
#define SVG_BOX_WIDTH 500
#define SVG_BOX_HEIGHT 250

/* ---- errors -------------------------------------------------------------- */

static int fail(xmi_error **error, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  xmi_set_error("%s", buf);
  if (error && !*error) {
    xmi_error *e = (xmi_error *)malloc(sizeof(xmi_error));
    if (e) {
      e->domain = XMI_XML_ERROR_DOMAIN;
      e->code = 1;
      e->message = strdup(buf);
    }
    *error = e;
  }
  return 0;
}

void xmi_error_free(xmi_error *error) {
  if (!error) return;
  free(error->message);
  free(error);
}

/* ---- element symbols ------------------------------------------------------ */

static const char *const k_symbols[] = {
    "",   "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",
    "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As",
    "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn",
    "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho",
    "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po",
    "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md",
    "No", "Lr"};

const char *xmi_element_symbol(int Z) {
  return Z >= 1 && Z < (int)(sizeof k_symbols / sizeof k_symbols[0]) ? k_symbols[Z] : NULL;
}

/* ---- reading helpers ------------------------------------------------------ */

static int is(xmlNodePtr n, const char *name) { return n && !xmlStrcmp(n->name, (const xmlChar *)name); }

/* text content of an element (malloc'ed, "" when empty) */
static char *text_of(xmlNodePtr n) {
  xmlChar *t = xmlNodeGetContent(n);
  char *s = strdup(t ? (const char *)t : "");
  if (t) xmlFree(t);
  return s;
}

static int scan_text(xmlNodePtr n, const char *fmt, void *out, const char *what, xmi_error **error) {
  char *s = text_of(n);
  const int ok = sscanf(s, fmt, out) == 1;
  free(s);
  return ok ? 1 : fail(error, "error reading in %s of xml file", what);
}
#define RD_DOUBLE(n, p, what) scan_text(n, "%lg", (void *)(p), what, error)
#define RD_INT(n, p, what) scan_text(n, "%i", (void *)(p), what, error)
#define RD_LONG(n, p, what) scan_text(n, "%li", (void *)(p), what, error)

static char *attr_of(xmlNodePtr n, const char *name) {
  xmlChar *v = xmlGetProp(n, (const xmlChar *)name);
  if (!v) return NULL;
  char *s = strdup((const char *)v);
  xmlFree(v);
  return s;
}

static xmlNodePtr first_el(xmlNodePtr n) { return n ? xmlFirstElementChild(n) : NULL; }
static xmlNodePtr next_el(xmlNodePtr n) { return xmlNextElementSibling(n); }

static int rd_xyz(xmlNodePtr n, double *v, const char *what, xmi_error **error) {
  char w[96];
  for (xmlNodePtr c = first_el(n); c; c = next_el(c)) {
    int k = is(c, "x") ? 0 : is(c, "y") ? 1 : is(c, "z") ? 2 : -1;
    if (k < 0) continue;
    snprintf(w, sizeof w, "%s %c", what, "xyz"[k]);
    if (!RD_DOUBLE(c, &v[k], w)) return 0;
  }
  return 1;
}

static void normalize3(double *v) {
  const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (n > 0.0)
    for (int k = 0; k < 3; k++) v[k] /= n;
}

/* ---- freeing -------------------------------------------------------------- */

static void free_layers(xmi_layer *l, int n) {
  if (!l) return;
  for (int i = 0; i < n; i++) {
    free(l[i].Z);
    free(l[i].weight);
  }
  free(l);
}

void xmi_input_free(xmi_input *input) {
  if (!input) return;
  if (input->general) {
    free(input->general->outputfile);
    free(input->general->comments);
    free(input->general);
  }
  if (input->composition) {
    free_layers(input->composition->layers, input->composition->n_layers);
    free(input->composition);
  }
  free(input->geometry);
  if (input->excitation) {
    free(input->excitation->discrete);
    free(input->excitation->continuous);
    free(input->excitation);
  }
  if (input->absorbers) {
    free_layers(input->absorbers->exc_layers, input->absorbers->n_exc_layers);
    free_layers(input->absorbers->det_layers, input->absorbers->n_det_layers);
    free(input->absorbers);
  }
  if (input->detector) {
    free_layers(input->detector->crystal_layers, input->detector->n_crystal_layers);
    free(input->detector);
  }
  free(input);
}

static void free_history(xmi_fluorescence_line_counts *h, int n) {
  if (!h) return;
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < h[i].n_lines; j++) {
      free(h[i].lines[j].line_type);
      free(h[i].lines[j].interactions);
    }
    free(h[i].lines);
  }
  free(h);
}

void xmi_output_free(xmi_output *output) {
  if (!output) return;
  free(output->inputfile);
  free(output->outputfile);
  free_history(output->brute_force_history, output->nbrute_force_history);
  free_history(output->var_red_history, output->nvar_red_history);
  for (int i = 0; i <= output->ninteractions; i++) {
    if (output->channels_conv) free(output->channels_conv[i]);
    if (output->channels_unconv) free(output->channels_unconv[i]);
  }
  free(output->channels_conv);
  free(output->channels_unconv);
  xmi_input_free(output->input);
  free(output);
}

/* ---- input sections (src/xmi_xml.c:364-1284) ------------------------------ */

/* a layer: elements with a zero weight are dropped, the rest sorted by Z
 * and their weights normalised to a sum of 1 (src/xmi_xml.c:1174-1284) */
static int rd_layer(xmlNodePtr node, xmi_layer *layer, xmi_error **error) {
  int n = 0, cap = 0, *Z = NULL;
  double *w = NULL;
  memset(layer, 0, sizeof(*layer));
  for (xmlNodePtr c = first_el(node); c; c = next_el(c)) {
    if (is(c, "element")) {
      int z = 0;
      double wt = 0.0;
      for (xmlNodePtr d = first_el(c); d; d = next_el(d)) {
        if (is(d, "atomic_number") && !RD_INT(d, &z, "atomic_number")) goto bad;
        if (is(d, "weight_fraction") && !RD_DOUBLE(d, &wt, "weight_fraction")) goto bad;
      }
      if (fabs(wt) < 1e-20) continue;
      if (n == cap) {
        cap = cap ? 2 * cap : 8;
        Z = (int *)realloc(Z, sizeof(int) * cap);
        w = (double *)realloc(w, sizeof(double) * cap);
      }
      Z[n] = z;
      w[n] = wt;
      n++;
    } else if (is(c, "density")) {
      if (!RD_DOUBLE(c, &layer->density, "density")) goto bad;
    } else if (is(c, "thickness")) {
      if (!RD_DOUBLE(c, &layer->thickness, "thickness")) goto bad;
    }
  }
  /* insertion sort by Z, stable */
  for (int i = 1; i < n; i++) {
    int z = Z[i];
    double wt = w[i];
    int j = i - 1;
    for (; j >= 0 && Z[j] > z; j--) {
      Z[j + 1] = Z[j];
      w[j + 1] = w[j];
    }
    Z[j + 1] = z;
    w[j + 1] = wt;
  }
  double sum = 0.0;
  for (int i = 0; i < n; i++) sum += w[i];
  for (int i = 0; i < n; i++) w[i] /= sum;
  layer->n_elements = n;
  layer->Z = Z;
  layer->weight = w;
  return 1;
bad:
  free(Z);
  free(w);
  return 0;
}

static int rd_layers(xmlNodePtr node, xmi_layer **out, int *n_out, xmi_error **error) {
  for (xmlNodePtr c = first_el(node); c; c = next_el(c)) {
    if (!is(c, "layer")) continue;
    xmi_layer *l = (xmi_layer *)realloc(*out, sizeof(xmi_layer) * (*n_out + 1));
    if (!l) return fail(error, "out of memory");
    *out = l;
    if (!rd_layer(c, &l[*n_out], error)) return 0;
    (*n_out)++;
  }
  return 1;
}

static int rd_general(xmlNodePtr node, xmi_general **out, xmi_error **error) {
  xmi_general *g = (xmi_general *)calloc(1, sizeof(xmi_general));
  *out = g;
  char *v = attr_of(node, "version");
  if (v) {
    const int ok = sscanf(v, "%f", &g->version) == 1;
    free(v);
    if (!ok) return fail(error, "error reading in version of xml file");
  }
  for (xmlNodePtr c = first_el(node); c; c = next_el(c)) {
    if (is(c, "outputfile")) {
      free(g->outputfile);
      g->outputfile = text_of(c);
    } else if (is(c, "n_photons_interval")) {
      if (!RD_LONG(c, &g->n_photons_interval, "n_photons_interval")) return 0;
    } else if (is(c, "n_photons_line")) {
      if (!RD_LONG(c, &g->n_photons_line, "n_photons_line")) return 0;
    } else if (is(c, "n_interactions_trajectory")) {
      if (!RD_INT(c, &g->n_interactions_trajectory, "n_interactions_trajectory")) return 0;
    } else if (is(c, "comments")) {
      free(g->comments);
      g->comments = text_of(c);
    }
  }
  if (!g->outputfile) g->outputfile = strdup("");
  if (!g->comments) g->comments = strdup("");
  return 1;
}

static int rd_composition(xmlNodePtr node, xmi_composition **out, xmi_error **error) {
  xmi_composition *c0 = (xmi_composition *)calloc(1, sizeof(xmi_composition));
  *out = c0;
  for (xmlNodePtr c = first_el(node); c; c = next_el(c)) {
    if (is(c, "layer")) {
      xmi_layer *l = (xmi_layer *)realloc(c0->layers, sizeof(xmi_layer) * (c0->n_layers + 1));
      if (!l) return fail(error, "out of memory");
      c0->layers = l;
      if (!rd_layer(c, &l[c0->n_layers], error)) return 0;
      c0->n_layers++;
    } else if (is(c, "reference_layer")) {
      if (!RD_INT(c, &c0->reference_layer, "reference_layer")) return 0;
      if (c0->reference_layer < 1 || c0->reference_layer > c0->n_layers)
        return fail(error, "invalid reference_layer value detected");
    }
  }
  return 1;
}

static int rd_geometry(xmlNodePtr node, xmi_geometry **out, xmi_error **error) {
  xmi_geometry *g = (xmi_geometry *)calloc(1, sizeof(xmi_geometry));
  *out = g;
  for (xmlNodePtr c = first_el(node); c; c = next_el(c)) {
    int ok = 1;
    if (is(c, "d_sample_source")) ok = RD_DOUBLE(c, &g->d_sample_source, "d_sample_source");
    else if (is(c, "n_sample_orientation")) ok = rd_xyz(c, g->n_sample_orientation, "n_sample_orientation", error);
    else if (is(c, "p_detector_window")) ok = rd_xyz(c, g->p_detector_window, "p_detector_window", error);
    else if (is(c, "n_detector_orientation")) ok = rd_xyz(c, g->n_detector_orientation, "n_detector_orientation", error);
    else if (is(c, "area_detector")) ok = RD_DOUBLE(c, &g->area_detector, "area_detector");
    else if (is(c, "collimator_height")) ok = RD_DOUBLE(c, &g->collimator_height, "collimator_height");
    else if (is(c, "collimator_diameter")) ok = RD_DOUBLE(c, &g->collimator_diameter, "collimator_diameter");
    else if (is(c, "d_source_slit")) ok = RD_DOUBLE(c, &g->d_source_slit, "d_source_slit");
    else if (is(c, "slit_size")) {
      for (xmlNodePtr d = first_el(c); d && ok; d = next_el(d)) {
        if (is(d, "slit_size_x")) ok = RD_DOUBLE(d, &g->slit_size_x, "slit_size_x");
        else if (is(d, "slit_size_y")) ok = RD_DOUBLE(d, &g->slit_size_y, "slit_size_y");
      }
    }
    if (!ok) return 0;
  }
  normalize3(g->n_sample_orientation);
  normalize3(g->n_detector_orientation);
  return 1;
}

/* the seven numbers shared by discrete and continuous entries */
typedef struct {
  double energy, h, v, sx, sxp, sy, syp;
} beam_entry;

static int rd_beam(xmlNodePtr c, beam_entry *b, xmi_error **error) {
  for (xmlNodePtr d = first_el(c); d; d = next_el(d)) {
    int ok = 1;
    if (is(d, "energy")) ok = RD_DOUBLE(d, &b->energy, "energy");
    else if (is(d, "horizontal_intensity")) ok = RD_DOUBLE(d, &b->h, "horizontal_intensity");
    else if (is(d, "vertical_intensity")) ok = RD_DOUBLE(d, &b->v, "vertical_intensity");
    else if (is(d, "sigma_x")) ok = RD_DOUBLE(d, &b->sx, "sigma_x");
    else if (is(d, "sigma_xp")) ok = RD_DOUBLE(d, &b->sxp, "sigma_xp");
    else if (is(d, "sigma_y")) ok = RD_DOUBLE(d, &b->sy, "sigma_y");
    else if (is(d, "sigma_yp")) ok = RD_DOUBLE(d, &b->syp, "sigma_yp");
    if (!ok) return 0;
  }
  return 1;
}

static int cmp_discrete(const void *a, const void *b) {
  const double x = ((const xmi_energy_discrete *)a)->energy, y = ((const xmi_energy_discrete *)b)->energy;
  return x < y ? -1 : x > y ? 1 : 0;
}
static int cmp_continuous(const void *a, const void *b) {
  const double x = ((const xmi_energy_continuous *)a)->energy, y = ((const xmi_energy_continuous *)b)->energy;
  return x < y ? -1 : x > y ? 1 : 0;
}

/* rules on the continuous spectrum's zero-intensity points, shared by the
 * reader and xmi_input_validate (src/xmi_xml.c:975-1006) */
static const char *continuous_check(const xmi_excitation *x) {
  const int n = x->n_continuous;
#define POS(i) ((x->continuous[i].horizontal_intensity + x->continuous[i].vertical_intensity) > 0.0)
  if (n == 2) {
    if (x->continuous[0].horizontal_intensity + x->continuous[0].vertical_intensity +
            x->continuous[1].horizontal_intensity + x->continuous[1].vertical_intensity ==
        0.0)
      return "With only two continuous intervals, both cannot have zero intensity";
  } else if (n > 2) {
    for (int i = 1; i < n - 1; i++) {
      const int before = POS(i - 1), current = POS(i), after = POS(i + 1);
      if (i == 1 && before + current == 0) return "First two continuous intensity densities cannot have a total intensity of zero";
      if (i == n - 2 && current + after == 0) return "Last two continuous intensity densities cannot have a total intensity of zero";
      if (before + current + after == 0)
        return "Three consecutive continuous intensity densities cannot have a total intensity of zero";
    }
  }
#undef POS
  return NULL;
}

/* src/xmi_xml.c:679-1009: a discrete energy given twice adds its
 * intensities to the first entry; a continuous energy given twice is an
 * error; both lists are sorted by energy */
static int rd_excitation(xmlNodePtr node, xmi_excitation **out, xmi_error **error) {
  xmi_excitation *x = (xmi_excitation *)calloc(1, sizeof(xmi_excitation));
  *out = x;
  for (xmlNodePtr c = first_el(node); c; c = next_el(c)) {
    if (is(c, "discrete")) {
      beam_entry b = {0};
      int dist = XMI_ENERGY_DISCRETE_DISTRIBUTION_MONOCHROMATIC;
      double scale = 0.0;
      if (!rd_beam(c, &b, error)) return 0;
      for (xmlNodePtr d = first_el(c); d; d = next_el(d)) {
        if (!is(d, "scale_parameter")) continue;
        char *t = attr_of(d, "distribution_type");
        if (t && !strcmp(t, "gaussian")) dist = XMI_ENERGY_DISCRETE_DISTRIBUTION_GAUSSIAN;
        else if (t && !strcmp(t, "lorentzian")) dist = XMI_ENERGY_DISCRETE_DISTRIBUTION_LORENTZIAN;
        free(t);
        if (dist != XMI_ENERGY_DISCRETE_DISTRIBUTION_MONOCHROMATIC && !RD_DOUBLE(d, &scale, "scale_parameter")) return 0;
      }
      if (!(b.energy > 0.0 && b.energy <= 200.0 && b.h >= 0.0 && b.v >= 0.0 && b.h + b.v > 0.0))
        return fail(error, "invalid discrete energy %g (intensities %g, %g)", b.energy, b.h, b.v);
      int k = 0;
      for (; k < x->n_discrete; k++)
        if (x->discrete[k].energy == b.energy) break;
      if (k < x->n_discrete) {
        x->discrete[k].horizontal_intensity += b.h;
        x->discrete[k].vertical_intensity += b.v;
        continue;
      }
      xmi_energy_discrete *p = (xmi_energy_discrete *)realloc(x->discrete, sizeof(*p) * (x->n_discrete + 1));
      if (!p) return fail(error, "out of memory");
      x->discrete = p;
      p[x->n_discrete++] = (xmi_energy_discrete){b.energy, b.h, b.v, b.sx, b.sxp, b.sy, b.syp,
                                                 (XmiEnergyDiscreteDistribution)dist, scale};
    } else if (is(c, "continuous")) {
      beam_entry b = {0};
      if (!rd_beam(c, &b, error)) return 0;
      for (int k = 0; k < x->n_continuous; k++)
        if (x->continuous[k].energy == b.energy) return fail(error, "continuous energy %g given twice", b.energy);
      if (!(b.energy >= 0.0 && b.energy <= 200.0 && b.h >= 0.0 && b.v >= 0.0))
        return fail(error, "invalid continuous energy %g (intensities %g, %g)", b.energy, b.h, b.v);
      xmi_energy_continuous *p = (xmi_energy_continuous *)realloc(x->continuous, sizeof(*p) * (x->n_continuous + 1));
      if (!p) return fail(error, "out of memory");
      x->continuous = p;
      p[x->n_continuous++] = (xmi_energy_continuous){b.energy, b.h, b.v, b.sx, b.sxp, b.sy, b.syp};
    }
  }
  if ((x->n_continuous < 2 && x->n_discrete == 0) || x->n_continuous == 1)
    return fail(error, "excitation needs a discrete energy or at least two continuous energies");
  if (x->n_continuous > 1) qsort(x->continuous, x->n_continuous, sizeof(xmi_energy_continuous), cmp_continuous);
  if (x->n_discrete > 1) qsort(x->discrete, x->n_discrete, sizeof(xmi_energy_discrete), cmp_discrete);
  const char *msg = continuous_check(x);
  if (msg) return fail(error, "%s", msg);
  return 1;
}

static int rd_absorbers(xmlNodePtr node, xmi_absorbers **out, xmi_error **error) {
  xmi_absorbers *a = (xmi_absorbers *)calloc(1, sizeof(xmi_absorbers));
  *out = a;
  for (xmlNodePtr c = first_el(node); c; c = next_el(c)) {
    if (is(c, "excitation_path") && !rd_layers(c, &a->exc_layers, &a->n_exc_layers, error)) return 0;
    if (is(c, "detector_path") && !rd_layers(c, &a->det_layers, &a->n_det_layers, error)) return 0;
  }
  return 1;
}

static const char *const k_det_names[] = {"SiLi", "Ge", "Si_SDD"};

static int rd_detector(xmlNodePtr node, xmi_detector **out, xmi_error **error) {
  xmi_detector *d = (xmi_detector *)calloc(1, sizeof(xmi_detector));
  *out = d;
  d->nchannels = 2048;
  for (xmlNodePtr c = first_el(node); c; c = next_el(c)) {
    int ok = 1;
    if (is(c, "detector_type")) {
      char *t = text_of(c);
      int k = 0;
      for (; k < 3; k++)
        if (!strcmp(t, k_det_names[k])) break;
      free(t);
      if (k == 3) return fail(error, "unknown detector_type");
      d->detector_type = (XmiDetectorConvolutionProfile)k;
    } else if (is(c, "live_time")) ok = RD_DOUBLE(c, &d->live_time, "live_time");
    else if (is(c, "pulse_width")) ok = RD_DOUBLE(c, &d->pulse_width, "pulse_width");
    else if (is(c, "gain")) ok = RD_DOUBLE(c, &d->gain, "gain");
    else if (is(c, "zero")) ok = RD_DOUBLE(c, &d->zero, "zero");
    else if (is(c, "fano")) ok = RD_DOUBLE(c, &d->fano, "fano");
    else if (is(c, "noise")) ok = RD_DOUBLE(c, &d->noise, "noise");
    else if (is(c, "nchannels")) ok = RD_INT(c, &d->nchannels, "nchannels");
    else if (is(c, "crystal")) ok = rd_layers(c, &d->crystal_layers, &d->n_crystal_layers, error);
    if (!ok) return 0;
  }
  return 1;
}

/* src/xmi_xml.c:2084-2122, plus the presence of every section the DTD requires */
static int rd_input_body(xmlNodePtr root, xmi_input *in, xmi_error **error) {
  for (xmlNodePtr c = first_el(root); c; c = next_el(c)) {
    int ok = 1;
    if (is(c, "general")) ok = !in->general && rd_general(c, &in->general, error);
    else if (is(c, "composition")) ok = !in->composition && rd_composition(c, &in->composition, error);
    else if (is(c, "geometry")) ok = !in->geometry && rd_geometry(c, &in->geometry, error);
    else if (is(c, "excitation")) ok = !in->excitation && rd_excitation(c, &in->excitation, error);
    else if (is(c, "absorbers")) ok = !in->absorbers && rd_absorbers(c, &in->absorbers, error);
    else if (is(c, "detector")) ok = !in->detector && rd_detector(c, &in->detector, error);
    if (!ok) return fail(error, "invalid <%s> section", (const char *)c->name);
  }
  if (!in->general || !in->composition || !in->geometry || !in->excitation || !in->absorbers || !in->detector)
    return fail(error, "document misses one of general, composition, geometry, excitation, absorbers, detector");
  return 1;
}

/* ---- validation (xmi_input_validate) -------------------------------------- */

static int layers_ok(const xmi_layer *l, int n) {
  for (int i = 0; i < n; i++) {
    double sum = 0.0;
    for (int j = 0; j < l[i].n_elements; j++) {
      if (l[i].Z[j] < 1 || l[i].Z[j] > 94) return 0;
      if (l[i].weight[j] < 0.0 || l[i].weight[j] > 1.0) return 0;
      sum += l[i].weight[j];
    }
    if (sum <= 0.0 || l[i].density <= 0.0 || l[i].thickness <= 0.0) return 0;
  }
  return 1;
}

int xmi_input_validate(xmi_input *in) {
  int rv = 0;
  if (!in) return XMI_INPUT_GENERAL | XMI_INPUT_COMPOSITION | XMI_INPUT_GEOMETRY | XMI_INPUT_EXCITATION |
                  XMI_INPUT_ABSORBERS | XMI_INPUT_DETECTOR;
  const xmi_general *g = in->general;
  if (!g || g->n_photons_interval <= 0 || g->n_photons_line <= 0 || g->n_interactions_trajectory <= 0 ||
      !g->outputfile || !g->outputfile[0])
    rv |= XMI_INPUT_GENERAL;
  const xmi_composition *c = in->composition;
  if (!c || c->n_layers < 1 || c->reference_layer < 1 || c->reference_layer > c->n_layers ||
      !layers_ok(c->layers, c->n_layers))
    rv |= XMI_INPUT_COMPOSITION;
  const xmi_geometry *ge = in->geometry;
  if (!ge || ge->n_sample_orientation[2] <= 0.0 || ge->d_sample_source <= 0.0 || ge->area_detector <= 0.0 ||
      ge->collimator_height < 0.0 || ge->collimator_diameter < 0.0 || ge->d_source_slit <= 0.0 ||
      ge->slit_size_x <= 0.0 || ge->slit_size_y <= 0.0)
    rv |= XMI_INPUT_GEOMETRY;
  const xmi_excitation *x = in->excitation;
  int xbad = !x || (x->n_discrete == 0 && x->n_continuous < 2) || x->n_continuous == 1;
  for (int i = 0; !xbad && i < x->n_discrete; i++) {
    const xmi_energy_discrete *d = &x->discrete[i];
    xbad = d->energy <= 0.0 || d->horizontal_intensity < 0.0 || d->vertical_intensity < 0.0 ||
           d->horizontal_intensity + d->vertical_intensity <= 0.0 || d->sigma_x < 0.0 || d->sigma_y < 0.0 ||
           d->distribution_type < XMI_ENERGY_DISCRETE_DISTRIBUTION_MONOCHROMATIC ||
           d->distribution_type > XMI_ENERGY_DISCRETE_DISTRIBUTION_LORENTZIAN ||
           (d->distribution_type != XMI_ENERGY_DISCRETE_DISTRIBUTION_MONOCHROMATIC && d->scale_parameter <= 0.0);
  }
  for (int i = 0; !xbad && i < x->n_continuous; i++) {
    const xmi_energy_continuous *d = &x->continuous[i];
    xbad = d->energy < 0.0 || d->horizontal_intensity < 0.0 || d->vertical_intensity < 0.0 ||
           d->horizontal_intensity + d->vertical_intensity < 0.0 || d->sigma_x < 0.0 || d->sigma_y < 0.0;
  }
  if (!xbad && continuous_check(x)) xbad = 1;
  if (xbad) rv |= XMI_INPUT_EXCITATION;
  const xmi_absorbers *a = in->absorbers;
  if (!a || !layers_ok(a->exc_layers, a->n_exc_layers) || !layers_ok(a->det_layers, a->n_det_layers))
    rv |= XMI_INPUT_ABSORBERS;
  const xmi_detector *d = in->detector;
  if (!d || d->live_time <= 0.0 || d->pulse_width <= 0.0 || d->gain <= 0.0 || d->fano <= 0.0 || d->noise <= 0.0 ||
      d->n_crystal_layers < 1 || d->nchannels < 10 || !layers_ok(d->crystal_layers, d->n_crystal_layers))
    rv |= XMI_INPUT_DETECTOR;
  return rv;
}

/* ---- parsing entry points ------------------------------------------------ */

static xmlDocPtr parse_doc(const char *file, const char *buf, xmi_error **error) {
  LIBXML_TEST_VERSION
  const int opts = XML_PARSE_NOBLANKS | XML_PARSE_NONET | XML_PARSE_NOERROR | XML_PARSE_NOWARNING;
  xmlDocPtr doc = file ? xmlReadFile(file, NULL, opts) : xmlReadMemory(buf, (int)strlen(buf), "input.xml", NULL, opts);
  if (!doc) {
    const xmlError *e = xmlGetLastError();
    fail(error, "%s: %s", file ? file : "string", e && e->message ? e->message : "not well-formed XML");
  }
  return doc;
}

static xmi_input *input_from_doc(xmlDocPtr doc, xmi_error **error) {
  xmlNodePtr root = xmlDocGetRootElement(doc);
  /* the reference accepts the backwards-compatible root name only
   * (src/xmi_xml.c:1320); the DTD also names xmimsim-input */
  if (!root || !is(root, "xmimsim")) {
    fail(error, "XML document is of the wrong type, expected xmimsim");
    return NULL;
  }
  xmi_input *in = (xmi_input *)calloc(1, sizeof(xmi_input));
  if (!rd_input_body(root, in, error)) {
    xmi_input_free(in);
    return NULL;
  }
  const int flags = xmi_input_validate(in);
  if (flags) {
    fail(error, "input file failed validation (sections 0x%x)", flags);
    xmi_input_free(in);
    return NULL;
  }
  return in;
}

xmi_input *xmi_input_read_from_xml_file(const char *xmsifile, xmi_error **error) {
  if (!xmsifile) {
    fail(error, "no file name");
    return NULL;
  }
  xmlDocPtr doc = parse_doc(xmsifile, NULL, error);
  if (!doc) return NULL;
  xmi_input *in = input_from_doc(doc, error);
  xmlFreeDoc(doc);
  return in;
}

xmi_input *xmi_input_read_from_xml_string(const char *xmsistring, xmi_error **error) {
  if (!xmsistring) {
    fail(error, "no string");
    return NULL;
  }
  xmlDocPtr doc = parse_doc(NULL, xmsistring, error);
  if (!doc) return NULL;
  xmi_input *in = input_from_doc(doc, error);
  xmlFreeDoc(doc);
  return in;
}

/* ---- writing helpers ------------------------------------------------------ */

static xmlNodePtr add(xmlNodePtr parent, const char *name, const char *fmt, ...) {
  if (!fmt) return xmlNewChild(parent, NULL, (const xmlChar *)name, NULL);
  char buf[256];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return xmlNewChild(parent, NULL, (const xmlChar *)name, (const xmlChar *)buf);
}

/* text children of arbitrary content go through xmlNewTextChild, which
 * escapes markup characters; an empty text gives an empty element */
static xmlNodePtr add_text(xmlNodePtr parent, const char *name, const char *text) {
  if (!text || !text[0]) return xmlNewChild(parent, NULL, (const xmlChar *)name, NULL);
  return xmlNewTextChild(parent, NULL, (const xmlChar *)name, (const xmlChar *)text);
}

static void prop(xmlNodePtr n, const char *name, const char *fmt, ...) {
  char buf[256];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  xmlNewProp(n, (const xmlChar *)name, (const xmlChar *)buf);
}

/* layers with weights renormalised and given in percent; zero weights are
 * left out (src/xmi_xml.c:2553-2576) */
static void wr_layers(xmlNodePtr parent, const xmi_layer *l, int n) {
  for (int i = 0; i < n; i++) {
    xmlNodePtr ln = add(parent, "layer", NULL);
    double sum = 0.0;
    for (int j = 0; j < l[i].n_elements; j++) sum += l[i].weight[j];
    for (int j = 0; j < l[i].n_elements; j++) {
      if (fabs(l[i].weight[j]) < 1e-20) continue;
      xmlNodePtr e = add(ln, "element", NULL);
      add(e, "atomic_number", "%i", l[i].Z[j]);
      add(e, "weight_fraction", "%g", l[i].weight[j] / sum * 100.0);
    }
    add(ln, "density", "%g", l[i].density);
    add(ln, "thickness", "%g", l[i].thickness);
  }
}

static void wr_xyz(xmlNodePtr parent, const char *name, const double *v) {
  xmlNodePtr n = add(parent, name, NULL);
  add(n, "x", "%g", v[0]);
  add(n, "y", "%g", v[1]);
  add(n, "z", "%g", v[2]);
}

static void wr_beam(xmlNodePtr n, double E, double h, double v, double sx, double sxp, double sy, double syp) {
  add(n, "energy", "%g", E);
  add(n, "horizontal_intensity", "%g", h);
  add(n, "vertical_intensity", "%g", v);
  add(n, "sigma_x", "%g", sx);
  add(n, "sigma_xp", "%g", sxp);
  add(n, "sigma_y", "%g", sy);
  add(n, "sigma_yp", "%g", syp);
}

/* src/xmi_xml.c:1679-1812 */
static int wr_input_body(xmlNodePtr root, const xmi_input *in, xmi_error **error) {
  if (!in || !in->general || !in->composition || !in->geometry || !in->excitation || !in->absorbers ||
      !in->detector)
    return fail(error, "incomplete input");
  const xmi_general *g = in->general;
  xmlNodePtr n = add(root, "general", NULL);
  prop(n, "version", "%s", XMI_XML_VERSION);
  add_text(n, "outputfile", g->outputfile);
  add(n, "n_photons_interval", "%li", g->n_photons_interval);
  add(n, "n_photons_line", "%li", g->n_photons_line);
  add(n, "n_interactions_trajectory", "%i", g->n_interactions_trajectory);
  add_text(n, "comments", g->comments);

  n = add(root, "composition", NULL);
  wr_layers(n, in->composition->layers, in->composition->n_layers);
  add(n, "reference_layer", "%i", in->composition->reference_layer);

  const xmi_geometry *ge = in->geometry;
  n = add(root, "geometry", NULL);
  add(n, "d_sample_source", "%g", ge->d_sample_source);
  wr_xyz(n, "n_sample_orientation", ge->n_sample_orientation);
  wr_xyz(n, "p_detector_window", ge->p_detector_window);
  wr_xyz(n, "n_detector_orientation", ge->n_detector_orientation);
  add(n, "area_detector", "%g", ge->area_detector);
  add(n, "collimator_height", "%g", ge->collimator_height);
  add(n, "collimator_diameter", "%g", ge->collimator_diameter);
  add(n, "d_source_slit", "%g", ge->d_source_slit);
  xmlNodePtr s = add(n, "slit_size", NULL);
  add(s, "slit_size_x", "%g", ge->slit_size_x);
  add(s, "slit_size_y", "%g", ge->slit_size_y);

  const xmi_excitation *x = in->excitation;
  n = add(root, "excitation", NULL);
  for (int i = 0; i < x->n_discrete; i++) {
    const xmi_energy_discrete *d = &x->discrete[i];
    xmlNodePtr e = add(n, "discrete", NULL);
    wr_beam(e, d->energy, d->horizontal_intensity, d->vertical_intensity, d->sigma_x, d->sigma_xp, d->sigma_y,
            d->sigma_yp);
    /* written only for a non-monochromatic line, as older readers expect */
    if (d->distribution_type != XMI_ENERGY_DISCRETE_DISTRIBUTION_MONOCHROMATIC) {
      xmlNodePtr sp = add(e, "scale_parameter", "%g", d->scale_parameter);
      if (d->distribution_type == XMI_ENERGY_DISCRETE_DISTRIBUTION_GAUSSIAN) prop(sp, "distribution_type", "%s", "gaussian");
      else if (d->distribution_type == XMI_ENERGY_DISCRETE_DISTRIBUTION_LORENTZIAN)
        prop(sp, "distribution_type", "%s", "lorentzian");
    }
  }
  for (int i = 0; i < x->n_continuous; i++) {
    const xmi_energy_continuous *d = &x->continuous[i];
    wr_beam(add(n, "continuous", NULL), d->energy, d->horizontal_intensity, d->vertical_intensity, d->sigma_x,
            d->sigma_xp, d->sigma_y, d->sigma_yp);
  }

  n = add(root, "absorbers", NULL);
  if (in->absorbers->n_exc_layers > 0)
    wr_layers(add(n, "excitation_path", NULL), in->absorbers->exc_layers, in->absorbers->n_exc_layers);
  if (in->absorbers->n_det_layers > 0)
    wr_layers(add(n, "detector_path", NULL), in->absorbers->det_layers, in->absorbers->n_det_layers);

  const xmi_detector *d = in->detector;
  n = add(root, "detector", NULL);
  if ((int)d->detector_type < 0 || (int)d->detector_type > 2) return fail(error, "invalid detector_type");
  add(n, "detector_type", "%s", k_det_names[d->detector_type]);
  add(n, "live_time", "%g", d->live_time);
  add(n, "pulse_width", "%g", d->pulse_width);
  add(n, "nchannels", "%i", d->nchannels);
  add(n, "gain", "%g", d->gain);
  add(n, "zero", "%g", d->zero);
  add(n, "fano", "%g", d->fano);
  add(n, "noise", "%g", d->noise);
  wr_layers(add(n, "crystal", NULL), d->crystal_layers, d->n_crystal_layers);
  return 1;
}

/* the two comments ahead of the root element (src/xmi_xml.c:2124-2145) */
static void wr_default_comments(xmlDocPtr doc, xmlNodePtr root) {
  char stamp[96], host[256] = "", text[1024];
  time_t now = time(NULL);
  struct tm lt;
  localtime_r(&now, &lt);
  strftime(stamp, sizeof stamp, "%F %H:%M:%S (%Z)", &lt);
  gethostname(host, sizeof host - 1);
  const struct passwd *pw = getpwuid(getuid());
  const char *user = pw && pw->pw_name ? pw->pw_name : "unknown";
  char real[256];
  snprintf(real, sizeof real, "%s", pw && pw->pw_gecos && pw->pw_gecos[0] ? pw->pw_gecos : user);
  char *comma = strchr(real, ',');
  if (comma) *comma = '\0';
  snprintf(text, sizeof text, " <Creator>%s (%s)</Creator>\n <Timestamp>%s</Timestamp>\n <Hostname>%s</Hostname>", real,
           user, stamp, host);
  xmlNodePtr c = xmlAddPrevSibling(root, xmlNewDocComment(doc, (const xmlChar *)text));
  xmlAddNextSibling(c, xmlNewDocComment(doc, (const xmlChar *)"DO NOT MODIFY THIS FILE UNLESS YOU KNOW WHAT YOU ARE DOING!"));
}

static xmlDocPtr new_doc(const char *root_name, xmlNodePtr *root) {
  LIBXML_TEST_VERSION
  xmlDocPtr doc = xmlNewDoc((const xmlChar *)"1.0");
  xmlCreateIntSubset(doc, (const xmlChar *)root_name, NULL, (const xmlChar *)XMI_DTD_URL);
  *root = xmlNewNode(NULL, (const xmlChar *)root_name);
  xmlDocSetRootElement(doc, *root);
  wr_default_comments(doc, *root);
  return doc;
}

static int save_doc(xmlDocPtr doc, const char *file, xmi_error **error) {
  xmlIndentTreeOutput = 1;
  const int rc = xmlSaveFormatFileEnc(file, doc, NULL, 1);
  xmlFreeDoc(doc);
  return rc == -1 ? fail(error, "could not write to %s", file) : 1;
}

int xmi_input_write_to_xml_file(xmi_input *input, const char *xmsifile, xmi_error **error) {
  xmlNodePtr root;
  xmlDocPtr doc = new_doc("xmimsim", &root);
  if (!wr_input_body(root, input, error)) {
    xmlFreeDoc(doc);
    return 0;
  }
  return save_doc(doc, xmsifile, error);
}

int xmi_input_write_to_xml_string(xmi_input *input, char **xmlstring, xmi_error **error) {
  xmlNodePtr root;
  xmlDocPtr doc = new_doc("xmimsim", &root);
  if (!xmlstring || !wr_input_body(root, input, error)) {
    xmlFreeDoc(doc);
    return xmlstring ? 0 : fail(error, "no output location");
  }
  xmlChar *buf = NULL;
  int len = 0;
  xmlIndentTreeOutput = 1;
  xmlDocDumpFormatMemory(doc, &buf, &len, 1);
  xmlFreeDoc(doc);
  if (!buf) return fail(error, "could not serialise the document");
  *xmlstring = strdup((const char *)buf);
  xmlFree(buf);
  return 1;
}

/* ---- output struct (src/xmi_data_structs.c:1371-1519) --------------------- */

static xmi_input *input_copy(const xmi_input *A) {
  xmi_input *B = (xmi_input *)calloc(1, sizeof(xmi_input));
#define DUP(dst, src, n)                            \
  do {                                              \
    (dst) = malloc((n) > 0 ? (n) : 1);              \
    if ((n) > 0) memcpy((dst), (src), (n));         \
  } while (0)
  B->general = (xmi_general *)malloc(sizeof(xmi_general));
  *B->general = *A->general;
  B->general->outputfile = strdup(A->general->outputfile ? A->general->outputfile : "");
  B->general->comments = strdup(A->general->comments ? A->general->comments : "");
#define COPY_LAYERS(dst, src, n)                                                       \
  do {                                                                               \
    DUP(dst, src, sizeof(xmi_layer) * (size_t)(n));                                  \
    for (int i_ = 0; i_ < (n); i_++) {                                               \
      DUP((dst)[i_].Z, (src)[i_].Z, sizeof(int) * (size_t)(src)[i_].n_elements);     \
      DUP((dst)[i_].weight, (src)[i_].weight, sizeof(double) * (size_t)(src)[i_].n_elements); \
    }                                                                                \
  } while (0)
  B->composition = (xmi_composition *)malloc(sizeof(xmi_composition));
  *B->composition = *A->composition;
  COPY_LAYERS(B->composition->layers, A->composition->layers, A->composition->n_layers);
  B->geometry = (xmi_geometry *)malloc(sizeof(xmi_geometry));
  *B->geometry = *A->geometry;
  B->excitation = (xmi_excitation *)malloc(sizeof(xmi_excitation));
  *B->excitation = *A->excitation;
  DUP(B->excitation->discrete, A->excitation->discrete, sizeof(xmi_energy_discrete) * (size_t)A->excitation->n_discrete);
  DUP(B->excitation->continuous, A->excitation->continuous,
      sizeof(xmi_energy_continuous) * (size_t)A->excitation->n_continuous);
  B->absorbers = (xmi_absorbers *)malloc(sizeof(xmi_absorbers));
  *B->absorbers = *A->absorbers;
  COPY_LAYERS(B->absorbers->exc_layers, A->absorbers->exc_layers, A->absorbers->n_exc_layers);
  COPY_LAYERS(B->absorbers->det_layers, A->absorbers->det_layers, A->absorbers->n_det_layers);
  B->detector = (xmi_detector *)malloc(sizeof(xmi_detector));
  *B->detector = *A->detector;
  COPY_LAYERS(B->detector->crystal_layers, A->detector->crystal_layers, A->detector->n_crystal_layers);
#undef COPY_LAYERS
#undef DUP
  return B;
}

static int cmp_int(const void *a, const void *b) { return *(const int *)a - *(const int *)b; }

/* per element with counts, per line with counts, per order with positive
 * counts; hist is double[100][385][n_int] (Z, line, order; 1-based Z and
 * line in the first two indices) */
static int history_from_raw(const double *hist, const int *uniqZ, int nuniq, int n_int,
                            xmi_fluorescence_line_counts **out, int *n_out) {
#define H(Z, j, k) hist[((size_t)((Z)-1) * XMI_HIST_ROWS + (size_t)((j)-1)) * (size_t)n_int + (size_t)((k)-1)]
  *out = NULL;
  *n_out = 0;
  for (int i = 0; i < nuniq; i++) {
    const int Z = uniqZ[i];
    double total = 0.0;
    for (int j = 1; j <= 383; j++)
      for (int k = 1; k <= n_int; k++) total += H(Z, j, k);
    if (total == 0.0) continue;
    xmi_fluorescence_line_counts *h =
        (xmi_fluorescence_line_counts *)realloc(*out, sizeof(xmi_fluorescence_line_counts) * (*n_out + 1));
    if (!h) return 0;
    *out = h;
    xmi_fluorescence_line_counts *e = &h[(*n_out)++];
    e->atomic_number = Z;
    e->total_counts = total;
    e->n_lines = 0;
    e->lines = NULL;
    for (int j = 1; j <= 383; j++) {
      double lt = 0.0;
      for (int k = 1; k <= n_int; k++) lt += H(Z, j, k);
      if (lt == 0.0) continue;
      xmi_fluorescence_line *ls = (xmi_fluorescence_line *)realloc(e->lines, sizeof(xmi_fluorescence_line) * (e->n_lines + 1));
      if (!ls) return 0;
      e->lines = ls;
      xmi_fluorescence_line *l = &ls[e->n_lines++];
      l->line_type = strdup(xmi_line_name(j));
      l->energy = xmi_provider_line_data(Z, 2, j);
      l->total_counts = lt;
      l->n_interactions = 0;
      l->interactions = (xmi_counts *)malloc(sizeof(xmi_counts) * (size_t)n_int);
      for (int k = 1; k <= n_int; k++) {
        if (H(Z, j, k) <= 0.0) continue;
        l->interactions[l->n_interactions].counts = H(Z, j, k);
        l->interactions[l->n_interactions].interaction_number = k;
        l->n_interactions++;
      }
    }
  }
#undef H
  return 1;
}

xmi_output *xmi_output_raw2struct(xmi_input *input, double *brute_history, double *var_red_history,
                                  double **channels_conv, double *channels_unconv, char *inputfile,
                                  int use_zero_interactions) {
  if (!input || !input->general || !input->composition || !input->detector || !brute_history || !channels_conv ||
      !channels_unconv) {
    xmi_set_error("xmi_output_raw2struct: missing argument");
    return NULL;
  }
  const int n_int = input->general->n_interactions_trajectory, nch = input->detector->nchannels;
  xmi_output *o = (xmi_output *)calloc(1, sizeof(xmi_output));
  o->version = (float)atof(XMI_XML_VERSION);
  o->input = input_copy(input);
  o->inputfile = strdup(inputfile ? inputfile : "");
  o->outputfile = strdup(input->general->outputfile ? input->general->outputfile : "");
  o->use_zero_interactions = use_zero_interactions;
  o->ninteractions = n_int;
  o->channels_conv = (double **)calloc((size_t)n_int + 1, sizeof(double *));
  o->channels_unconv = (double **)calloc((size_t)n_int + 1, sizeof(double *));
  for (int i = 0; i <= n_int; i++) {
    o->channels_conv[i] = (double *)calloc((size_t)nch, sizeof(double));
    o->channels_unconv[i] = (double *)calloc((size_t)nch, sizeof(double));
    if (i == 0 && !use_zero_interactions) continue;
    for (int j = 0; j < nch; j++) {
      o->channels_unconv[i][j] = channels_unconv[(size_t)i * nch + j];
      o->channels_conv[i][j] = channels_conv[i][j];
    }
  }
  /* unique atomic numbers of the sample, ascending */
  int n = 0;
  for (int l = 0; l < input->composition->n_layers; l++) n += input->composition->layers[l].n_elements;
  int *uniq = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1)), nu = 0;
  for (int l = 0; l < input->composition->n_layers; l++)
    for (int e = 0; e < input->composition->layers[l].n_elements; e++) {
      const int Z = input->composition->layers[l].Z[e];
      int k = 0;
      while (k < nu && uniq[k] != Z) k++;
      if (k == nu) uniq[nu++] = Z;
    }
  qsort(uniq, nu, sizeof(int), cmp_int);
  int ok = history_from_raw(brute_history, uniq, nu, n_int, &o->brute_force_history, &o->nbrute_force_history);
  if (ok && var_red_history)
    ok = history_from_raw(var_red_history, uniq, nu, n_int, &o->var_red_history, &o->nvar_red_history);
  free(uniq);
  if (!ok) {
    xmi_output_free(o);
    xmi_set_error("xmi_output_raw2struct: out of memory");
    return NULL;
  }
  return o;
}

/* ---- output writer (src/xmi_xml.c:1453-1636, 1880-2031) ------------------- */

static void wr_spectrum(xmlNodePtr parent, const char *name, const xmi_output *o, double *const *ch) {
  const xmi_detector *d = o->input->detector;
  xmlNodePtr n = add(parent, name, NULL);
  for (int j = 0; j < d->nchannels; j++) {
    xmlNodePtr c = add(n, "channel", NULL);
    add(c, "channelnr", "%i", j);
    add(c, "energy", "%g", d->gain * j + d->zero);
    for (int i = o->use_zero_interactions ? 0 : 1; i <= o->input->general->n_interactions_trajectory; i++)
      prop(add(c, "counts", "%g", ch[i][j]), "interaction_number", "%i", i);
  }
}

static int wr_history(xmlNodePtr parent, const xmi_fluorescence_line_counts *h, int n, xmi_error **error) {
  for (int i = 0; i < n; i++) {
    xmlNodePtr e = add(parent, "fluorescence_line_counts", NULL);
    const char *sym = xmi_element_symbol(h[i].atomic_number);
    if (!sym) return fail(error, "no symbol for atomic number %d", h[i].atomic_number);
    prop(e, "atomic_number", "%i", h[i].atomic_number);
    prop(e, "symbol", "%s", sym);
    prop(e, "total_counts", "%g", h[i].total_counts);
    for (int j = 0; j < h[i].n_lines; j++) {
      const xmi_fluorescence_line *l = &h[i].lines[j];
      xmlNodePtr ln = add(e, "fluorescence_line", NULL);
      prop(ln, "type", "%s", l->line_type);
      prop(ln, "energy", "%g", l->energy);
      prop(ln, "total_counts", "%g", l->total_counts);
      for (int k = 0; k < l->n_interactions; k++)
        prop(add(ln, "counts", "%g", l->interactions[k].counts), "interaction_number", "%i",
             l->interactions[k].interaction_number);
    }
  }
  return 1;
}

/* plot coordinates of the preview graphs, in single precision as the
 * reference computes them */
static float svg_x(double E, const double *energies, int last) {
  float x = (float)(SVG_BOX_WIDTH * (E - energies[0]) / (energies[last - 1] - energies[0]));
  return x < 0 ? 0.0f : x > SVG_BOX_WIDTH ? (float)SVG_BOX_WIDTH : x;
}

static float svg_y(double I, double max_log, double min_log) {
  const double Ic = I < 1 ? 1 : I;
  float y = (float)SVG_BOX_HEIGHT * (log10(Ic) - min_log) / (max_log - min_log);
  return y < 0 ? 0.0f : y > SVG_BOX_HEIGHT ? (float)SVG_BOX_HEIGHT : y;
}

static void wr_graph(xmlNodePtr parent, const xmi_input *in, const char *name, int interaction, const double *ch,
                     double maximum) {
  const int nch = in->detector->nchannels;
  const double minimum = 1, max_log = log10(maximum);
  double min_log = log10(minimum);
  int max_channel = 0;
  for (int i = nch - 1; i >= 0; i--)
    if (ch[i] >= 1) {
      max_channel = i;
      break;
    }
  double *E = (double *)malloc(sizeof(double) * (size_t)nch);
  for (int i = 0; i < nch; i++) E[i] = (double)i * in->detector->gain + in->detector->zero;
  xmlNodePtr g = add(parent, "graphic", NULL);
  xmlNodePtr id = add(g, "id", NULL);
  add_text(id, "name", name);
  add(id, "interaction", "%i", interaction);
  xmlNodePtr rect = add(g, "rect", NULL);
  add(rect, "view", NULL);
  xmlNodePtr size = add(rect, "size", NULL);
  add(size, "width", "%i", SVG_BOX_WIDTH);
  add(size, "height", "%i", SVG_BOX_HEIGHT);
  add(size, "min_energy", "%g", E[0]);
  /* the channel below the last one with a count, as the reference writes
   * (an empty spectrum indexes channel -1 there; 0 here) */
  add(size, "max_energy", "%g", E[max_channel > 0 ? max_channel - 1 : 0]);
  xmlNodePtr ax = add(rect, "x-axis", NULL);
  add(ax, "name", "%s", "Energy (keV)");
  for (int k = 0; (double)(5 * k) <= E[max_channel]; k++) {
    xmlNodePtr ix = add(ax, "index", NULL);
    add(ix, "value", "%g", (double)svg_x(5.0 * k, E, max_channel));
    add(ix, "name", "%.1f", 5.0 * k);
  }
  xmlNodePtr ay = add(rect, "y-axis", NULL);
  add(ay, "name", "%s", "Intensity (counts)");
  if (min_log < 0.0) min_log = 0.0;
  for (double I = 1.0; I <= maximum; I *= 10.0) {
    if (I < minimum) continue;
    xmlNodePtr ix = add(ay, "index", NULL);
    add(ix, "value", "%g", (double)svg_y(I, max_log, min_log));
    add(ix, "name", "%.0f", I);
  }
  xmlNodePtr pts = add(g, "points", NULL);
  add(pts, "color", "%s", "blue");
  for (int i = 0; i <= max_channel; i++) {
    xmlNodePtr p = add(pts, "point", NULL);
    add(p, "x", "%g", (double)svg_x(E[i], E, max_channel));
    add(p, "y", "%g", (double)svg_y(ch[i], max_log, min_log));
  }
  free(E);
}

static double spectra_max(const xmi_output *o, double *const *ch) {
  double m = 0.0;
  for (int i = o->use_zero_interactions ? 0 : 1; i <= o->input->general->n_interactions_trajectory; i++)
    for (int j = 0; j < o->input->detector->nchannels; j++)
      if (ch[i][j] > m) m = ch[i][j];
  return m;
}

static int wr_output_body(xmlNodePtr root, const xmi_output *o, int with_svg, xmi_error **error) {
  if (!o || !o->input || !o->channels_conv || !o->channels_unconv) return fail(error, "incomplete output");
  prop(root, "version", "%s", XMI_XML_VERSION);
  add_text(root, "inputfile", o->inputfile);
  wr_spectrum(root, "spectrum_conv", o, o->channels_conv);
  wr_spectrum(root, "spectrum_unconv", o, o->channels_unconv);
  if (!wr_history(add(root, "brute_force_history", NULL), o->brute_force_history, o->nbrute_force_history, error))
    return 0;
  xmlNodePtr vr = add(root, "variance_reduction_history", NULL);
  if (o->var_red_history && !wr_history(vr, o->var_red_history, o->nvar_red_history, error)) return 0;
  if (!wr_input_body(add(root, "xmimsim-input", NULL), o->input, error)) return 0;
  if (!with_svg) return 1;
  xmlNodePtr svg = add(root, "svg_graphs", NULL);
  const double conv_max = spectra_max(o, o->channels_conv), unconv_max = spectra_max(o, o->channels_unconv);
  for (int i = o->use_zero_interactions ? 0 : 1; i <= o->input->general->n_interactions_trajectory; i++) {
    wr_graph(svg, o->input, "convoluted", i, o->channels_conv[i], conv_max);
    wr_graph(svg, o->input, "unconvoluted", i, o->channels_unconv[i], unconv_max);
  }
  return 1;
}

int xmi_output_write_to_xml_file(xmi_output *output, const char *xmsofile, xmi_error **error) {
  xmlNodePtr root;
  xmlDocPtr doc = new_doc("xmimsim-results", &root);
  if (!wr_output_body(root, output, 1, error)) {
    xmlFreeDoc(doc);
    return 0;
  }
  return save_doc(doc, xmsofile, error);
}

/* ---- output reader (src/xmi_xml.c:165-362, 2033-2308) --------------------- */

static int rd_spectrum(xmlNodePtr node, xmi_output *o, double ***dst, xmi_error **error) {
  const int nch = (int)xmlChildElementCount(node);
  if (nch < 1) return fail(error, "spectrum without channels");
  double **ch = (double **)calloc((size_t)o->ninteractions + 1, sizeof(double *));
  *dst = ch;
  for (int i = 0; i <= o->ninteractions; i++) ch[i] = (double *)calloc((size_t)nch, sizeof(double));
  int j = 0;
  for (xmlNodePtr c = first_el(node); c; c = next_el(c)) {
    if (!is(c, "channel")) continue;
    if (j >= nch) return fail(error, "too many channels");
    for (xmlNodePtr k = first_el(c); k; k = next_el(k)) {
      if (!is(k, "counts")) continue;
      char *a = attr_of(k, "interaction_number");
      const int it = a ? (int)strtol(a, NULL, 10) : -1;
      free(a);
      if (it < 0 || it > o->ninteractions) return fail(error, "interaction_number out of range");
      if (!RD_DOUBLE(k, &ch[it][j], "counts")) return 0;
    }
    j++;
  }
  return 1;
}

static int rd_history(xmlNodePtr node, xmi_fluorescence_line_counts **out, int *n_out, xmi_error **error) {
  const int n = (int)xmlChildElementCount(node);
  *n_out = 0;
  *out = NULL;
  if (n == 0) return 1;
  xmi_fluorescence_line_counts *h = (xmi_fluorescence_line_counts *)calloc((size_t)n, sizeof(*h));
  *out = h;
  *n_out = n;
  int i = 0;
  for (xmlNodePtr e = first_el(node); e; e = next_el(e), i++) {
    char *a = attr_of(e, "atomic_number"), *t = attr_of(e, "total_counts");
    const int ok = a && t && sscanf(a, "%i", &h[i].atomic_number) == 1 && sscanf(t, "%lg", &h[i].total_counts) == 1;
    free(a);
    free(t);
    if (!ok) return fail(error, "invalid fluorescence_line_counts attributes");
    h[i].n_lines = (int)xmlChildElementCount(e);
    h[i].lines = (xmi_fluorescence_line *)calloc((size_t)(h[i].n_lines > 0 ? h[i].n_lines : 1), sizeof(xmi_fluorescence_line));
    int j = 0;
    for (xmlNodePtr l = first_el(e); l; l = next_el(l), j++) {
      xmi_fluorescence_line *fl = &h[i].lines[j];
      char *ty = attr_of(l, "type"), *en = attr_of(l, "energy"), *tc = attr_of(l, "total_counts");
      const int ok2 = ty && en && tc && sscanf(en, "%lg", &fl->energy) == 1 && sscanf(tc, "%lg", &fl->total_counts) == 1;
      fl->line_type = ty;
      free(en);
      free(tc);
      if (!ok2) return fail(error, "invalid fluorescence_line attributes");
      fl->n_interactions = (int)xmlChildElementCount(l);
      fl->interactions = (xmi_counts *)calloc((size_t)(fl->n_interactions > 0 ? fl->n_interactions : 1), sizeof(xmi_counts));
      int k = 0;
      for (xmlNodePtr c = first_el(l); c; c = next_el(c), k++) {
        char *in = attr_of(c, "interaction_number");
        const int ok3 = in && sscanf(in, "%i", &fl->interactions[k].interaction_number) == 1;
        free(in);
        if (!ok3 || !RD_DOUBLE(c, &fl->interactions[k].counts, "counts")) return fail(error, "invalid counts");
      }
    }
  }
  return 1;
}

xmi_output *xmi_output_read_from_xml_file(const char *xmsofile, xmi_error **error) {
  if (!xmsofile) {
    fail(error, "no file name");
    return NULL;
  }
  xmlDocPtr doc = parse_doc(xmsofile, NULL, error);
  if (!doc) return NULL;
  xmlNodePtr root = xmlDocGetRootElement(doc);
  xmi_output *o = NULL;
  if (!root || !is(root, "xmimsim-results")) {
    fail(error, "XML document %s is of the wrong type, expected xmimsim-results", xmsofile);
    goto done;
  }
  o = (xmi_output *)calloc(1, sizeof(xmi_output));
  char *v = attr_of(root, "version");
  if (v && sscanf(v, "%f", &o->version) != 1) {
    free(v);
    fail(error, "error reading in version of xml file");
    goto bad;
  }
  free(v);
  xmlNodePtr inp = NULL;
  for (xmlNodePtr c = first_el(root); c; c = next_el(c))
    if (is(c, "xmimsim-input")) inp = c;
  if (!inp) {
    fail(error, "no xmimsim-input element");
    goto bad;
  }
  o->input = (xmi_input *)calloc(1, sizeof(xmi_input));
  if (!rd_input_body(inp, o->input, error)) goto bad;
  if (xmi_input_validate(o->input)) {
    fail(error, "the input of %s failed validation", xmsofile);
    goto bad;
  }
  o->ninteractions = o->input->general->n_interactions_trajectory;
  for (xmlNodePtr c = first_el(root); c; c = next_el(c)) {
    if (is(c, "inputfile")) {
      free(o->inputfile);
      o->inputfile = text_of(c);
    } else if (is(c, "spectrum_conv")) {
      if (!rd_spectrum(c, o, &o->channels_conv, error)) goto bad;
      double sum = 0.0;
      for (int i = 0; i < o->input->detector->nchannels && i < (int)xmlChildElementCount(c); i++) sum += o->channels_conv[0][i];
      o->use_zero_interactions = sum != 0.0;
    } else if (is(c, "spectrum_unconv")) {
      if (!rd_spectrum(c, o, &o->channels_unconv, error)) goto bad;
    } else if (is(c, "brute_force_history")) {
      if (!rd_history(c, &o->brute_force_history, &o->nbrute_force_history, error)) goto bad;
    } else if (is(c, "variance_reduction_history")) {
      if (!rd_history(c, &o->var_red_history, &o->nvar_red_history, error)) goto bad;
    }
  }
  if (!o->channels_conv || !o->channels_unconv) {
    fail(error, "missing spectra");
    goto bad;
  }
  if (!o->inputfile) o->inputfile = strdup("");
  o->outputfile = strdup(xmsofile);
  goto done;
bad:
  xmi_output_free(o);
  o = NULL;
done:
  xmlFreeDoc(doc);
  return o;
}
