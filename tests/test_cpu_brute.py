"""CPU tests of the oracle's brute-force path (use_variance_reduction = 0):
the analogue step against the closed-form transmission of a layer stack, the
detector tests against exact solid angles and hand-built segments, and the
whole path, cascades included, against the variance-reduced estimator on the
same input (src/xmi_main.F90:435-538,1188-1685,1920-1984,2413-4783;
src/xmi_aux_f.F90:1622-1833).  Each check is shown to reject a planted
mistake."""
import math

import numpy as np
import pytest

from xmimsim_amd import api
from xmimsim_amd.api import DiscreteLine, Input, Layer


def _brute_options(**kw):
    return api.default_options(use_variance_reduction=0, **kw)


def _runs(orc, prep, seeds):
    return [orc.run(prep, seed=s) for s in seeds]


# ---- the analogue step: transmission through a stack ------------------------

def _transmission_input(n):
    """A pencil beam through three layers into a detector on the beam axis
    behind the sample (transmission geometry)."""
    return Input(layers=[Layer([13], [1.0], 2.7, 0.05), Layer([26, 28], [0.6, 0.4], 7.9, 0.002),
                         Layer([6, 8], [0.5, 0.5], 1.1, 0.2)],
                 discrete=[DiscreteLine(20.0, 1e9, 0.0)], n_photons_line=n, n_interactions_trajectory=2,
                 p_detector_window=(0.0, 0.0, 110.0), n_detector_orientation=(0.0, 0.0, -1.0), area_detector=1.0,
                 gain=0.02, nchannels=1200)


def _uncollided(orc, prep):
    """Weight in the line's channel of the photons that reach the detector
    without interacting (channels row 0)."""
    ch, vr, bf, st = orc.run(prep, seed=11)
    assert vr is None
    return ch[0, int(20.0 / 0.02)], ch[0].sum()


def test_brute_transmission_closed_form(orc):
    """Uncollided photons reach the detector with probability
    exp(-sum mu_i rho_i t_i / cos a): the analogue step's layer walk (the
    running transmission blbs of src/xmi_main.F90:1331-1400) reproduces it
    to binomial precision; dropping the transmission of the layers already
    crossed does not."""
    n = 400_000
    inp = _transmission_input(n)
    prep = api.Prepared(inp, options=_brute_options())
    d = orc.input_derived(prep)
    cosa = abs(d["n_sample_orientation"][2])
    tau = 0.0
    from xmimsim_amd import lines
    for L in inp.layers:
        mu = sum(w * lines.cs_total(Z, 20.0) for Z, w in zip(L.Z, L.weight))
        tau += mu * L.density * L.thickness / cosa
    T = math.exp(-tau)
    assert 0.2 < T < 0.8
    line, total = _uncollided(orc, prep)
    assert line == pytest.approx(total)  # nothing but the primary line is uncollided
    sigma = math.sqrt(T * (1 - T) / n)
    assert abs(line / 1e9 - T) < 4 * sigma, (line / 1e9, T, sigma)
    with orc.planted("brute_blbs"):
        bad, _ = _uncollided(orc, prep)
    assert abs(bad / 1e9 - T) > 10 * sigma
    prep.close()


# ---- the detector tests ------------------------------------------------------

def _disk_solid_angle(R, rho, h, coll=None, n=2_000_000, seed=0):
    """Solid angle of the window disk (radius R) from a point at height h
    above the window plane and rho off the axis, integrated over uniform
    points of the disk; with a collimator (Rc, H) a disk point counts when the
    line to it crosses the plane at height H within Rc."""
    rng = np.random.default_rng(seed)
    r = R * np.sqrt(rng.random(n))
    a = 2 * math.pi * rng.random(n)
    dx, dy = r * np.cos(a), r * np.sin(a) - rho
    f = h / (dx * dx + dy * dy + h * h) ** 1.5
    if coll is not None:
        Rc, H = coll
        t = (h - H) / h
        f = f * ((t * dx) ** 2 + (rho + t * dy) ** 2 <= Rc * Rc)
    return math.pi * R * R * f.mean()


@pytest.mark.parametrize("coll", [None, (0.2, 0.4)])
def test_detector_hit_solid_angle(orc, coll):
    """The fraction of isotropic directions from a point that pass the
    detector test of an escaped photon (src/xmi_main.F90:1920-1984) is the
    detector's solid angle over 4 pi, with and without a conical
    collimator."""
    kw = {} if coll is None else {"collimator_diameter": 2 * coll[0], "collimator_height": coll[1]}
    inp = Input(layers=[Layer([14], [1.0], 2.33, 0.1)], discrete=[DiscreteLine(20.0, 1e9, 0.0)],
                area_detector=0.5, **kw)
    prep = api.Prepared(inp, options=_brute_options())
    d = orc.input_derived(prep)
    R = d["detector_radius"]
    new = d["new"]
    p0 = np.array(inp.p_detector_window)
    rng = np.random.default_rng(3)
    n = 2_000_000
    for (h, rho) in ((1.0, 0.0), (0.8, 0.25), (1.5, -0.4)):
        point = p0 + new @ np.array([h, rho, 0.0])
        v = rng.normal(size=(n, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        hits = orc.probe_detector_hit(prep, np.repeat(point[None], n, axis=0), v)
        frac = hits.mean()
        exact = _disk_solid_angle(R, rho, h, coll) / (4 * math.pi)
        sigma = math.sqrt(exact * (1 - exact) / n)
        assert abs(frac - exact) < 4 * sigma + 2e-3 * exact, (h, rho, frac, exact)
    if coll is not None:
        with orc.planted("no_collimator"):
            point = p0 + new @ np.array([1.0, 0.0, 0.0])
            hits = orc.probe_detector_hit(prep, np.repeat(point[None], n, axis=0), v)
        exact = _disk_solid_angle(R, 0.0, 1.0, coll) / (4 * math.pi)
        assert abs(hits.mean() - exact) > 20 * math.sqrt(exact / n)
    prep.close()


def test_detector_segment_codes(orc):
    """Hand-built path segments through the detector test of
    src/xmi_aux_f.F90:1622-1833, in the detector frame (x along the detector
    axis towards the sample, window at x = 0):
    - without a collimator the reference gives the line the direction
      (L, L, L), L the segment length, so a crossing counts where
      (0, y_e - x_e, z_e - x_e) lies within the radius, and always as one
      from behind;
    - with a collimator, crossing the cone's wall below its height is a
      collimator hit, entering the window from the sample side a window hit,
      from behind a bad one."""
    def frame(prep, inp):
        d = orc.input_derived(prep)
        new, p0 = d["new"], np.array(inp.p_detector_window)
        return d, (lambda x: p0 + new @ np.asarray(x, dtype=float))

    inp = Input(layers=[Layer([14], [1.0], 2.33, 0.1)], discrete=[DiscreteLine(20.0, 1e9, 0.0)], area_detector=0.5)
    prep = api.Prepared(inp, options=_brute_options())
    d, lab = frame(prep, inp)
    R = d["detector_radius"]
    segs = [
        ((0.3, 0.0, 0.0), (0.5, 0.0, 0.0), 0),       # does not reach the plane
        ((0.2, 0.0, 0.0), (-0.1, 0.0, 0.0), 3),      # on the axis: (0, 0.1, 0.1) within R, from behind
        ((0.2, 0.05, 0.0), (-0.5, 0.05, 0.0), 0),    # crossing at y = 0.05, but (0, 0.55, 0.5) lies outside R
        ((0.05, 0.0, 0.0), (-0.05, 0.9 * R, 0.0), 0),  # inside R at the crossing, garbled point (0, 0.9R+0.05, 0.05)
        ((-0.1, 0.0, 0.0), (0.1, 0.0, 0.0), 3),      # moving towards the sample, same garbled point rule
    ]
    # the garbled point of case 4 is (0, 0.9 R + 0.05, 0.05): outside R when R < 0.53
    assert R < 0.5
    codes = orc.probe_detector_segment(prep, np.array([lab(b) for b, e, c in segs]),
                                       np.array([lab(e) for b, e, c in segs]))
    assert codes.tolist() == [c for _, _, c in segs]
    prep.close()

    Rc, H = 0.2, 0.4
    inp = Input(layers=[Layer([14], [1.0], 2.33, 0.1)], discrete=[DiscreteLine(20.0, 1e9, 0.0)], area_detector=0.5,
                collimator_diameter=2 * Rc, collimator_height=H)
    prep = api.Prepared(inp, options=_brute_options())
    d, lab = frame(prep, inp)
    R = d["detector_radius"]
    # cone wall radius at height x: R - (R - Rc) x / H
    wall = lambda x: R - (R - Rc) * x / H
    segs = [
        ((1.0, 0.05, 0.0), (-0.01, 0.05, 0.0), 2),                # parallel to the axis into the window
        ((-0.01, 0.05, 0.0), (1.0, 0.05, 0.0), 3),                # out of the window, from behind
        ((0.15, 0.0, 0.0), (0.25, 2 * wall(0.2), 0.0), 1),        # through the wall at half height
        ((0.55, 0.0, 0.0), (0.65, 2 * wall(0.6), 0.0), 0),        # through the cone above the collimator's top
        ((2.0, 3.0, 0.0), (2.5, 3.0, 0.5), 0),                    # far away
    ]
    codes = orc.probe_detector_segment(prep, np.array([lab(b) for b, e, c in segs]),
                                       np.array([lab(e) for b, e, c in segs]))
    assert codes.tolist() == [c for _, _, c in segs], codes
    prep.close()


# ---- the whole path against the variance-reduced estimator ------------------

def _fe_sn(n, pol=(5e8, 5e8)):
    """10 um Fe/Sn at 40 keV, a 1 cm^2 x 3 detector 1 cm away (a seventh of
    the sphere, so that brute force collects enough photons)."""
    return Input(layers=[Layer([26, 50], [0.5, 0.5], 5.0, 0.001)], discrete=[DiscreteLine(40.0, *pol)],
                 n_photons_line=n, n_interactions_trajectory=1, area_detector=3.0, gain=0.02, nchannels=2500)


LINES = [(50, 3), (50, 2), (26, 3), (50, 90), (50, 89), (50, 63)]  # Sn KL3 KL2, Fe KL3, Sn L3M5 L3M4 L2M4


_TABLES = {}


def _brute_lines(orc, inp, opts, seeds):
    # a brute-force run reads no cascade-dependent table: one set serves all
    if "fe_sn" not in _TABLES:
        _TABLES["fe_sn"] = api.SharedTables([26, 50])
    prep = api.Prepared(inp, options=opts, tables=_TABLES["fe_sn"])
    bf = np.array([r[2] for r in _runs(orc, prep, seeds)])
    prep.close()
    return {(Z, l): (bf[:, Z - 1, l - 1, 0].mean(), bf[:, Z - 1, l - 1, 0].std(ddof=1) / math.sqrt(len(seeds)))
            for Z, l in LINES}


_VR_CACHE = {}


def _vr_lines(orc, inp, aug, rad):
    """The estimator's first-order line intensities (one run per cascade
    setting, kept for the tests of this module)."""
    key = (aug, rad)
    if key not in _VR_CACHE:
        prep = api.Prepared(inp, options=api.default_options(use_cascade_auger=aug, use_cascade_radiative=rad))
        orc.solid_angle(prep, 128, 5)
        _, vr, _, _ = orc.run(prep, seed=1)
        prep.close()
        _VR_CACHE[key] = {(Z, l): vr[Z - 1, l - 1, 0] for Z, l in LINES}
    return _VR_CACHE[key]


@pytest.mark.parametrize("aug,rad", [(1, 1), (0, 0), (1, 0), (0, 1)])
def test_brute_lines_match_estimator_with_cascades(orc, aug, rad):
    """First-order line intensities of the brute-force run (analogue yields,
    explicit Auger and radiative cascades with their offspring photons)
    equal the estimator's (cascade-corrected cross sections), within 3 % and
    4 sigma of the brute-force scatter, for every cascade setting.  The Sn L
    lines at 40 keV come mostly from cascades of K vacancies (a factor of
    about 8 between no cascades and both), so the transition tables, the
    vacancy bookkeeping and the offspring tallies are all on the line."""
    inp = _fe_sn(1_000_000)
    seeds = range(16)
    b = _brute_lines(orc, inp, _brute_options(use_cascade_auger=aug, use_cascade_radiative=rad), seeds)
    v = _vr_lines(orc, inp, aug, rad)
    for key in LINES:
        m, se = b[key]
        assert abs(m - v[key]) < 4 * se + 0.03 * v[key], (aug, rad, key, m, se, v[key])


def test_brute_cascade_plants_rejected(orc):
    """Dropping the second Auger vacancy (the Sn L lines fed by Auger
    cascades lose half their excess), or mapping a K line's vacancy one shell
    too low (the L3 lines lose their radiative feeding), moves the summed
    lines off the estimator's."""
    inp = _fe_sn(1_000_000)
    seeds = range(100, 132)
    for plant, aug, rad, keys in (("auger_one_vacancy", 1, 0, [(50, 90), (50, 89), (50, 63)]),
                                  ("radiative_shell", 0, 1, [(50, 90), (50, 89)])):
        v = sum(_vr_lines(orc, inp, aug, rad)[k] for k in keys)
        with orc.planted(plant):
            b = _brute_lines(orc, inp, _brute_options(use_cascade_auger=aug, use_cascade_radiative=rad), seeds)
        m = sum(b[k][0] for k in keys)
        se = math.sqrt(sum(b[k][1] ** 2 for k in keys))
        assert abs(m - v) > 4 * se + 0.03 * v, (plant, m, se, v)


def test_auger_table_reproduces_vacancy_transfer():
    """The provider's Auger transitions leave, per decay, exactly the
    vacancies its cascade-corrected cross sections assume (auger_KL, auger_KM,
    auger_LM): the expected count of each shell summed over the transition
    list, for every element; each vacancy's rates sum to one where the
    provider defines a transfer."""
    from xmimsim_amd import lines
    for Z in list(range(3, 93, 4)) + [92]:
        a = lines.atom(Z)
        kl, km, lm = np.zeros(3), np.zeros(5), np.zeros((3, 5))
        tot = np.zeros(4)
        for r, s0, s1, s2 in lines.auger_transitions(Z):
            tot[s0] += r
            for s in (s1, s2):
                if s0 == 0 and 1 <= s <= 3:
                    kl[s - 1] += r
                elif s0 == 0 and 4 <= s <= 8:
                    km[s - 4] += r
                elif s0 >= 1 and 4 <= s <= 8:
                    lm[s0 - 1, s - 4] += r
        assert np.allclose(kl, a["auger_KL"], atol=1e-12), Z
        assert np.allclose(km, a["auger_KM"], atol=1e-12), Z
        assert np.allclose(lm, a["auger_LM"].reshape(3, 5), atol=1e-12), Z
        assert tot[0] == pytest.approx(1.0) or (a["auger_KL"].sum() + a["auger_KM"].sum() == 0)
        for i in range(3):
            assert tot[1 + i] == pytest.approx(1.0) or a["auger_LM"].reshape(3, 5)[i].sum() == 0


def test_brute_compton_polarisation(orc):
    """First-order Compton photons of a fully polarised beam in a plastic
    scatterer: the brute-force count over a detector of half-angle alpha is
    the estimator's value at the detector's centre times the mean of the
    Thomson polarisation factor 1 - (k.e)^2 over the detector's cone (the
    centre at k.e = 0): 1 - <sin^2 psi>/2.  With alpha = 44 degrees that
    factor is 0.87, which tells a polarisation-blind or wrongly rotated
    transport apart."""
    n = 2_000_000
    inp = Input(layers=[Layer([6, 8], [0.55, 0.45], 1.2, 0.1)], discrete=[DiscreteLine(20.0, 1e9, 0.0)],
                n_photons_line=n, n_interactions_trajectory=1, area_detector=3.0, gain=0.02, nchannels=1200)
    prep = api.Prepared(inp, options=_brute_options())
    bf = np.array([r[2] for r in _runs(orc, prep, range(8))])
    d = orc.input_derived(prep)
    prep.close()
    b = bf[:, [5, 7], 384, 0].sum(axis=1)
    m, se = b.mean(), b.std(ddof=1) / math.sqrt(len(b))
    prep = api.Prepared(inp, options=api.default_options())
    orc.solid_angle(prep, 256, 5)
    _, vr, _, _ = orc.run(prep, seed=1)
    prep.close()
    v = vr[[5, 7], 384, 0].sum()
    D = np.linalg.norm(np.array(inp.p_detector_window) - np.array([0.0, 0.0, inp.d_sample_source]))
    ca = math.cos(math.atan(d["detector_radius"] / D))
    mean_sin2 = (2.0 / 3.0 - ca + ca ** 3 / 3.0) / (1.0 - ca)
    factor = 1.0 - mean_sin2 / 2.0
    assert 0.85 < factor < 0.89
    assert abs(m - factor * v) < 4 * se + 0.03 * v, (m, se, v, factor)
