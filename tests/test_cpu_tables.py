"""CPU tests of the physics provider, the interaction tables and the C ABI
library (loads and exports every declared symbol; no device calls)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from xmimsim_amd import api, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(lib):
    header = open(os.path.join(ROOT, "include", "xmimsim_amd.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)                 # comments
    header = re.sub(r"typedef\s+(struct|enum)[^{;]*\{.*?\}[^;]*;", "", header, flags=re.S)  # types
    declared = set(re.findall(r"\b(xmi_\w+)\s*\(", header))
    declared |= set(re.findall(r"^\s*extern\s+[\w\s\*]+?\b(xmi_\w+)\s*;", header, re.M))
    declared -= {"xmi_inputFPtr", "xmi_hdf5FPtr"}
    assert declared, "no declarations parsed"
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(api.EXPORTED_SYMBOLS) <= declared


def test_default_options(lib):
    o = api.default_options()
    assert (o.use_M_lines, o.use_cascade_auger, o.use_cascade_radiative, o.use_variance_reduction) == (1, 1, 1, 1)
    assert (o.use_escape_peaks, o.use_advanced_compton, o.use_default_seeds) == (1, 0, 0)


def test_struct_layouts_match_reference_c_types():
    # sizes of the reference's public structs on x86-64 (same member order/types)
    assert C.sizeof(api.xmi_layer) == 40
    assert C.sizeof(api.xmi_geometry) == 16 * 8
    assert C.sizeof(api.xmi_energy_discrete) == 9 * 8
    assert C.sizeof(api.xmi_main_options) == 12 * 4 + 8 + 2 * 4
    assert C.sizeof(api.xmi_solid_angle) == 6 * 8


def test_prepare_and_geometry(lib, orc):
    """xmi_init_input on the steel input: detector radius sqrt(A/pi) and the
    on-axis solid angle 2 pi (1 - cos atan(R/D)) with D the distance from the
    sample point (0, 0, d_sample_source) to the window (the general case,
    with collimator, tilted sample and a middle reference layer, is pinned in
    test_cpu_oracle_pins.py)"""
    inp = workloads.single_layer_steel(n_photons_line=10)
    prep = api.Prepared(inp)
    g = orc.input_derived(prep)
    R = math.sqrt(inp.area_detector / math.pi)
    assert g["detector_radius"] == pytest.approx(R, rel=1e-15)
    D = math.dist(inp.p_detector_window, (0.0, 0.0, inp.d_sample_source))
    assert g["detector_solid_angle"] == pytest.approx(2 * math.pi * (1 - math.cos(math.atan(R / D))), rel=1e-13)
    assert g["collimator_present"] == 0
    assert g["Z_coord_begin"][0] == pytest.approx(inp.d_sample_source)
    assert g["Z_coord_end"][0] == pytest.approx(inp.d_sample_source + 0.5 / math.sqrt(0.5), rel=1e-12)
    prep.close()


def test_interaction_order_bound_is_an_error(lib):
    """The random-stream counter keeps the interaction order in 8 bits:
    more than 255 orders per trajectory is refused (before any device work)"""
    prep = api.Prepared(workloads.single_layer_steel(n_photons_line=10, n_interactions=256))
    grid = api.xmi_solid_angle()
    with pytest.raises(RuntimeError, match="n_interactions_trajectory must be in 1..255"):
        prep.solid_angle = C.pointer(grid)
        prep._sa_owner = "test"
        prep.run(seed=1)
    prep.solid_angle = None
    prep.close()


def test_line_numbering(lib):
    """xraylib line numbering used for the history rows (KL1=1, KL3=3,
    KP5=29, L3M5=90, M5P5=219, P3P5=383, 384 Rayleigh, 385 Compton)."""
    import xmimsim_amd.lines as L
    names = L.line_names()
    assert len(names) == 386
    assert names[1] == "KL1" and names[3] == "KL3" and names[29] == "KP5"
    assert names[90] == "L3M5" and names[219] == "M5P5" and names[383] == "P3P5"
    assert names[384] == "Rayleigh" and names[385] == "Compton"
    assert len(set(names[1:384])) == 383


def test_provider_invariants():
    import xmimsim_amd.lines as L
    for Z in (6, 14, 26, 42, 56, 79, 92):
        a = L.atom(Z)
        edges = a["edge"][:9]
        present = edges[edges > 0]
        assert np.all(np.diff(present) < 0), (Z, edges)
        fy = a["fluor_yield"]
        assert np.all((fy >= 0) & (fy <= 1))
        for shell, (first, last) in enumerate(L.SHELL_LINES):
            s = a["rad_rate"][first:last + 1].sum()
            assert s == pytest.approx(1.0, abs=1e-12) or s == 0.0, (Z, shell, s)
        # photo cross section: K jump ratio is the model's 125/Z + 3.5
        eK = edges[0]
        if eK > 0.2:
            hi = L.cs_photo_total(Z, eK * (1 + 1e-9))
            lo = L.cs_photo_total(Z, eK * (1 - 1e-9))
            assert hi / lo == pytest.approx(125.0 / Z + 3.5, rel=1e-6)


def test_rayleigh_compton_tables_integrate_dcs():
    """The tabulated integrated cross sections equal a fresh numerical
    integration of the differential cross sections (independent quadrature)."""
    import xmimsim_amd.lines as L
    th = np.linspace(0, math.pi, 20001)
    for Z in (8, 26, 82):
        for E in (1.5, 10.0, 33.0, 120.0):
            for kind in ("rayl", "compt"):
                f = np.array([L.dcs(kind, Z, E, t) for t in th[::10]]) * np.sin(th[::10])
                integ = 2 * math.pi * np.trapezoid(f, th[::10])
                assert L.cs(kind, Z, E) == pytest.approx(integ, rel=2e-4), (Z, E, kind)


def test_generated_tables_invert_their_cdfs(lib):
    """Angle ICDFs (Rayleigh), azimuth ICDF and Compton-profile ICDF are the
    inverses of independently evaluated CDFs."""
    import xmimsim_amd.lines as L
    t = L.tables_for([26])
    E = t["Energies"]
    rs = t["RandomNumbers"]
    th = np.linspace(0, math.pi, 40001)
    for ei in (20, 150, 390):
        f = np.array([L.dcs("rayl", 26, E[ei], x) for x in th]) * np.sin(th)
        cdf = np.concatenate([[0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(th))])
        cdf /= cdf[-1]
        for ri in (100, 1000, 1900):
            theta = t["RayleighTheta_ICDF"][ei, ri]
            assert np.interp(theta, th, cdf) == pytest.approx(rs[ri], abs=2e-3)
    # azimuth: CDF(phi) = (phi - t sin 2 phi) / (2 pi)
    phi = t["Phi_ICDF"]
    T = t["Thetas"]
    for ti in (0, 77, 199):
        for ri in (1, 500, 1999):
            p = phi[ti, ri]
            assert (p - T[ti] * math.sin(2 * p)) / (2 * math.pi) == pytest.approx(rs[ri], abs=1e-9)
    # interaction probabilities
    e = t["ip_energies"]
    k = len(e) // 3
    assert t["ip_rayl"][k] == pytest.approx(L.cs("rayl", 26, e[k]) / L.cs_total(26, e[k]), rel=1e-12)


def test_hdf5_roundtrip(lib, tmp_path):
    """xmi_db writes the xmimsimdata.h5 layout; xmi_init_from_hdf5 reads back
    exactly the tables the in-memory provider path builds."""
    inp = workloads.single_layer_steel(n_photons_line=10)
    fn = str(tmp_path / "xmimsimdata-test.h5")
    Zs = (C.c_int * 4)(*inp.unique_Z())
    assert lib.xmi_db(fn.encode(), Zs, 4) == 1
    a = api.Prepared(inp, hdf5_file=fn)
    b = api.Prepared(inp)
    import xmimsim_amd.lines as L
    ta, tb = L.tables_of(a), L.tables_of(b)
    for k in ta:
        assert np.array_equal(ta[k], tb[k]), k
    a.close()
    b.close()


def test_missing_element_is_an_error(lib, tmp_path):
    fn = str(tmp_path / "db.h5")
    Zs = (C.c_int * 1)(26)
    assert lib.xmi_db(fn.encode(), Zs, 1) == 1
    with pytest.raises(RuntimeError):
        api.Prepared(workloads.single_layer_steel(n_photons_line=10), hdf5_file=fn)


def test_no_gpu_is_a_loud_error(lib):
    """Without a HIP device the engine refuses instead of falling back."""
    import ctypes as C2
    try:
        n = C2.c_int()
        hip = C2.CDLL("libamdhip64.so")
        if hip.hipGetDeviceCount(C2.byref(n)) == 0 and n.value > 0:
            pytest.skip("a device is present")
    except OSError:
        pass
    prep = api.Prepared(workloads.single_layer_steel(n_photons_line=10))
    with pytest.raises(RuntimeError):
        prep.compute_solid_angle(hits_per_single=10)


def test_host_table_generator_properties(lib):
    """Host generator (the one the GPU generator is compared with): inverse
    CDFs are monotone with the documented end points, and the angle inverse
    CDF of an element agrees with an independent inversion of its DCS."""
    lib.xmi_tables_set_device(0)
    try:
        T = api.SharedTables([14, 82])
        for Z in (14, 82):
            for which in (0, 1):
                t = T.table(Z, which).reshape(400, 2000)
                assert np.all(np.diff(t, axis=1) >= 0.0)
                assert np.all(t[:, 0] == 0.0) and np.all(t[:, -1] == math.pi)
            cp = T.table(Z, 2)
            assert cp[0] == 0.0 and cp[-1] == 100.0 and np.all(np.diff(cp) >= 0.0)
            qs = T.table(Z, 3).reshape(-1, 10001)
            assert np.all(qs[:, 0] == 0.0) and np.all(qs[:, -1] == 100.0)
        phi = T.table(14, 4).reshape(200, 2000)
        assert np.all(phi[:, 0] == 0.0) and np.allclose(phi[:, -1], 2.0 * math.pi)
        # the azimuth CDF (phi - t sin 2 phi) / 2 pi at the table's phi values
        t = np.linspace(0.0, 0.5, 200)[:, None]
        r = np.linspace(0.0, 1.0, 2000)[None, :]
        assert np.max(np.abs((phi - t * np.sin(2.0 * phi)) / (2.0 * math.pi) - r)) < 1e-12
        # Rayleigh angle ICDF of Si at 20 keV (energy row 39.9..: pick the row
        # nearest 20 keV) against a direct trapezoid inversion in numpy
        E = np.linspace(0.1, 200.0, 400)
        j = int(np.argmin(np.abs(E - 20.0)))
        th = np.linspace(0.0, math.pi, 10000)
        f = np.array([lib.xmi_provider_dcs(0, 14, E[j], x) for x in th]) * np.sin(th)
        cdf = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * (th[1] - th[0]))])
        want = np.interp(r[0] * cdf[-1], cdf, th)
        got = T.table(14, 0).reshape(400, 2000)[j]
        assert np.max(np.abs(got[1:-1] - want[1:-1])) < 1e-6
        T.close()
    finally:
        lib.xmi_tables_set_device(-1)


def test_bench_refuses_several_gpus():
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
