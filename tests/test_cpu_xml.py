"""CPU tests of the XMSI/XMSO files (src/xmi_xml.c, src/xmi_data_structs.c:1371-1519
of the reference), through the library's C ABI.

Fixtures: the reference's example inputs (examples/*.xmsi) and two of its
example outputs (examples/In.xmso, examples/srm1155.xmso), gzip'ed by
tests/golden/make_xml_fixtures.sh.  Values are checked against an
independent parse of the same files with Python's xml.etree, and the writer
against the reference's own text of the outputs.
"""
import gzip
import math
import os
import re
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from xmimsim_amd import api, workloads, xmlio

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
XMSI = ["In", "srm1132", "srm1155", "srm1412"]


def _unzip(name, tmp_path):
    dst = tmp_path / name
    dst.write_bytes(gzip.decompress(open(os.path.join(GOLD, name + ".gz"), "rb").read()))
    return str(dst)


def _f(node, path):
    return float(node.find(path).text)


@pytest.mark.parametrize("name", XMSI)
def test_read_reference_inputs(lib, tmp_path, name):
    path = _unzip(name + ".xmsi", tmp_path)
    inp = xmlio.read_xmsi(path)
    root = ET.parse(path).getroot()
    gen = root.find("general")
    assert inp.outputfile == gen.find("outputfile").text
    assert inp.n_photons_interval == int(gen.find("n_photons_interval").text)
    assert inp.n_photons_line == int(gen.find("n_photons_line").text)
    assert inp.n_interactions_trajectory == int(gen.find("n_interactions_trajectory").text)
    # layers: zero weights dropped, sorted by Z, normalised
    xl = root.find("composition").findall("layer")
    assert len(inp.layers) == len(xl)
    for l, x in zip(inp.layers, xl):
        pairs = sorted((int(e.find("atomic_number").text), float(e.find("weight_fraction").text))
                       for e in x.findall("element") if abs(float(e.find("weight_fraction").text)) >= 1e-20)
        assert list(l.Z) == [z for z, _ in pairs]
        tot = sum(w for _, w in pairs)
        np.testing.assert_allclose(l.weight, [w / tot for _, w in pairs], rtol=1e-14)
        assert l.density == _f(x, "density") and l.thickness == _f(x, "thickness")
    assert inp.reference_layer == int(root.find("composition/reference_layer").text)
    # orientations normalised
    geo = root.find("geometry")
    for key in ("n_sample_orientation", "n_detector_orientation"):
        v = np.array([_f(geo, f"{key}/{c}") for c in "xyz"])
        np.testing.assert_allclose(getattr(inp, key), v / np.linalg.norm(v), rtol=1e-14)
    assert tuple(inp.p_detector_window) == tuple(_f(geo, f"p_detector_window/{c}") for c in "xyz")
    # discrete energies: merged when repeated, sorted
    disc = {}
    for d in root.find("excitation").findall("discrete"):
        E = _f(d, "energy")
        h, v = disc.get(E, (0.0, 0.0))
        disc[E] = (h + _f(d, "horizontal_intensity"), v + _f(d, "vertical_intensity"))
    assert [d.energy for d in inp.discrete] == sorted(disc)
    for d in inp.discrete:
        assert (d.horizontal_intensity, d.vertical_intensity) == pytest.approx(disc[d.energy], rel=1e-14)
    det = root.find("detector")
    assert inp.nchannels == (int(det.find("nchannels").text) if det.find("nchannels") is not None else 2048)
    assert inp.detector_type == ["SiLi", "Ge", "Si_SDD"].index(det.find("detector_type").text)
    assert (inp.gain, inp.zero, inp.fano, inp.noise) == tuple(_f(det, k) for k in ("gain", "zero", "fano", "noise"))
    assert len(inp.crystal) == len(det.findall("crystal/layer"))


@pytest.mark.parametrize("name", XMSI)
def test_input_round_trip(lib, tmp_path, name):
    a = xmlio.read_xmsi(_unzip(name + ".xmsi", tmp_path))
    out = str(tmp_path / "rt.xmsi")
    xmlio.write_xmsi(a, out)
    b = xmlio.read_xmsi(out)
    c = xmlio.read_xmsi_string(xmlio.xmsi_string(a))
    for x in (b, c):
        # numbers pass through %g (6 significant digits)
        for fa, fb in zip(_flat(a), _flat(x)):
            assert fa == pytest.approx(fb, rel=2e-6, abs=1e-300), (fa, fb)
    assert '<xmimsim>' in xmlio.xmsi_string(a) and 'version="8.1"' in xmlio.xmsi_string(a)


def _flat(inp):
    out = []
    for v in inp.__dict__.values():
        if isinstance(v, (int, float)):
            out.append(float(v))
        elif isinstance(v, (tuple, list)):
            for e in v:
                if isinstance(e, (int, float)):
                    out.append(float(e))
                elif hasattr(e, "__dict__"):
                    for w in e.__dict__.values():
                        out += [float(t) for t in (w if isinstance(w, (list, tuple)) else [w])]
    return out


def _lines(path):
    t = re.sub(r"<!--.*?-->", "", open(path).read(), flags=re.S)
    lines = [re.sub(r' version="[^"]*"', ' version="V"', l.strip()) for l in t.splitlines() if l.strip()]
    i = lines.index("<svg_graphs>")
    return lines[:i], lines[i:]


@pytest.mark.parametrize("name", ["In", "srm1155"])
def test_output_writer_reproduces_reference_text(lib, tmp_path, name):
    """Read the reference's .xmso and write it again: every element outside
    the preview graphs comes out with the reference's text (%g numbers,
    attribute order, names, empty elements); the graphs have the same
    elements, and their plot coordinates agree to 2e-4 plus one unit of the
    sixth significant digit (the graphs are drawn from counts that the file
    holds to 6 digits, and are printed with 6)."""
    src = _unzip(name + ".xmso", tmp_path)
    dst = str(tmp_path / "out.xmso")
    xmlio.rewrite_xmso(src, dst)
    (a, sa), (b, sb) = _lines(src), _lines(dst)
    assert a == b
    assert len(sa) == len(sb)
    tag = lambda l: re.sub(r">[^<]*<", "><", l)
    assert [tag(l) for l in sa] == [tag(l) for l in sb]
    num = re.compile(r">([-0-9.e+]+)<")
    for la, lb in zip(sa, sb):
        for u, v in zip(num.findall(la), num.findall(lb)):
            digit = 10.0 ** (math.floor(math.log10(abs(float(u)))) - 5) if float(u) else 0.0
            assert abs(float(u) - float(v)) <= 2e-4 + 1.001 * digit, (la, lb)
    head = open(dst).read(400)
    assert head.startswith('<?xml version="1.0"?>\n<!DOCTYPE xmimsim-results SYSTEM '
                           '"http://www.xmi.UGent.be/xml/xmimsim-1.0.dtd">')
    assert "DO NOT MODIFY THIS FILE UNLESS YOU KNOW WHAT YOU ARE DOING!" in head


def test_output_reader_against_etree(lib, tmp_path):
    src = _unzip("In.xmso", tmp_path)
    o = xmlio.read_xmso(src)
    root = ET.parse(src).getroot()
    chans = root.find("spectrum_conv").findall("channel")
    assert o.channels_conv.shape == (5, len(chans))
    for c in chans[::97]:
        j = int(c.find("channelnr").text)
        for k in c.findall("counts"):
            assert o.channels_conv[int(k.get("interaction_number")), j] == float(k.text)
    assert not o.use_zero_interactions and not o.channels_conv[0].any()
    assert o.inputfile == "In.xmsi" and o.version == pytest.approx(1.0)
    for e in root.find("variance_reduction_history").findall("fluorescence_line_counts"):
        Z = int(e.get("atomic_number"))
        assert e.get("symbol") == lib.xmi_element_symbol(Z).decode()
        assert o.element_totals[Z] == float(e.get("total_counts"))
        for l in e.findall("fluorescence_line"):
            got = o.var_red_history[Z][l.get("type")]
            assert got == {int(k.get("interaction_number")): float(k.text) for k in l.findall("counts")}
            assert o.line_energies[Z][l.get("type")] == float(l.get("energy"))
            # a line's total is the sum of its orders, to the file's 6 digits
            assert sum(got.values()) == pytest.approx(float(l.get("total_counts")), rel=1e-5)
    assert o.brute_force_history == {}


def test_raw2struct_and_written_record(lib, tmp_path):
    inp = workloads.single_layer_steel(n_photons_line=1000)
    inp.outputfile = "steel.xmso"
    inp.comments = "a <comment> & more"
    n_int, nch = inp.n_interactions_trajectory, inp.nchannels
    rng = np.random.default_rng(5)
    unconv = np.cumsum(rng.random((n_int + 1, nch)) * (rng.random((n_int + 1, nch)) > 0.3), axis=0)
    conv = unconv * 0.9
    vr = np.zeros((100, 385, n_int))
    vr[25, 2, :] = [5.0, 1.0, 0.0, 0.25]     # Fe KL3, order 3 empty
    vr[25, 383, :] = [2.0, 0.0, 0.0, 0.0]    # Fe row 384 (Rayleigh): not a line, left out
    vr[23, 0, 1] = 3.5                       # Cr KL1, order 2 only
    vr[50, 2, 0] = 9.0                       # Sn: not in the sample, not written
    path = str(tmp_path / "steel.xmso")
    rec = xmlio.write_xmso(path, inp, unconv, conv, vr, inputfile="steel.xmsi")
    back = xmlio.read_xmso(path)
    for o in (rec, back):
        assert set(o.var_red_history) == {24, 26}
        assert o.var_red_history[26] == {"KL3": {1: 5.0, 2: 1.0, 4: 0.25}}
        assert o.var_red_history[24] == {"KL1": {2: 3.5}}
        assert o.element_totals[26] == 6.25 and o.brute_force_history == {}
        assert o.line_energies[26]["KL3"] == pytest.approx(lib.xmi_provider_line_data(26, 2, 3), rel=1e-5)
    # orders 1..n as written (%g, 6 digits); order 0 is not written
    np.testing.assert_allclose(back.channels_unconv[1:], unconv[1:], rtol=5e-6, atol=0)
    np.testing.assert_allclose(back.channels_conv[1:], conv[1:], rtol=5e-6, atol=0)
    assert not back.channels_unconv[0].any()
    assert back.input.comments == inp.comments and back.input.outputfile == "steel.xmso"
    text = open(path).read()
    assert "&lt;comment&gt; &amp; more" in text
    assert text.count("<graphic>") == 2 * n_int


def test_invalid_inputs_raise(lib, tmp_path):
    good = xmlio.xmsi_string(workloads.single_layer_steel(n_photons_line=10))
    assert xmlio.read_xmsi_string(good.replace("<outputfile/>", "<outputfile>x.xmso</outputfile>"))
    cases = {
        "reference_layer": good.replace("<reference_layer>1</reference_layer>",
                                        "<reference_layer>3</reference_layer>"),
        "detector_type": good.replace("<detector_type>SiLi</detector_type>", "<detector_type>CdTe</detector_type>"),
        "not well-formed": good.replace("</general>", ""),
        "wrong root": good.replace("<xmimsim>", "<xmimsim-results>").replace("</xmimsim>", "</xmimsim-results>"),
        "validation": good,  # empty outputfile fails xmi_input_validate's general section
        "n_photons_line": good.replace("<n_photons_line>10</n_photons_line>", "<n_photons_line>ten</n_photons_line>"),
    }
    for what, text in cases.items():
        with pytest.raises(RuntimeError) as e:
            xmlio.read_xmsi_string(text)
        assert str(e.value), what


def test_continuous_spectrum_rules(lib):
    base = workloads.single_layer_steel(n_photons_line=10)
    base.outputfile = "o.xmso"
    base.discrete = []
    pts = [api.ContinuousPoint(E, I) for E, I in [(1.0, 1.0), (2.0, 0.0), (3.0, 0.0), (4.0, 0.0), (5.0, 1.0)]]
    base.continuous = pts
    assert xmlio.validate(base) == 8  # three consecutive zero densities: XMI_INPUT_EXCITATION
    with pytest.raises(RuntimeError):
        xmlio.read_xmsi_string(xmlio.xmsi_string(base))
    base.continuous = [api.ContinuousPoint(E, 1.0) for E in (1.0, 2.0, 3.0)]
    assert xmlio.validate(base) == 0
    text = xmlio.xmsi_string(base)
    dup = text.replace("<energy>3</energy>", "<energy>2</energy>")
    with pytest.raises(RuntimeError, match="given twice"):
        xmlio.read_xmsi_string(dup)
    base.layers[0].density = 0.0
    assert xmlio.validate(base) == 2  # XMI_INPUT_COMPOSITION


def test_error_struct_has_gerror_layout(lib):
    import ctypes as C
    err = C.POINTER(api.xmi_error)()
    p = lib.xmi_input_read_from_xml_file(b"/nonexistent/file.xmsi", C.byref(err))
    assert not p and err and err.contents.message and err.contents.code == 1
    lib.xmi_error_free(err)
    assert C.sizeof(api.xmi_error) == 16 and api.xmi_error.message.offset == 8


def test_element_symbols(lib):
    assert [lib.xmi_element_symbol(z).decode() for z in (1, 8, 26, 49, 82, 92, 94)] == \
        ["H", "O", "Fe", "In", "Pb", "U", "Pu"]
    assert lib.xmi_element_symbol(0) is None and lib.xmi_element_symbol(200) is None
