"""GPU test of the reference harness's file path (bin/xmimsim.c): an .xmsi
input (the reference's examples/srm1155.xmsi, photons reduced) through the
engine and the detector response into an .xmso, read back and compared with
the engine's arrays."""
import gzip
import os

import numpy as np
import pytest

from xmimsim_amd import api, xmlio

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_xmsi_to_xmso(tmp_path):
    src = tmp_path / "srm1155.xmsi"
    src.write_bytes(gzip.decompress(open(os.path.join(GOLD, "srm1155.xmsi.gz"), "rb").read()))
    inp = xmlio.read_xmsi(str(src))
    inp.n_photons_line = 20000
    print("srm1155: layers", [len(l.Z) for l in inp.layers], "lines", len(inp.discrete), flush=True)
    prep = api.Prepared(inp)
    prep.compute_solid_angle(hits_per_single=100, seed=11)
    ch, vr, _, stats = prep.run(seed=12)
    assert stats.histories == 20000 * len(inp.discrete)
    o = api.default_options()
    o.use_escape_peaks, o.use_default_seeds = 0, 1
    vr_det = vr.copy()
    conv = api.detector_response(prep, ch, vr_det, options=o)
    out = str(tmp_path / "srm1155.xmso")
    xmlio.write_xmso(out, inp, ch, conv, vr_det, inputfile=str(src))
    back = xmlio.read_xmso(out)
    np.testing.assert_allclose(back.channels_unconv[1:], ch[1:], rtol=5e-6, atol=0)
    np.testing.assert_allclose(back.channels_conv[1:], conv[1:], rtol=5e-6, atol=0)
    # every element of the sample with counts, every line and order with counts
    n_int = inp.n_interactions_trajectory
    zs = sorted({z for l in inp.layers for z in l.Z})
    expect = {}
    for Z in zs:
        rows = {}
        for j in range(383):
            orders = {k + 1: vr_det[Z - 1, j, k] for k in range(n_int) if vr_det[Z - 1, j, k] > 0}
            if vr_det[Z - 1, j].sum() != 0:
                rows[api.load_library().xmi_line_name(j + 1).decode()] = orders
        if rows:
            expect[Z] = rows
    assert set(back.var_red_history) == set(expect)
    assert {z: 0 for z in (26, 28, 24)}.keys() <= set(expect)  # Fe, Ni, Cr lines of the steel
    for Z, rows in expect.items():
        assert set(back.var_red_history[Z]) == set(rows)
        for name, orders in rows.items():
            got = back.var_red_history[Z][name]
            assert set(got) == set(orders)
            for k, v in orders.items():
                assert got[k] == pytest.approx(v, rel=5e-6)
    prep.close()
