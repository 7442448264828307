"""The drop-in chain of the reference harness (bin/xmimsim.c:262-651) run by a
C program compiled against include/xmimsim_amd.h (tests/harness/xmimsim_chain.c):
XMSI -> C2F -> xmi_init_input -> tables written by xmi_db and read by
xmi_init_from_hdf5 -> solid-angle grid through its HDF5 cache ->
xmi_main_msim -> escape ratios through their cache -> detector response ->
xmi_output_raw2struct -> XMSO.  A second run finds both caches and writes the
same file; the unconvoluted spectra equal those of the same chain driven from
Python with the same seeds."""
import gzip
import os
import subprocess

import numpy as np
import pytest

from xmimsim_amd import api, xmlio

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "harness", "xmimsim_chain")


def _run(tmp_path, out, n_photons=20000):
    args = [BIN, str(tmp_path / "srm1155.xmsi"), str(tmp_path / out), str(tmp_path / "tables.h5"),
            str(tmp_path / "grids.h5"), str(tmp_path / "escape.h5"), str(n_photons), "11", "12", "100", "20000"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_c_harness_runs_the_reference_chain(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("tests/harness/xmimsim_chain is not built (make -C tests/harness)")
    src = tmp_path / "srm1155.xmsi"
    src.write_bytes(gzip.decompress(open(os.path.join(HERE, "golden", "srm1155.xmsi.gz"), "rb").read()))
    first = _run(tmp_path, "a.xmso")
    assert "solid-angle grid computed and stored" in first and "escape ratios computed and stored" in first
    second = _run(tmp_path, "b.xmso")
    assert "solid-angle grid found" in second and "escape ratios found" in second
    a, b = (tmp_path / "a.xmso").read_text(), (tmp_path / "b.xmso").read_text()
    strip = lambda t: "\n".join(l for l in t.splitlines()
                                if not any(k in l for k in ("inputfile", "outputfile", "Timestamp")))
    assert strip(a) == strip(b)

    back = xmlio.read_xmso(str(tmp_path / "a.xmso"))
    inp = xmlio.read_xmsi(str(src))
    inp.n_photons_line = 20000
    prep = api.Prepared(inp, hdf5_file=str(tmp_path / "tables.h5"))
    prep.compute_solid_angle(hits_per_single=100, seed=11)
    ch, _, _, st = prep.run(seed=12)
    assert st.histories == 20000 * len(inp.discrete)
    np.testing.assert_allclose(back.channels_unconv[1:], ch[1:], rtol=5e-6, atol=0)
    assert back.channels_conv[1:].sum() > 0
    prep.close()
