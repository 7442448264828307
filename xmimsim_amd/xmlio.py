"""XMSI input and XMSO output files through the library's C ABI.

Python side of xmi_input_read_from_xml_file / xmi_input_write_to_xml_file
(src/xmi_xml.c:1286, :1404), xmi_output_raw2struct
(src/xmi_data_structs.c:1371-1519) and xmi_output_write_to_xml_file /
xmi_output_read_from_xml_file (src/xmi_xml.c:1638, :2033).  Errors raise
RuntimeError with the library's message.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from . import api


def _raise(err, what: str):
    lib = api.load_library()
    msg = err.contents.message.decode() if err else lib.xmi_msim_last_error().decode()
    if err:
        lib.xmi_error_free(err)
    raise RuntimeError(f"{what}: {msg}")


def _layers(ptr, n) -> list:
    out = []
    for i in range(n):
        l = ptr[i]
        out.append(api.Layer([l.Z[k] for k in range(l.n_elements)], [l.weight[k] for k in range(l.n_elements)],
                             l.density, l.thickness))
    return out


def input_from_c(p) -> api.Input:
    """api.Input from an xmi_input * (the C struct is not freed)."""
    s = p.contents
    g, c, ge, x, a, d = (s.general.contents, s.composition.contents, s.geometry.contents, s.excitation.contents,
                         s.absorbers.contents, s.detector.contents)
    disc = [api.DiscreteLine(e.energy, e.horizontal_intensity, e.vertical_intensity, e.sigma_x, e.sigma_xp,
                             e.sigma_y, e.sigma_yp, e.distribution_type, e.scale_parameter)
            for e in (x.discrete[i] for i in range(x.n_discrete))]
    cont = [api.ContinuousPoint(e.energy, e.horizontal_intensity, e.vertical_intensity, e.sigma_x, e.sigma_xp,
                                e.sigma_y, e.sigma_yp)
            for e in (x.continuous[i] for i in range(x.n_continuous))]
    return api.Input(
        layers=_layers(c.layers, c.n_layers), reference_layer=c.reference_layer,
        outputfile=(g.outputfile or b"").decode(), comments=(g.comments or b"").decode(),
        discrete=disc, continuous=cont, n_photons_interval=g.n_photons_interval, n_photons_line=g.n_photons_line,
        n_interactions_trajectory=g.n_interactions_trajectory, d_sample_source=ge.d_sample_source,
        n_sample_orientation=tuple(ge.n_sample_orientation), p_detector_window=tuple(ge.p_detector_window),
        n_detector_orientation=tuple(ge.n_detector_orientation), area_detector=ge.area_detector,
        collimator_height=ge.collimator_height, collimator_diameter=ge.collimator_diameter,
        d_source_slit=ge.d_source_slit, slit_size_x=ge.slit_size_x, slit_size_y=ge.slit_size_y,
        exc_absorbers=_layers(a.exc_layers, a.n_exc_layers), det_absorbers=_layers(a.det_layers, a.n_det_layers),
        detector_type=d.detector_type, live_time=d.live_time, pulse_width=d.pulse_width, gain=d.gain, zero=d.zero,
        fano=d.fano, noise=d.noise, nchannels=d.nchannels, crystal=_layers(d.crystal_layers, d.n_crystal_layers))


def read_xmsi(path: str) -> api.Input:
    lib = api.load_library()
    err = C.POINTER(api.xmi_error)()
    p = lib.xmi_input_read_from_xml_file(path.encode(), C.byref(err))
    if not p:
        _raise(err, path)
    try:
        return input_from_c(p)
    finally:
        lib.xmi_input_free(p)


def read_xmsi_string(text: str) -> api.Input:
    lib = api.load_library()
    err = C.POINTER(api.xmi_error)()
    p = lib.xmi_input_read_from_xml_string(text.encode(), C.byref(err))
    if not p:
        _raise(err, "xmsi string")
    try:
        return input_from_c(p)
    finally:
        lib.xmi_input_free(p)


def validate(inp: api.Input) -> int:
    """xmi_input_validate: 0, or the OR of the XMI_INPUT_* flags of the failing sections."""
    ci = api.CInput(inp)
    return api.load_library().xmi_input_validate(C.byref(ci.input))


def write_xmsi(inp: api.Input, path: str) -> None:
    lib = api.load_library()
    ci = api.CInput(inp)
    err = C.POINTER(api.xmi_error)()
    if not lib.xmi_input_write_to_xml_file(C.byref(ci.input), path.encode(), C.byref(err)):
        _raise(err, path)


def xmsi_string(inp: api.Input) -> str:
    lib = api.load_library()
    ci = api.CInput(inp)
    err = C.POINTER(api.xmi_error)()
    out = C.c_void_p()
    if not lib.xmi_input_write_to_xml_string(C.byref(ci.input), C.byref(out), C.byref(err)):
        _raise(err, "xmsi string")
    try:
        return C.string_at(out.value).decode()
    finally:
        lib.xmi_free(out)


@dataclass
class Xmso:
    """Contents of an .xmso file; histories map Z -> line name -> {order: counts}."""
    input: api.Input
    version: float
    inputfile: str
    channels_conv: np.ndarray      # [n_interactions + 1][nchannels]
    channels_unconv: np.ndarray    # [n_interactions + 1][nchannels]
    use_zero_interactions: bool
    var_red_history: Dict[int, Dict[str, Dict[int, float]]] = field(default_factory=dict)
    brute_force_history: Dict[int, Dict[str, Dict[int, float]]] = field(default_factory=dict)
    line_energies: Dict[int, Dict[str, float]] = field(default_factory=dict)
    element_totals: Dict[int, float] = field(default_factory=dict)


def _history(ptr, n, energies, totals) -> dict:
    out = {}
    for i in range(n):
        e = ptr[i]
        lines = {}
        totals[e.atomic_number] = e.total_counts
        en = energies.setdefault(e.atomic_number, {})
        for j in range(e.n_lines):
            l = e.lines[j]
            name = l.line_type.decode()
            lines[name] = {l.interactions[k].interaction_number: l.interactions[k].counts
                           for k in range(l.n_interactions)}
            en[name] = l.energy
        out[e.atomic_number] = lines
    return out


def _output_from_c(p) -> Xmso:
    o = p.contents
    n_int, nch = o.ninteractions, o.input.contents.detector.contents.nchannels
    conv = np.array([[o.channels_conv[i][j] for j in range(nch)] for i in range(n_int + 1)])
    unconv = np.array([[o.channels_unconv[i][j] for j in range(nch)] for i in range(n_int + 1)])
    energies, totals = {}, {}
    vr = _history(o.var_red_history, o.nvar_red_history, energies, totals) if o.var_red_history else {}
    bf = _history(o.brute_force_history, o.nbrute_force_history, {}, {}) if o.brute_force_history else {}
    return Xmso(input_from_c(o.input), float(o.version), (o.inputfile or b"").decode(), conv, unconv,
                bool(o.use_zero_interactions), vr, bf, energies, totals)


def read_xmso(path: str) -> Xmso:
    lib = api.load_library()
    err = C.POINTER(api.xmi_error)()
    p = lib.xmi_output_read_from_xml_file(path.encode(), C.byref(err))
    if not p:
        _raise(err, path)
    try:
        return _output_from_c(p)
    finally:
        lib.xmi_output_free(p)


def write_xmso(path: str, inp: api.Input, channels_unconv: np.ndarray, channels_conv: Optional[np.ndarray] = None,
               var_red_history: Optional[np.ndarray] = None, brute_history: Optional[np.ndarray] = None,
               inputfile: str = "", use_zero_interactions: bool = False) -> Xmso:
    """xmi_output_raw2struct + xmi_output_write_to_xml_file.

    channels_unconv: [n_int + 1][nchannels] (xmi_main_msim's channels);
    channels_conv: the detector-convolved spectra, same shape (defaults to
    channels_unconv); histories: [100][385][n_int] (brute_history defaults to
    zeros, as with variance reduction on).  Returns the record as written."""
    lib = api.load_library()
    n_int, nch = inp.n_interactions_trajectory, inp.nchannels
    unconv = np.ascontiguousarray(channels_unconv, dtype=np.float64).reshape(n_int + 1, nch)
    conv = unconv if channels_conv is None else np.ascontiguousarray(channels_conv, dtype=np.float64)
    conv = conv.reshape(n_int + 1, nch)
    shape = (100, 385, n_int)
    bh = np.zeros(shape) if brute_history is None else np.ascontiguousarray(brute_history, dtype=np.float64)
    vr = None if var_red_history is None else np.ascontiguousarray(var_red_history, dtype=np.float64)
    if bh.shape != shape or (vr is not None and vr.shape != shape):
        raise ValueError(f"histories must have shape {shape}")
    rows = (C.POINTER(C.c_double) * (n_int + 1))(*[conv[i].ctypes.data_as(C.POINTER(C.c_double))
                                                    for i in range(n_int + 1)])
    ci = api.CInput(inp)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None
    p = lib.xmi_output_raw2struct(C.byref(ci.input), dp(bh), dp(vr), rows, dp(unconv), inputfile.encode(),
                                  int(use_zero_interactions))
    if not p:
        _raise(None, "xmi_output_raw2struct")
    try:
        err = C.POINTER(api.xmi_error)()
        if not lib.xmi_output_write_to_xml_file(p, path.encode(), C.byref(err)):
            _raise(err, path)
        return _output_from_c(p)
    finally:
        lib.xmi_output_free(p)


def rewrite_xmso(src: str, dst: str) -> None:
    """Reads an .xmso and writes the same record again (xmi_output_read_from_xml_file
    followed by xmi_output_write_to_xml_file, preview graphs redrawn)."""
    lib = api.load_library()
    err = C.POINTER(api.xmi_error)()
    p = lib.xmi_output_read_from_xml_file(src.encode(), C.byref(err))
    if not p:
        _raise(err, src)
    try:
        if not lib.xmi_output_write_to_xml_file(p, dst.encode(), C.byref(err)):
            _raise(err, dst)
    finally:
        lib.xmi_output_free(p)
