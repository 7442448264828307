"""HDF5 caches of solid-angle grids and escape ratios through the C ABI.

The caller side of xmi_find_solid_angle_match / xmi_update_solid_angle_hdf5_file
(src/xmi_solid_angle.c:236-785) and xmi_find_escape_ratios_match /
xmi_update_escape_ratios_hdf5_file (src/xmi_detector.c:174-555), in the order
bin/xmimsim.c uses them: look up, compute on a miss, store.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Optional

from . import api, xmlio

_declared = False


def _lib():
    global _declared
    lib = api.load_library()
    if not _declared:
        P = C.POINTER
        I = P(api.xmi_input)
        for name in ("xmi_create_empty_solid_angle_hdf5_file", "xmi_create_empty_escape_ratios_hdf5_file"):
            getattr(lib, name).argtypes = [C.c_char_p]
            getattr(lib, name).restype = C.c_int
        lib.xmi_update_solid_angle_hdf5_file.argtypes = [C.c_char_p, P(api.xmi_solid_angle)]
        lib.xmi_update_solid_angle_hdf5_file.restype = C.c_int
        lib.xmi_update_escape_ratios_hdf5_file.argtypes = [C.c_char_p, P(api.xmi_escape_ratios)]
        lib.xmi_update_escape_ratios_hdf5_file.restype = C.c_int
        for name in ("xmi_check_solid_angle_match", "xmi_check_escape_ratios_match"):
            getattr(lib, name).argtypes = [I, I]
            getattr(lib, name).restype = C.c_int
        lib.xmi_find_solid_angle_match.argtypes = [C.c_char_p, I, P(P(api.xmi_solid_angle)),
                                                   P(api.xmi_main_options)]
        lib.xmi_find_solid_angle_match.restype = C.c_int
        lib.xmi_find_escape_ratios_match.argtypes = [C.c_char_p, I, P(P(api.xmi_escape_ratios)),
                                                     P(api.xmi_main_options)]
        lib.xmi_find_escape_ratios_match.restype = C.c_int
        for name in ("xmi_get_solid_angle_file", "xmi_get_escape_ratios_file"):
            getattr(lib, name).argtypes = [P(C.c_void_p), C.c_int]
            getattr(lib, name).restype = C.c_int
        _declared = True
    return lib


def _fail(what):
    raise RuntimeError(f"{what}: {api.last_error()}")


def xmsi_text(inp: api.Input) -> bytes:
    """The XMSI text stored with a cached result (an output file name is
    required by the input's validation; a placeholder is used when unset)."""
    if not inp.outputfile:
        inp = dataclasses.replace(inp, outputfile="xmimsim.xmso")
    return xmlio.xmsi_string(inp).encode()


def default_file(kind: str, create: bool = True) -> str:
    """xmi_get_solid_angle_file / xmi_get_escape_ratios_file with no name:
    the user data directory's XMI-MSIM file of that kind."""
    lib = _lib()
    p = C.c_void_p()
    fn = lib.xmi_get_solid_angle_file if kind == "solid_angle" else lib.xmi_get_escape_ratios_file
    if not fn(C.byref(p), int(create)):
        _fail("default cache file")
    try:
        return C.string_at(p.value).decode()
    finally:
        lib.xmi_free(p)


def create(path: str, kind: str) -> None:
    lib = _lib()
    fn = (lib.xmi_create_empty_solid_angle_hdf5_file if kind == "solid_angle"
          else lib.xmi_create_empty_escape_ratios_hdf5_file)
    if not fn(path.encode()):
        _fail(path)


def solid_angle_match(a: api.Input, b: api.Input) -> bool:
    ca, cb = api.CInput(a), api.CInput(b)
    return bool(_lib().xmi_check_solid_angle_match(C.byref(ca.input), C.byref(cb.input)))


def escape_ratios_match(a: api.Input, b: api.Input) -> bool:
    ca, cb = api.CInput(a), api.CInput(b)
    return bool(_lib().xmi_check_escape_ratios_match(C.byref(ca.input), C.byref(cb.input)))


def store_solid_angle(path: str, inp: api.Input, grid, r_vals, theta_vals) -> None:
    """Adds a grid ([theta_n][r_n] with its axes) computed for inp to the cache."""
    import numpy as np
    lib = _lib()
    if not os.path.exists(path):
        create(path, "solid_angle")
    g = np.ascontiguousarray(grid, dtype=np.float64)
    r = np.ascontiguousarray(r_vals, dtype=np.float64)
    t = np.ascontiguousarray(theta_vals, dtype=np.float64)
    text = C.create_string_buffer(xmsi_text(inp))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    sa = api.xmi_solid_angle(dp(g), r.size, t.size, dp(r), dp(t), C.cast(text, C.c_char_p))
    if not lib.xmi_update_solid_angle_hdf5_file(path.encode(), C.byref(sa)):
        _fail(path)


def find_solid_angle(path: str, inp: api.Input, options: Optional[api.xmi_main_options] = None):
    """The newest cached grid that serves inp, as (grid, r_vals, theta_vals,
    pointer), or None.  The pointer is the library's xmi_solid_angle (hand it
    to Prepared.set_solid_angle(ptr, owner="engine"), or release it with
    xmi_free_solid_angle)."""
    import numpy as np
    lib = _lib()
    ci = api.CInput(inp)
    rv = C.POINTER(api.xmi_solid_angle)()
    opts = options or api.default_options()
    if not lib.xmi_find_solid_angle_match(path.encode(), C.byref(ci.input), C.byref(rv), C.byref(opts)):
        _fail(path)
    if not rv:
        return None
    s = rv.contents
    nt, nr = s.grid_dims_theta_n, s.grid_dims_r_n
    grid = np.ctypeslib.as_array(s.solid_angles, shape=(nt, nr)).copy()
    r = np.ctypeslib.as_array(s.grid_dims_r_vals, shape=(nr,)).copy()
    t = np.ctypeslib.as_array(s.grid_dims_theta_vals, shape=(nt,)).copy()
    return grid, r, t, rv


def store_escape_ratios(path: str, inp: api.Input, handle) -> None:
    """Adds the escape ratios of api.escape_ratios (its handle) computed for
    inp's detector crystal to the cache."""
    lib = _lib()
    if not os.path.exists(path):
        create(path, "escape_ratios")
    # the record's own string pointer is swapped for the XMSI text during
    # the call and put back afterwards (the library frees it with the record)
    slot = C.c_void_p.from_address(C.addressof(handle.contents) + api.xmi_escape_ratios.xmi_input_string.offset)
    keep = slot.value
    text = C.create_string_buffer(xmsi_text(inp))
    slot.value = C.cast(text, C.c_void_p).value
    try:
        if not lib.xmi_update_escape_ratios_hdf5_file(path.encode(), handle):
            _fail(path)
    finally:
        slot.value = keep


def find_escape_ratios(path: str, inp: api.Input, options: Optional[api.xmi_main_options] = None):
    """The newest cached escape ratios for inp's crystal (the library's
    xmi_escape_ratios pointer, for api.detector_response; release with
    api.free_escape_ratios), or None."""
    lib = _lib()
    ci = api.CInput(inp)
    rv = C.POINTER(api.xmi_escape_ratios)()
    opts = options or api.default_options()
    if not lib.xmi_find_escape_ratios_match(path.encode(), C.byref(ci.input), C.byref(rv), C.byref(opts)):
        _fail(path)
    return rv if rv else None
