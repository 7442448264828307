"""Line names, provider values and loaded tables as numpy arrays.

Line numbering is xraylib's (KL1 = 1 ... P3P5 = 383), the row index of the
reference's history arrays (src/xmi_lines.c); 384 is Rayleigh, 385 Compton.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List

import numpy as np

from . import api

NSHELLS = 31
NLINES = 383
NCS = 1024
NCK = 14
# first/last line of each vacancy shell K..M5 (xraylib numbering)
SHELL_LINES = [(1, 29), (30, 58), (59, 85), (86, 113), (114, 136), (137, 158), (159, 180), (181, 200), (201, 219)]


class XrlAtom(C.Structure):
    """Mirror of xrl_atom (xmimsim_amd/csrc/xrl_model.h)."""
    _fields_ = [("Z", C.c_int), ("n_occupied", C.c_int), ("A", C.c_double),
                ("edge", C.c_double * NSHELLS), ("occupancy", C.c_double * NSHELLS),
                ("fluor_yield", C.c_double * 9), ("ck", C.c_double * NCK),
                ("rad_rate", C.c_double * (NLINES + 1)), ("line_energy", C.c_double * (NLINES + 1)),
                ("photo_B", C.c_double * 9), ("photo_B_outer", C.c_double), ("auger_KL", C.c_double * 3),
                ("auger_KM", C.c_double * 5), ("auger_LM", C.c_double * 15), ("ff_x0", C.c_double),
                ("sf_x0", C.c_double), ("profile_p", C.c_double * NSHELLS), ("photo_Bsum", C.c_double * 10),
                ("shell_C", C.c_double * 360), ("cs_rayl", C.c_double * NCS), ("cs_compt", C.c_double * NCS),
                ("cs_rc", C.c_double * NCS), ("inv_ff_x0", C.c_double), ("inv_sf_x0", C.c_double),
                ("barn_per_A", C.c_double)]


class HDF5Z(C.Structure):
    """Mirror of xmi_hdf5_ZF (xmimsim_amd/csrc/xmi_internal.h)."""
    _fields_ = [("Z", C.c_int), ("Zindex", C.c_int), ("n_e", C.c_int), ("n_r", C.c_int),
                ("Energies", C.POINTER(C.c_double)), ("RandomNumbers", C.POINTER(C.c_double)),
                ("RayleighTheta_ICDF", C.POINTER(C.c_double)), ("ComptonTheta_ICDF", C.POINTER(C.c_double)),
                ("FluorYieldsCorr", C.c_double * 9), ("ip_len", C.c_int),
                ("ip_energies", C.POINTER(C.c_double)), ("ip_rayl", C.POINTER(C.c_double)),
                ("ip_rayl_compt", C.POINTER(C.c_double)), ("cp_nshells", C.c_int), ("cp_len", C.c_int),
                ("cp_shell_indices", C.POINTER(C.c_int)), ("cp_Qs", C.POINTER(C.c_double)),
                ("cp_partial_cdf", C.POINTER(C.c_double)), ("cp_partial_cdf_inv", C.POINTER(C.c_double)),
                ("cp_Qs_inv", C.POINTER(C.c_double)), ("cp_random_numbers", C.POINTER(C.c_double)),
                ("cp_total_icdf", C.POINTER(C.c_double)), ("precalc_xrf_cs", C.POINTER(C.c_double))]


class HDF5(C.Structure):
    """Mirror of xmi_hdf5F."""
    _fields_ = [("n_Z", C.c_int), ("Zs", C.POINTER(HDF5Z)), ("phi_n_theta", C.c_int), ("phi_n_r", C.c_int),
                ("Phi_ICDF", C.POINTER(C.c_double)), ("Thetas", C.POINTER(C.c_double)),
                ("RandomNumbers", C.POINTER(C.c_double)), ("uniqZ", C.c_int * 95), ("cascade_type", C.c_int),
                ("use_M_lines", C.c_int)]


def _lib():
    lib = api.load_library()
    if not getattr(lib, "_lines_ready", False):
        lib.xmi_line_name.restype = C.c_char_p
        lib.xmi_line_name.argtypes = [C.c_int]
        lib.xmi_provider_atom.argtypes = [C.c_int, C.c_void_p]
        lib.xmi_provider_cs.argtypes = [C.c_int, C.c_int, C.c_double]
        lib.xmi_provider_cs.restype = C.c_double
        lib.xmi_provider_dcs.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        lib.xmi_provider_dcs.restype = C.c_double
        lib.xmi_provider_shell_cross_sections.argtypes = [C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double)]
        lib.xmi_provider_auger.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int)]
        lib.xmi_provider_auger.restype = C.c_double
        lib._lines_ready = True
    return lib


def line_names() -> List[str]:
    lib = _lib()
    return [lib.xmi_line_name(i).decode() for i in range(386)]


def atom(Z: int) -> Dict[str, np.ndarray]:
    a = XrlAtom()
    if not _lib().xmi_provider_atom(Z, C.byref(a)):
        raise ValueError(Z)
    out = {}
    for name, _ in XrlAtom._fields_:
        v = getattr(a, name)
        out[name] = np.array(v) if hasattr(v, "__len__") else v
    return out


def cs_total(Z: int, E: float) -> float:
    return _lib().xmi_provider_cs(0, Z, E)


def cs_photo_total(Z: int, E: float) -> float:
    return _lib().xmi_provider_cs(1, Z, E)


def cs(kind: str, Z: int, E: float) -> float:
    return _lib().xmi_provider_cs({"rayl": 2, "compt": 3}[kind], Z, E)


def dcs(kind: str, Z: int, E: float, theta: float) -> float:
    return _lib().xmi_provider_dcs({"rayl": 0, "compt": 1}[kind], Z, E, theta)


def shell_cross_sections(Z: int, E: float, cascade: int = 4) -> np.ndarray:
    out = (C.c_double * 9)()
    _lib().xmi_provider_shell_cross_sections(Z, cascade, E, out)
    return np.array(out)


def auger_transitions(Z: int):
    """The provider's Auger transitions of element Z as (rate, vacancy
    shell, first new vacancy, second new vacancy) rows (shell indices 0 K ..
    8 M5, 9 beyond M5)."""
    out = []
    sh = (C.c_int * 3)()
    t = 0
    while True:
        r = _lib().xmi_provider_auger(Z, t, sh)
        if r < 0:
            break
        out.append((r, sh[0], sh[1], sh[2]))
        t += 1
    return out


def line_energy(Z: int, line: int) -> float:
    return float(atom(Z)["line_energy"][line])


def tables_of(prep) -> Dict[str, np.ndarray]:
    """Arrays of the tables behind a Prepared's hdf5 handle (all elements concatenated)."""
    h = C.cast(prep.hdf5F, C.POINTER(HDF5)).contents
    out: Dict[str, list] = {}
    for i in range(h.n_Z):
        z = h.Zs[i]
        ne, nr = z.n_e, z.n_r
        arrs = {
            "Energies": np.ctypeslib.as_array(z.Energies, (ne,)),
            "RandomNumbers": np.ctypeslib.as_array(z.RandomNumbers, (nr,)),
            "RayleighTheta_ICDF": np.ctypeslib.as_array(z.RayleighTheta_ICDF, (ne, nr)),
            "ComptonTheta_ICDF": np.ctypeslib.as_array(z.ComptonTheta_ICDF, (ne, nr)),
            "FluorYieldsCorr": np.array(z.FluorYieldsCorr),
            "ip_energies": np.ctypeslib.as_array(z.ip_energies, (z.ip_len,)),
            "ip_rayl": np.ctypeslib.as_array(z.ip_rayl, (z.ip_len,)),
            "ip_rayl_compt": np.ctypeslib.as_array(z.ip_rayl_compt, (z.ip_len,)),
            "cp_total_icdf": np.ctypeslib.as_array(z.cp_total_icdf, (z.cp_len,)),
            "cp_Qs_inv": np.ctypeslib.as_array(z.cp_Qs_inv, (z.cp_nshells, z.cp_len)),
            "cp_partial_cdf": np.ctypeslib.as_array(z.cp_partial_cdf, (z.cp_nshells, z.cp_len)),
            "precalc_xrf_cs": np.ctypeslib.as_array(z.precalc_xrf_cs, (9, h.n_Z, 219)),
        }
        for k, v in arrs.items():
            out.setdefault(k, []).append(np.array(v))
    res = {k: np.concatenate([x.ravel() for x in v]) for k, v in out.items()}
    res["Phi_ICDF"] = np.ctypeslib.as_array(h.Phi_ICDF, (h.phi_n_theta, h.phi_n_r)).copy()
    res["Thetas"] = np.ctypeslib.as_array(h.Thetas, (h.phi_n_theta,)).copy()
    return res


def tables_for(Zs: List[int]) -> Dict[str, np.ndarray]:
    """Tables of one element (first of Zs) with their native shapes."""
    inp = api.Input(layers=[api.Layer(list(Zs), [1.0 / len(Zs)] * len(Zs), 1.0, 0.1)],
                    discrete=[api.DiscreteLine(20.0, 1e9)], n_photons_line=10)
    prep = api.Prepared(inp)
    h = C.cast(prep.hdf5F, C.POINTER(HDF5)).contents
    z = h.Zs[0]
    ne, nr = z.n_e, z.n_r
    out = {
        "Energies": np.ctypeslib.as_array(z.Energies, (ne,)).copy(),
        "RandomNumbers": np.ctypeslib.as_array(z.RandomNumbers, (nr,)).copy(),
        "RayleighTheta_ICDF": np.ctypeslib.as_array(z.RayleighTheta_ICDF, (ne, nr)).copy(),
        "ComptonTheta_ICDF": np.ctypeslib.as_array(z.ComptonTheta_ICDF, (ne, nr)).copy(),
        "ip_energies": np.ctypeslib.as_array(z.ip_energies, (z.ip_len,)).copy(),
        "ip_rayl": np.ctypeslib.as_array(z.ip_rayl, (z.ip_len,)).copy(),
        "cp_total_icdf": np.ctypeslib.as_array(z.cp_total_icdf, (z.cp_len,)).copy(),
        "Phi_ICDF": np.ctypeslib.as_array(h.Phi_ICDF, (h.phi_n_theta, h.phi_n_r)).copy(),
        "Thetas": np.ctypeslib.as_array(h.Thetas, (h.phi_n_theta,)).copy(),
    }
    prep.close()
    return out
