"""ctypes mirror of the engine's C ABI (include/xmimsim_amd.h).

The classes below have the member order and types of the reference's public C
structs (include/xmi_data_structs.h:36-495 and include/xmi_solid_angle.h:27-35
of XMI-MSIM 8.1), so a caller of the reference library can hand the same
structures over.  The Python helpers build inputs, run the GPU engine and
return numpy arrays with the reference's layouts:

    channels          [n_interactions + 1][nchannels]   cumulative over order
    var_red_history   [100][385][n_interactions]        Z-1, line-1 (384 = Rayleigh, 385 = Compton)

The engine is native code; when libxmimsim_amd.so is missing this module
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxmimsim_amd.so")


class xmi_general(C.Structure):
    _fields_ = [("version", C.c_float), ("outputfile", C.c_char_p), ("n_photons_interval", C.c_long),
                ("n_photons_line", C.c_long), ("n_interactions_trajectory", C.c_int), ("comments", C.c_char_p)]


class xmi_layer(C.Structure):
    _fields_ = [("n_elements", C.c_int), ("Z", C.POINTER(C.c_int)), ("weight", C.POINTER(C.c_double)),
                ("density", C.c_double), ("thickness", C.c_double)]


class xmi_composition(C.Structure):
    _fields_ = [("n_layers", C.c_int), ("layers", C.POINTER(xmi_layer)), ("reference_layer", C.c_int)]


class xmi_geometry(C.Structure):
    _fields_ = [("d_sample_source", C.c_double), ("n_sample_orientation", C.c_double * 3),
                ("p_detector_window", C.c_double * 3), ("n_detector_orientation", C.c_double * 3),
                ("area_detector", C.c_double), ("collimator_height", C.c_double),
                ("collimator_diameter", C.c_double), ("d_source_slit", C.c_double), ("slit_size_x", C.c_double),
                ("slit_size_y", C.c_double)]


class xmi_energy_discrete(C.Structure):
    _fields_ = [("energy", C.c_double), ("horizontal_intensity", C.c_double), ("vertical_intensity", C.c_double),
                ("sigma_x", C.c_double), ("sigma_xp", C.c_double), ("sigma_y", C.c_double),
                ("sigma_yp", C.c_double), ("distribution_type", C.c_int), ("scale_parameter", C.c_double)]


class xmi_energy_continuous(C.Structure):
    _fields_ = [("energy", C.c_double), ("horizontal_intensity", C.c_double), ("vertical_intensity", C.c_double),
                ("sigma_x", C.c_double), ("sigma_xp", C.c_double), ("sigma_y", C.c_double),
                ("sigma_yp", C.c_double)]


class xmi_excitation(C.Structure):
    _fields_ = [("n_discrete", C.c_int), ("discrete", C.POINTER(xmi_energy_discrete)), ("n_continuous", C.c_int),
                ("continuous", C.POINTER(xmi_energy_continuous))]


class xmi_absorbers(C.Structure):
    _fields_ = [("n_exc_layers", C.c_int), ("exc_layers", C.POINTER(xmi_layer)), ("n_det_layers", C.c_int),
                ("det_layers", C.POINTER(xmi_layer))]


class xmi_detector(C.Structure):
    _fields_ = [("detector_type", C.c_int), ("live_time", C.c_double), ("pulse_width", C.c_double),
                ("gain", C.c_double), ("zero", C.c_double), ("fano", C.c_double), ("noise", C.c_double),
                ("nchannels", C.c_int), ("n_crystal_layers", C.c_int), ("crystal_layers", C.POINTER(xmi_layer))]


class xmi_input(C.Structure):
    _fields_ = [("general", C.POINTER(xmi_general)), ("composition", C.POINTER(xmi_composition)),
                ("geometry", C.POINTER(xmi_geometry)), ("excitation", C.POINTER(xmi_excitation)),
                ("absorbers", C.POINTER(xmi_absorbers)), ("detector", C.POINTER(xmi_detector))]


class xmi_main_options(C.Structure):
    _fields_ = [("use_M_lines", C.c_int), ("use_cascade_auger", C.c_int), ("use_cascade_radiative", C.c_int),
                ("use_variance_reduction", C.c_int), ("use_sum_peaks", C.c_int), ("use_escape_peaks", C.c_int),
                ("escape_ratios_mode", C.c_int), ("verbose", C.c_int), ("use_poisson", C.c_int),
                ("use_gpu", C.c_int), ("omp_num_threads", C.c_int), ("extra_verbose", C.c_int),
                ("custom_detector_response", C.c_char_p), ("use_advanced_compton", C.c_int),
                ("use_default_seeds", C.c_int)]


class xmi_solid_angle(C.Structure):
    _fields_ = [("solid_angles", C.POINTER(C.c_double)), ("grid_dims_r_n", C.c_long),
                ("grid_dims_theta_n", C.c_long), ("grid_dims_r_vals", C.POINTER(C.c_double)),
                ("grid_dims_theta_vals", C.POINTER(C.c_double)), ("xmi_input_string", C.c_char_p)]


class xmi_msim_stats(C.Structure):
    _fields_ = [("transport_ms", C.c_double), ("total_ms", C.c_double), ("histories", C.c_int64),
                ("launches", C.c_int64), ("vr_kernel_ms", C.c_double), ("vr_events", C.c_int64),
                ("step_kernel_ms", C.c_double), ("interaction_kernel_ms", C.c_double),
                ("solid_angle_fallbacks", C.c_int64), ("vr_geom_ms", C.c_double), ("vr_cone_ms", C.c_double),
                ("vr_scatter_ms", C.c_double), ("vr_lines_ms", C.c_double),
                ("grid_uploads", C.c_int64), ("photon_steps", C.c_int64), ("detector_hits", C.c_int64),
                ("batches", C.c_int64), ("graph_replays", C.c_int64)]


class xmi_msim_launch(C.Structure):
    _fields_ = [("batch_log2", C.c_int), ("use_graph", C.c_int), ("lines_kp", C.c_int), ("wg_per_cu", C.c_int)]


class xmi_error(C.Structure):
    _fields_ = [("domain", C.c_uint32), ("code", C.c_int), ("message", C.c_char_p)]


class xmi_counts(C.Structure):
    _fields_ = [("counts", C.c_double), ("interaction_number", C.c_int)]


class xmi_fluorescence_line(C.Structure):
    _fields_ = [("line_type", C.c_char_p), ("energy", C.c_double), ("total_counts", C.c_double),
                ("n_interactions", C.c_int), ("interactions", C.POINTER(xmi_counts))]


class xmi_fluorescence_line_counts(C.Structure):
    _fields_ = [("atomic_number", C.c_int), ("total_counts", C.c_double), ("n_lines", C.c_int),
                ("lines", C.POINTER(xmi_fluorescence_line))]


class xmi_output(C.Structure):
    _fields_ = [("version", C.c_float), ("inputfile", C.c_char_p), ("outputfile", C.c_char_p),
                ("input", C.POINTER(xmi_input)),
                ("brute_force_history", C.POINTER(xmi_fluorescence_line_counts)),
                ("var_red_history", C.POINTER(xmi_fluorescence_line_counts)),
                ("nbrute_force_history", C.c_int), ("nvar_red_history", C.c_int),
                ("channels_conv", C.POINTER(C.POINTER(C.c_double))),
                ("channels_unconv", C.POINTER(C.POINTER(C.c_double))), ("ninteractions", C.c_int),
                ("use_zero_interactions", C.c_int)]


# exported symbols of include/xmimsim_amd.h
EXPORTED_SYMBOLS = [
    "xmi_main_options_new", "xmi_input_C2F", "xmi_init_input", "xmi_free_input_F", "xmi_db",
    "xmi_init_from_hdf5", "xmi_init_from_provider", "xmi_update_input_from_hdf5", "xmi_free_hdf5_F",
    "xmi_solid_angle_calculation_cl", "xmi_free_solid_angle", "xmi_hits_per_single", "xmi_main_msim",
    "xmi_free", "xmi_set_seed", "xmi_msim_get_stats", "xmi_msim_last_error", "xmi_msim_set_profiling",
    "xmi_msim_get_layer_events", "xmi_msim_set_launch", "xmi_msim_get_launch", "xmi_provider_active_lines", "xmi_detector_convolute_all",
    "xmi_detector_convolute_all_custom", "xmi_surrogate_create", "xmi_surrogate_train_step", "xmi_main_msim_batch",
    "xmi_tables_set_device", "xmi_tables_copy", "xmi_escape_ratios_calculation", "xmi_free_escape_ratios",
    "xmi_get_default_escape_ratios_options", "xmi_input_read_from_xml_file", "xmi_input_read_from_xml_string",
    "xmi_input_write_to_xml_file", "xmi_input_write_to_xml_string", "xmi_input_validate", "xmi_input_free",
    "xmi_output_raw2struct", "xmi_output_write_to_xml_file", "xmi_output_read_from_xml_file", "xmi_output_free",
    "xmi_error_free", "xmi_element_symbol", "xmi_create_empty_solid_angle_hdf5_file",
    "xmi_update_solid_angle_hdf5_file", "xmi_check_solid_angle_match", "xmi_find_solid_angle_match",
    "xmi_get_solid_angle_file", "xmi_create_empty_escape_ratios_hdf5_file", "xmi_update_escape_ratios_hdf5_file",
    "xmi_check_escape_ratios_match", "xmi_find_escape_ratios_match", "xmi_get_escape_ratios_file",
]

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libxmimsim_amd.so and declares the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(path)
    P = C.POINTER
    lib.xmi_main_options_new.restype = P(xmi_main_options)
    lib.xmi_input_C2F.argtypes = [P(xmi_input), P(C.c_void_p)]
    lib.xmi_input_C2F.restype = C.c_int
    lib.xmi_init_input.argtypes = [P(C.c_void_p)]
    lib.xmi_init_input.restype = C.c_int
    lib.xmi_free_input_F.argtypes = [P(C.c_void_p)]
    lib.xmi_db.argtypes = [C.c_char_p, P(C.c_int), C.c_int]
    lib.xmi_db.restype = C.c_int
    lib.xmi_init_from_hdf5.argtypes = [C.c_char_p, C.c_void_p, P(C.c_void_p), P(xmi_main_options)]
    lib.xmi_init_from_hdf5.restype = C.c_int
    lib.xmi_init_from_provider.argtypes = [C.c_void_p, P(C.c_void_p), P(xmi_main_options)]
    lib.xmi_init_from_provider.restype = C.c_int
    lib.xmi_tables_set_device.argtypes = [C.c_int]
    lib.xmi_tables_set_device.restype = None
    lib.xmi_tables_copy.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double), C.c_long]
    lib.xmi_tables_copy.restype = C.c_long
    lib.xmi_escape_ratios_calculation.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, C.c_char_p, C.c_void_p,
                                                  xmi_escape_ratios_options]
    lib.xmi_escape_ratios_calculation.restype = None
    lib.xmi_free_escape_ratios.argtypes = [C.c_void_p]
    lib.xmi_get_default_escape_ratios_options.restype = xmi_escape_ratios_options
    lib.xmi_update_input_from_hdf5.argtypes = [C.c_void_p, C.c_void_p]
    lib.xmi_update_input_from_hdf5.restype = C.c_int
    lib.xmi_free_hdf5_F.argtypes = [P(C.c_void_p)]
    lib.xmi_solid_angle_calculation_cl.argtypes = [C.c_void_p, P(P(xmi_solid_angle)), C.c_char_p,
                                                   P(xmi_main_options)]
    lib.xmi_solid_angle_calculation_cl.restype = C.c_int
    lib.xmi_free_solid_angle.argtypes = [P(xmi_solid_angle)]
    lib.xmi_main_msim.argtypes = [C.c_void_p, C.c_void_p, C.c_int, P(P(C.c_double)), P(xmi_main_options),
                                  P(P(C.c_double)), P(P(C.c_double)), P(xmi_solid_angle)]
    lib.xmi_main_msim.restype = C.c_int
    lib.xmi_free.argtypes = [C.c_void_p]
    lib.xmi_set_seed.argtypes = [C.c_uint64]
    lib.xmi_msim_get_stats.argtypes = [P(xmi_msim_stats)]
    lib.xmi_msim_last_error.restype = C.c_char_p
    lib.xmi_msim_set_profiling.argtypes = [C.c_int]
    lib.xmi_msim_get_layer_events.argtypes = [P(C.c_int64), C.c_int]
    lib.xmi_msim_get_layer_events.restype = C.c_int
    lib.xmi_msim_set_launch.argtypes = [P(xmi_msim_launch)]
    lib.xmi_msim_set_launch.restype = C.c_int
    lib.xmi_msim_get_launch.argtypes = [P(xmi_msim_launch)]
    lib.xmi_msim_get_launch.restype = None
    lib.xmi_provider_active_lines.argtypes = [C.c_int, C.c_int]
    lib.xmi_provider_active_lines.restype = C.c_int
    lib.xmi_provider_dcs.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
    lib.xmi_provider_dcs.restype = C.c_double
    lib.xmi_provider_cs.argtypes = [C.c_int, C.c_int, C.c_double]
    lib.xmi_provider_cs.restype = C.c_double
    lib.xmi_provider_line_data.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.xmi_provider_line_data.restype = C.c_double
    lib.xmi_provider_shell_cross_sections.argtypes = [C.c_int, C.c_int, C.c_double, P(C.c_double)]
    lib.xmi_provider_shell_cross_sections.restype = C.c_int
    PE = P(P(xmi_error))
    lib.xmi_input_read_from_xml_file.argtypes = [C.c_char_p, PE]
    lib.xmi_input_read_from_xml_file.restype = P(xmi_input)
    lib.xmi_input_read_from_xml_string.argtypes = [C.c_char_p, PE]
    lib.xmi_input_read_from_xml_string.restype = P(xmi_input)
    lib.xmi_input_write_to_xml_file.argtypes = [P(xmi_input), C.c_char_p, PE]
    lib.xmi_input_write_to_xml_file.restype = C.c_int
    lib.xmi_input_write_to_xml_string.argtypes = [P(xmi_input), P(C.c_void_p), PE]
    lib.xmi_input_write_to_xml_string.restype = C.c_int
    lib.xmi_input_validate.argtypes = [P(xmi_input)]
    lib.xmi_input_validate.restype = C.c_int
    lib.xmi_input_free.argtypes = [P(xmi_input)]
    lib.xmi_input_free.restype = None
    lib.xmi_output_raw2struct.argtypes = [P(xmi_input), P(C.c_double), P(C.c_double), P(P(C.c_double)),
                                          P(C.c_double), C.c_char_p, C.c_int]
    lib.xmi_output_raw2struct.restype = P(xmi_output)
    lib.xmi_output_write_to_xml_file.argtypes = [P(xmi_output), C.c_char_p, PE]
    lib.xmi_output_write_to_xml_file.restype = C.c_int
    lib.xmi_output_read_from_xml_file.argtypes = [C.c_char_p, PE]
    lib.xmi_output_read_from_xml_file.restype = P(xmi_output)
    lib.xmi_output_free.argtypes = [P(xmi_output)]
    lib.xmi_output_free.restype = None
    lib.xmi_error_free.argtypes = [P(xmi_error)]
    lib.xmi_error_free.restype = None
    lib.xmi_line_name.argtypes = [C.c_int]
    lib.xmi_line_name.restype = C.c_char_p
    lib.xmi_element_symbol.argtypes = [C.c_int]
    lib.xmi_element_symbol.restype = C.c_char_p
    _lib = lib
    return lib


def get_stats() -> "xmi_msim_stats":
    """Timing and counters of the last engine call on this thread."""
    lib = load_library()
    st = xmi_msim_stats()
    lib.xmi_msim_get_stats(C.byref(st))
    return st


def set_launch(batch_log2: int = 0, use_graph: int = 0, lines_kp: int = 0, wg_per_cu: int = 0):
    """Launch shape of the following engine calls (xmi_msim_set_launch);
    zeros take the defaults.  Raises ValueError for a shape out of range."""
    lib = load_library()
    L = xmi_msim_launch(int(batch_log2), int(use_graph), int(lines_kp), int(wg_per_cu))
    if not lib.xmi_msim_set_launch(C.byref(L)):
        raise ValueError(last_error())


def get_launch() -> dict:
    """The launch shape in force, as keyword arguments of set_launch."""
    lib = load_library()
    L = xmi_msim_launch()
    lib.xmi_msim_get_launch(C.byref(L))
    return {f: getattr(L, f) for f, _ in xmi_msim_launch._fields_}


def last_error() -> str:
    return load_library().xmi_msim_last_error().decode()


# ---------------------------------------------------------------------------
# input description


@dataclass
class Layer:
    Z: Sequence[int]
    weight: Sequence[float]
    density: float
    thickness: float


@dataclass
class DiscreteLine:
    energy: float
    horizontal_intensity: float
    vertical_intensity: float = 0.0
    sigma_x: float = 0.0
    sigma_xp: float = 0.0
    sigma_y: float = 0.0
    sigma_yp: float = 0.0
    distribution_type: int = 0
    scale_parameter: float = 0.0


@dataclass
class ContinuousPoint:
    energy: float
    horizontal_intensity: float
    vertical_intensity: float = 0.0
    sigma_x: float = 0.0
    sigma_xp: float = 0.0
    sigma_y: float = 0.0
    sigma_yp: float = 0.0


@dataclass
class Input:
    """Python view of xmi_input; defaults are xmi_input_init_empty
    (src/xmi_data_structs.c:551-654)."""
    layers: List[Layer]
    reference_layer: int = 1
    outputfile: str = ""
    comments: str = ""
    discrete: List[DiscreteLine] = field(default_factory=list)
    continuous: List[ContinuousPoint] = field(default_factory=list)
    n_photons_interval: int = 10000
    n_photons_line: int = 100000
    n_interactions_trajectory: int = 4
    d_sample_source: float = 100.0
    n_sample_orientation: Sequence[float] = (0.0, -0.7071067811865476, 0.7071067811865476)
    p_detector_window: Sequence[float] = (0.0, 1.0, 100.0)
    n_detector_orientation: Sequence[float] = (0.0, -1.0, 0.0)
    area_detector: float = 0.3
    collimator_height: float = 0.0
    collimator_diameter: float = 0.0
    d_source_slit: float = 100.0
    slit_size_x: float = 0.001
    slit_size_y: float = 0.001
    exc_absorbers: List[Layer] = field(default_factory=list)
    det_absorbers: List[Layer] = field(default_factory=lambda: [Layer([4], [1.0], 1.85, 0.002)])
    detector_type: int = 0
    live_time: float = 1.0
    pulse_width: float = 10e-6
    gain: float = 0.02
    zero: float = 0.0
    fano: float = 0.12
    noise: float = 0.1
    nchannels: int = 2048
    crystal: List[Layer] = field(default_factory=lambda: [Layer([14], [1.0], 2.33, 0.5)])

    def unique_Z(self) -> List[int]:
        return sorted({z for l in self.layers for z in l.Z})


class CInput:
    """Owns the ctypes objects of one xmi_input (keeps every buffer alive)."""

    def __init__(self, inp: Input):
        self._keep = []

        def layers(ls: List[Layer]):
            if not ls:
                return 0, None
            arr = (xmi_layer * len(ls))()
            for i, l in enumerate(ls):
                z = (C.c_int * len(l.Z))(*l.Z)
                w = (C.c_double * len(l.weight))(*l.weight)
                self._keep += [z, w]
                arr[i] = xmi_layer(len(l.Z), z, w, l.density, l.thickness)
            self._keep.append(arr)
            return len(ls), arr

        self.general = xmi_general(8.1, inp.outputfile.encode(), inp.n_photons_interval, inp.n_photons_line,
                                   inp.n_interactions_trajectory, inp.comments.encode())
        nl, la = layers(inp.layers)
        self.composition = xmi_composition(nl, la, inp.reference_layer)
        g = xmi_geometry()
        g.d_sample_source = inp.d_sample_source
        g.n_sample_orientation[:] = list(inp.n_sample_orientation)
        g.p_detector_window[:] = list(inp.p_detector_window)
        g.n_detector_orientation[:] = list(inp.n_detector_orientation)
        g.area_detector = inp.area_detector
        g.collimator_height = inp.collimator_height
        g.collimator_diameter = inp.collimator_diameter
        g.d_source_slit = inp.d_source_slit
        g.slit_size_x = inp.slit_size_x
        g.slit_size_y = inp.slit_size_y
        self.geometry = g
        disc = (xmi_energy_discrete * max(1, len(inp.discrete)))()
        for i, d in enumerate(inp.discrete):
            disc[i] = xmi_energy_discrete(d.energy, d.horizontal_intensity, d.vertical_intensity, d.sigma_x,
                                          d.sigma_xp, d.sigma_y, d.sigma_yp, d.distribution_type,
                                          d.scale_parameter)
        cont = (xmi_energy_continuous * max(1, len(inp.continuous)))()
        for i, c in enumerate(inp.continuous):
            cont[i] = xmi_energy_continuous(c.energy, c.horizontal_intensity, c.vertical_intensity, c.sigma_x,
                                            c.sigma_xp, c.sigma_y, c.sigma_yp)
        self._keep += [disc, cont]
        self.excitation = xmi_excitation(len(inp.discrete), disc, len(inp.continuous), cont)
        ne, ea = layers(inp.exc_absorbers)
        nd, da = layers(inp.det_absorbers)
        self.absorbers = xmi_absorbers(ne, ea, nd, da)
        ncr, cr = layers(inp.crystal)
        self.detector = xmi_detector(inp.detector_type, inp.live_time, inp.pulse_width, inp.gain, inp.zero, inp.fano,
                                     inp.noise, inp.nchannels, ncr, cr)
        self.input = xmi_input(C.pointer(self.general), C.pointer(self.composition), C.pointer(self.geometry),
                               C.pointer(self.excitation), C.pointer(self.absorbers), C.pointer(self.detector))


def default_options(**kw) -> xmi_main_options:
    lib = load_library()
    p = lib.xmi_main_options_new()
    o = xmi_main_options()
    C.memmove(C.byref(o), p, C.sizeof(o))
    lib.xmi_free(p)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Prepared:
    """An input after xmi_input_C2F + xmi_init_input + table loading (the
    handles xmi_main_msim takes), plus an optional solid-angle grid."""

    def __init__(self, inp: Input, options: Optional[xmi_main_options] = None, hdf5_file: Optional[str] = None,
                 tables: Optional["SharedTables"] = None):
        lib = load_library()
        self.lib = lib
        self.inp = inp
        self.options = options or default_options()
        self.cinput = CInput(inp)
        self.inputF = C.c_void_p()
        self.hdf5F = C.c_void_p()
        self._owns_tables = tables is None
        self.solid_angle = None
        if not lib.xmi_input_C2F(C.byref(self.cinput.input), C.byref(self.inputF)):
            raise RuntimeError("xmi_input_C2F failed: " + last_error())
        if not lib.xmi_init_input(C.byref(self.inputF)):
            raise RuntimeError("xmi_init_input failed: " + last_error())
        if tables is not None:
            self.hdf5F = tables.hdf5F
            if not lib.xmi_update_input_from_hdf5(self.inputF, self.hdf5F):
                raise RuntimeError("xmi_update_input_from_hdf5 failed: " + last_error())
            return
        if hdf5_file is not None:
            ok = lib.xmi_init_from_hdf5(hdf5_file.encode(), self.inputF, C.byref(self.hdf5F), C.byref(self.options))
        else:
            ok = lib.xmi_init_from_provider(self.inputF, C.byref(self.hdf5F), C.byref(self.options))
        if not ok:
            raise RuntimeError("table loading failed: " + last_error())
        if not lib.xmi_update_input_from_hdf5(self.inputF, self.hdf5F):
            raise RuntimeError("xmi_update_input_from_hdf5 failed: " + last_error())

    def compute_solid_angle(self, hits_per_single: int = 5000, seed: Optional[int] = None):
        lib = self.lib
        C.c_long.in_dll(lib, "xmi_hits_per_single").value = hits_per_single
        if seed is not None:
            lib.xmi_set_seed(seed)
        sa = C.POINTER(xmi_solid_angle)()
        if not lib.xmi_solid_angle_calculation_cl(self.inputF, C.byref(sa), None, C.byref(self.options)):
            raise RuntimeError("xmi_solid_angle_calculation_cl failed: " + last_error())
        self.set_solid_angle(sa, owner="engine")
        return self.solid_angle_array()

    def set_solid_angle(self, sa_ptr, owner: str):
        self.free_solid_angle()
        self.solid_angle = sa_ptr
        self._sa_owner = owner

    def solid_angle_array(self):
        sa = self.solid_angle.contents
        nr, nt = sa.grid_dims_r_n, sa.grid_dims_theta_n
        grid = np.ctypeslib.as_array(sa.solid_angles, shape=(nt, nr)).copy()
        r = np.ctypeslib.as_array(sa.grid_dims_r_vals, shape=(nr,)).copy()
        t = np.ctypeslib.as_array(sa.grid_dims_theta_vals, shape=(nt,)).copy()
        return grid, r, t

    def free_solid_angle(self):
        if self.solid_angle is not None and getattr(self, "_sa_owner", "") == "engine":
            self.lib.xmi_free_solid_angle(self.solid_angle)
        self.solid_angle = None

    def run(self, seed: int, n_mpi_hosts: int = 1):
        """xmi_main_msim on the GPU; returns (channels, var_red_history,
        brute_history, stats).  Without variance reduction (the brute-force
        path) var_red_history is None and no solid-angle grid is needed."""
        lib = self.lib
        if self.solid_angle is None and self.options.use_variance_reduction:
            raise RuntimeError("no solid-angle grid; call compute_solid_angle() first")
        lib.xmi_set_seed(seed)
        ch = C.POINTER(C.c_double)()
        bh = C.POINTER(C.c_double)()
        vr = C.POINTER(C.c_double)()
        ok = lib.xmi_main_msim(self.inputF, self.hdf5F, n_mpi_hosts, C.byref(ch), C.byref(self.options), C.byref(bh),
                               C.byref(vr), self.solid_angle)
        if not ok:
            raise RuntimeError("xmi_main_msim failed: " + last_error())
        stats = xmi_msim_stats()
        lib.xmi_msim_get_stats(C.byref(stats))
        ev = (C.c_int64 * 16)()
        lib.xmi_msim_get_layer_events(ev, 16)
        self.layer_events = list(ev)[: len(self.inp.layers)]
        return _collect(lib, self.inp, ch, bh, vr) + (stats,)

    def close(self):
        lib = self.lib
        self.free_solid_angle()
        if self.hdf5F and self._owns_tables:
            lib.xmi_free_hdf5_F(C.byref(self.hdf5F))
        self.hdf5F = C.c_void_p()
        if self.inputF:
            lib.xmi_free_input_F(C.byref(self.inputF))
            self.inputF = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class xmi_escape_ratios(C.Structure):
    _fields_ = [("n_elements", C.c_int), ("n_fluo_input_energies", C.c_int), ("n_compton_input_energies", C.c_int),
                ("n_compton_output_energies", C.c_int), ("Z", C.POINTER(C.c_int)),
                ("fluo_escape_ratios", C.POINTER(C.c_double)), ("fluo_escape_input_energies", C.POINTER(C.c_double)),
                ("compton_escape_ratios", C.POINTER(C.c_double)),
                ("compton_escape_input_energies", C.POINTER(C.c_double)),
                ("compton_escape_output_energies", C.POINTER(C.c_double)), ("xmi_input_string", C.c_char_p)]


class xmi_escape_ratios_options(C.Structure):
    _fields_ = [("n_input_energies", C.c_long), ("n_compton_output_energies", C.c_long), ("n_photons", C.c_long),
                ("input_energy_min", C.c_double), ("input_energy_delta", C.c_double),
                ("compton_output_energy_min", C.c_double), ("compton_output_energy_delta", C.c_double)]


def escape_input(inp: Input) -> Input:
    """The input of an escape-ratio run (src/xmi_detector.c:98-116): the
    detector crystal as the sample (reference layer 1), source 1 cm away
    behind a 1 um slit, sample normal along z, one interaction per photon."""
    import dataclasses
    return dataclasses.replace(inp, layers=list(inp.crystal), reference_layer=1, d_sample_source=1.0,
                               d_source_slit=1.0, slit_size_x=0.0001, slit_size_y=0.0001,
                               n_sample_orientation=(0.0, 0.0, 1.0), n_interactions_trajectory=1,
                               n_photons_line=100000)


def escape_ratios(inp: Input, n_photons: int = 500000, n_input_energies: int = 1990, input_energy_min: float = 1.0,
                  input_energy_delta: float = 0.1, n_compton_output_energies: int = 1999,
                  compton_output_energy_min: float = 0.1, compton_output_energy_delta: float = 0.1,
                  seed: Optional[int] = None, hdf5_file: Optional[str] = None):
    """xmi_escape_ratios_calculation of the input's detector crystal on the
    GPU.  Returns (Z list, fluo [n_el][109][nE] as the reference's Fortran
    array indexed (element, line, energy), compton [nE][nOut] indexed
    (input, output energy), handle) -- pass the handle to detector_response
    and release it with free_escape_ratios."""
    lib = load_library()
    ero = xmi_escape_ratios_options(n_input_energies, n_compton_output_energies, n_photons, input_energy_min,
                                    input_energy_delta, compton_output_energy_min, compton_output_energy_delta)
    o = default_options()
    if seed is not None:
        lib.xmi_set_seed(seed)
    cin = CInput(inp)
    out = C.POINTER(xmi_escape_ratios)()
    lib.xmi_escape_ratios_calculation(C.byref(cin.input), C.byref(out), None,
                                      hdf5_file.encode() if hdf5_file else None, C.byref(o), ero)
    if not out:
        raise RuntimeError("xmi_escape_ratios_calculation failed: " + last_error())
    r = out.contents
    n_el, nE, nOut = r.n_elements, r.n_fluo_input_energies, r.n_compton_output_energies
    Z = [r.Z[k] for k in range(n_el)]
    fluo = np.ctypeslib.as_array(r.fluo_escape_ratios, shape=(nE, 109, n_el)).transpose(2, 1, 0).copy()
    compton = np.ctypeslib.as_array(r.compton_escape_ratios, shape=(nOut, nE)).T.copy()
    return Z, fluo, compton, out


def free_escape_ratios(handle):
    if handle:
        load_library().xmi_free_escape_ratios(handle)


def detector_response(prep: "Prepared", channels: np.ndarray, var_red: Optional[np.ndarray] = None,
                      escape_ratios=None, zero_interaction: bool = False, options: Optional[xmi_main_options] = None):
    """xmi_detector_convolute_all on the spectra of every order (channels
    [n_int+1][nch] from Prepared.run); returns the convoluted spectra and
    applies the detector absorption to var_red in place."""
    lib = load_library()
    n1, nch = channels.shape
    ch = [np.ascontiguousarray(channels[k], np.float64) for k in range(n1)]
    noconv = (C.POINTER(C.c_double) * n1)(*[c.ctypes.data_as(C.POINTER(C.c_double)) for c in ch])
    conv = (C.POINTER(C.c_double) * n1)()
    vr_ptr = None
    if var_red is not None:
        assert var_red.flags["C_CONTIGUOUS"] and var_red.dtype == np.float64
        vr_ptr = var_red.ctypes.data_as(C.POINTER(C.c_double))
    opts = options or prep.options
    lib.xmi_detector_convolute_all.argtypes = [C.c_void_p, C.POINTER(C.POINTER(C.c_double)),
                                               C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(xmi_main_options),
                                               C.POINTER(xmi_escape_ratios), C.c_int, C.c_int]
    lib.xmi_detector_convolute_all(prep.inputF, noconv, conv, None, vr_ptr, C.byref(opts),
                                   C.byref(escape_ratios) if escape_ratios is not None else None, n1 - 1,
                                   1 if zero_interaction else 0)
    if not all(conv[k] for k in range(n1)):
        raise RuntimeError("xmi_detector_convolute_all failed: " + last_error())
    out = np.zeros((n1, nch))
    for k in range(n1):
        out[k] = np.ctypeslib.as_array(conv[k], shape=(nch,))
        lib.xmi_free(conv[k])
    return out


def detector_response_noalloc(prep: "Prepared", channels: np.ndarray, var_red: Optional[np.ndarray] = None,
                              escape_ratios=None, zero_interaction: bool = False,
                              options: Optional[xmi_main_options] = None):
    """xmi_detector_convolute_all_noalloc: the same response as
    detector_response, with caller-owned outputs and no device allocation."""
    lib = load_library()
    n1, nch = channels.shape
    ch = [np.ascontiguousarray(channels[k], np.float64) for k in range(n1)]
    out = np.zeros((n1, nch))
    noconv = (C.POINTER(C.c_double) * n1)(*[c.ctypes.data_as(C.POINTER(C.c_double)) for c in ch])
    conv = (C.POINTER(C.c_double) * n1)(*[out[k].ctypes.data_as(C.POINTER(C.c_double)) for k in range(n1)])
    vr_ptr = None
    if var_red is not None:
        assert var_red.flags["C_CONTIGUOUS"] and var_red.dtype == np.float64
        vr_ptr = var_red.ctypes.data_as(C.POINTER(C.c_double))
    opts = options or prep.options
    f = lib.xmi_detector_convolute_all_noalloc
    f.argtypes = [C.c_void_p, C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.POINTER(C.c_double)),
                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(xmi_main_options),
                  C.POINTER(xmi_escape_ratios), C.c_int, C.c_int]
    f.restype = C.c_int
    if not f(prep.inputF, noconv, conv, None, vr_ptr, C.byref(opts),
             C.byref(escape_ratios) if escape_ratios is not None else None, n1 - 1, 1 if zero_interaction else 0):
        raise RuntimeError("xmi_detector_convolute_all_noalloc failed: " + last_error())
    return out


def run_batch(preps: Sequence["Prepared"], tables: "SharedTables", seed: int, n_streams: int = 4,
              options: Optional[xmi_main_options] = None, seeds: Optional[Sequence[int]] = None):
    """xmi_main_msim_batch over inputs bound to `tables`, each with a
    solid-angle grid (seeds: one per sample, else `seed` for all); returns
    ([(channels, var_red_history)], done_ms) and leaves the batch totals in
    batch_stats()."""
    lib = load_library()
    n = len(preps)
    P = C.POINTER
    lib.xmi_main_msim_batch.argtypes = [C.c_int, P(C.c_void_p), C.c_void_p, C.c_int, P(xmi_main_options),
                                        P(P(xmi_solid_angle)), P(P(C.c_double)), P(P(C.c_double)), C.c_int,
                                        P(C.c_double), P(C.c_uint64)]
    lib.xmi_main_msim_batch.restype = C.c_int
    inputs = (C.c_void_p * n)(*[p.inputF.value for p in preps])
    sas = (P(xmi_solid_angle) * n)(*[p.solid_angle for p in preps])
    ch = (P(C.c_double) * n)()
    vr = (P(C.c_double) * n)()
    done = (C.c_double * n)()
    opts = options or preps[0].options
    lib.xmi_set_seed(seed)
    sd = (C.c_uint64 * n)(*[int(x) for x in seeds]) if seeds is not None else None
    if not lib.xmi_main_msim_batch(n, inputs, tables.hdf5F, 1, C.byref(opts), sas, ch, vr, n_streams, done, sd):
        raise RuntimeError("xmi_main_msim_batch failed: " + last_error())
    out = []
    for k, p in enumerate(preps):
        n_int, nch = p.inp.n_interactions_trajectory, p.inp.nchannels
        c = np.ctypeslib.as_array(ch[k], shape=(n_int + 1, nch)).copy()
        v = np.ctypeslib.as_array(vr[k], shape=(100, 385, n_int)).copy()
        lib.xmi_free(ch[k])
        lib.xmi_free(vr[k])
        out.append((c, v))
    return out, np.array(done)


def batch_stats() -> "xmi_msim_stats":
    """Totals of the last xmi_main_msim_batch call on this thread."""
    st = xmi_msim_stats()
    load_library().xmi_msim_get_stats(C.byref(st))
    return st


class SharedTables:
    """Interaction tables of a set of elements, built once and bound to many
    inputs (each Prepared made with tables=... uses them without a copy; the
    engine keeps one device copy per table handle)."""

    def __init__(self, Zs: Sequence[int], hdf5_file: Optional[str] = None):
        Zs = sorted(set(int(z) for z in Zs))
        self.Zs = Zs
        self._holder = Prepared(Input(layers=[Layer(Zs, [1.0 / len(Zs)] * len(Zs), 1.0, 0.1)],
                                      discrete=[DiscreteLine(20.0, 1e9)], n_photons_line=10), hdf5_file=hdf5_file)
        self.hdf5F = self._holder.hdf5F

    def table(self, Z: int, which: int) -> np.ndarray:
        """One table as numpy (which: 0/1 Rayleigh/Compton angle inverse CDF
        [400][2000], 2 total Compton-profile inverse CDF, 3 partial-profile
        inversions [shells][10001], 4 azimuth inverse CDF [200][2000])."""
        lib = load_library()
        n = lib.xmi_tables_copy(self.hdf5F, Z, which, None, 0)
        if n <= 0:
            raise KeyError((Z, which))
        out = np.zeros(n, np.float64)
        lib.xmi_tables_copy(self.hdf5F, Z, which, out.ctypes.data_as(C.POINTER(C.c_double)), n)
        return out

    def close(self):
        if self._holder is not None:
            self._holder.close()
            self._holder = None
            self.hdf5F = C.c_void_p()


def _collect(lib, inp: Input, ch, bh, vr):
    n_int = inp.n_interactions_trajectory
    nch = inp.nchannels
    channels = np.ctypeslib.as_array(ch, shape=(n_int + 1, nch)).copy()
    # NULL without variance reduction (src/xmi_main.F90:934-943)
    var_red = np.ctypeslib.as_array(vr, shape=(100, 385, n_int)).copy() if vr else None
    brute = np.ctypeslib.as_array(bh, shape=(100, 385, n_int)).copy()
    lib.xmi_free(ch)
    lib.xmi_free(bh)
    lib.xmi_free(vr)
    return channels, var_red, brute
