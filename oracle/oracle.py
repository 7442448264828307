"""ctypes binding of the oracle (liboracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  It runs the CPU restatement of the
reference engine on the handles made by the product library (same process).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
RNG_PER_HISTORY = 0
RNG_PER_THREAD = 1

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class OracleWork(C.Structure):
    _fields_ = [("events", C.c_int64), ("elements", C.c_int64), ("element_layers", C.c_int64),
                ("element2_layers", C.c_int64), ("doppler_trials", C.c_int64), ("lines", C.c_int64),
                ("line_layers", C.c_int64)]


def work() -> dict:
    """Estimator work counters of the last run() (sums over scored events)."""
    w = OracleWork()
    load().xmi_oracle_get_work(C.byref(w))
    return {name: getattr(w, name) for name, _ in OracleWork._fields_}


class OracleStats(C.Structure):
    _fields_ = [("histories", C.c_int64), ("interactions", C.c_int64), ("vr_calls", C.c_int64),
                ("solid_angle_fallbacks", C.c_int64), ("seconds", C.c_double)]


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    from xmimsim_amd import api
    lib = C.CDLL(LIB)
    P = C.POINTER
    lib.xmi_oracle_main_msim.argtypes = [C.c_void_p, C.c_void_p, C.c_int, P(P(C.c_double)),
                                         P(api.xmi_main_options), P(P(C.c_double)), P(P(C.c_double)),
                                         P(api.xmi_solid_angle), C.c_int, C.c_uint64, C.c_int]
    lib.xmi_oracle_main_msim.restype = C.c_int
    lib.xmi_oracle_get_stats.argtypes = [P(OracleStats)]
    lib.xmi_oracle_get_work.argtypes = [P(OracleWork)]
    lib.xmi_oracle_solid_angle_calculation.argtypes = [C.c_void_p, P(P(api.xmi_solid_angle)), C.c_long, C.c_uint64,
                                                       C.c_int]
    lib.xmi_oracle_solid_angle_calculation.restype = C.c_int
    lib.xmi_oracle_free_solid_angle.argtypes = [P(api.xmi_solid_angle)]
    lib.xmi_oracle_solid_angle_extents.argtypes = [C.c_void_p] + [P(C.c_double)] * 4
    lib.xmi_oracle_single_solid_angle.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_long, C.c_uint64,
                                                  C.c_uint64]
    lib.xmi_oracle_single_solid_angle.restype = C.c_double
    lib.xmi_oracle_fpm_first_order.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    lib.xmi_oracle_fpm_first_order.restype = C.c_double
    lib.xmi_oracle_scatter_first_order.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.xmi_oracle_scatter_first_order.restype = C.c_double
    lib.xmi_oracle_detector_convolute.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
    lib.xmi_oracle_poly_solve_quadratic.argtypes = [C.c_double] * 3 + [P(C.c_double)] * 2
    lib.xmi_oracle_poly_solve_quadratic.restype = C.c_int
    lib.xmi_oracle_energy_from_q.argtypes = [C.c_double] * 3
    lib.xmi_oracle_energy_from_q.restype = C.c_double
    lib.xmi_oracle_q_from_energy.argtypes = [C.c_double] * 3
    lib.xmi_oracle_q_from_energy.restype = C.c_double
    lib.xmi_oracle_ran_trap.argtypes = [C.c_double] * 5
    lib.xmi_oracle_ran_trap.restype = C.c_double
    lib.xmi_oracle_philox.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    lib.xmi_oracle_set_hits_per_single.argtypes = [C.c_long]
    lib.xmi_oracle_free.argtypes = [C.c_void_p]
    lib.xmi_oracle_escape_ratios.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_double, C.c_double,
                                             C.c_long, C.c_double, C.c_double, C.c_long, C.c_uint64, C.c_int,
                                             P(C.c_double), P(C.c_double)]
    lib.xmi_oracle_escape_ratios.restype = C.c_int
    D = P(C.c_double)
    lib.xmi_oracle_probe_scatter.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, D, D,
                                             C.c_long, C.c_uint64, D]
    lib.xmi_oracle_probe_scatter.restype = C.c_int
    lib.xmi_oracle_probe_fluorescence.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                                  C.c_long, C.c_uint64, P(C.c_int), D, D]
    lib.xmi_oracle_probe_fluorescence.restype = C.c_int
    lib.xmi_oracle_probe_rotate.argtypes = [D, D, C.c_double, C.c_double, D]
    lib.xmi_oracle_probe_phi0.argtypes = [D, D]
    lib.xmi_oracle_probe_phi0.restype = C.c_double
    lib.xmi_oracle_probe_source.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_long, C.c_uint64, D]
    lib.xmi_oracle_probe_source.restype = C.c_int
    lib.xmi_oracle_input_derived.argtypes = [C.c_void_p, D]
    lib.xmi_oracle_probe_doppler.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double,
                                             C.c_long, C.c_uint64, D]
    lib.xmi_oracle_probe_doppler.restype = C.c_int
    lib.xmi_oracle_set_plant.argtypes = [C.c_int]
    lib.xmi_oracle_probe_detector_hit.argtypes = [C.c_void_p, D, D, C.c_long, C.POINTER(C.c_int)]
    lib.xmi_oracle_probe_detector_segment.argtypes = [C.c_void_p, D, D, C.c_long, C.POINTER(C.c_int)]
    _lib = lib
    return lib


def run(prepared, seed: int, rng_mode: int = RNG_PER_HISTORY, n_threads: int = 0, n_mpi_hosts: int = 1):
    """xmi_main_msim restated on the CPU; returns (channels, var_red, brute, stats);
    var_red is None for a run without variance reduction."""
    lib = load()
    if n_threads <= 0:
        n_threads = os.cpu_count() or 1
    # the reference shares hits_per_single between the grid and the in-estimator
    # Monte Carlo fallback: use the engine's current value
    from xmimsim_amd import api
    lib.xmi_oracle_set_hits_per_single(C.c_long.in_dll(api.load_library(), "xmi_hits_per_single").value)
    ch = C.POINTER(C.c_double)()
    bh = C.POINTER(C.c_double)()
    vr = C.POINTER(C.c_double)()
    ok = lib.xmi_oracle_main_msim(prepared.inputF, prepared.hdf5F, n_mpi_hosts, C.byref(ch),
                                  C.byref(prepared.options), C.byref(bh), C.byref(vr), prepared.solid_angle,
                                  rng_mode, seed, n_threads)
    if not ok:
        raise RuntimeError("oracle run failed")
    st = OracleStats()
    lib.xmi_oracle_get_stats(C.byref(st))
    n_int = prepared.inp.n_interactions_trajectory
    nch = prepared.inp.nchannels
    channels = np.ctypeslib.as_array(ch, shape=(n_int + 1, nch)).copy()
    # NULL without variance reduction, as in the reference
    var_red = np.ctypeslib.as_array(vr, shape=(100, 385, n_int)).copy() if vr else None
    brute = np.ctypeslib.as_array(bh, shape=(100, 385, n_int)).copy()
    for p in (ch, bh, vr):
        lib.xmi_oracle_free(p)
    return channels, var_red, brute, st


def escape_ratios(prepared, n_photons: int, n_input_energies: int, input_energy_min: float,
                  input_energy_delta: float, n_compton_output_energies: int, compton_output_energy_min: float,
                  compton_output_energy_delta: float, seed: int, n_threads: int = 0):
    """Escape-ratio run restated on the CPU (prepared: the escape input, see
    api.escape_input, with escape options); returns (fluo [n_el][109][nE],
    compton [nE][nOut])."""
    lib = load()
    n_el = len(prepared.inp.unique_Z())
    fluo = np.zeros(n_el * 109 * n_input_energies)
    comp = np.zeros(n_input_energies * n_compton_output_energies)
    ok = lib.xmi_oracle_escape_ratios(prepared.inputF, prepared.hdf5F, C.byref(prepared.options), n_input_energies,
                                      input_energy_min, input_energy_delta, n_compton_output_energies,
                                      compton_output_energy_min, compton_output_energy_delta, n_photons, seed,
                                      n_threads or (os.cpu_count() or 1),
                                      fluo.ctypes.data_as(C.POINTER(C.c_double)),
                                      comp.ctypes.data_as(C.POINTER(C.c_double)))
    if not ok:
        raise RuntimeError("oracle escape-ratio run failed")
    fluo = fluo.reshape(n_input_energies, 109, n_el).transpose(2, 1, 0).copy()
    comp = comp.reshape(n_compton_output_energies, n_input_energies).T.copy()
    return fluo, comp


def solid_angle(prepared, hits: int, seed: int, n_threads: int = 0):
    """Oracle grid; installs it on `prepared` (owner: oracle) and returns arrays."""
    lib = load()
    from xmimsim_amd import api
    sa = C.POINTER(api.xmi_solid_angle)()
    C.c_long.in_dll(api.load_library(), "xmi_hits_per_single").value = hits
    if not lib.xmi_oracle_solid_angle_calculation(prepared.inputF, C.byref(sa), hits, seed,
                                                  n_threads or (os.cpu_count() or 1)):
        raise RuntimeError("oracle solid angle failed")
    prepared.set_solid_angle(sa, owner="oracle")
    prepared._oracle_sa = sa
    return prepared.solid_angle_array()


def fpm_first_order(prepared, Z: int, line: int, cascade: int = 4) -> float:
    return load().xmi_oracle_fpm_first_order(prepared.inputF, Z, line, cascade)


def scatter_first_order(prepared, Z: int, kind: str) -> float:
    """Closed-form first-order Rayleigh ("rayl") or Compton ("compt") intensity."""
    return load().xmi_oracle_scatter_first_order(prepared.inputF, Z, {"rayl": 0, "compt": 1}[kind])


def detector_convolute(prepared, noconv):
    """Absorption and Gaussian response of one order's spectrum."""
    x = np.ascontiguousarray(noconv, np.float64)
    out = np.zeros_like(x)
    load().xmi_oracle_detector_convolute(prepared.inputF, x.ctypes.data_as(C.POINTER(C.c_double)),
                                         out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


# ---- probes of single restated routines (pinning tests) --------------------
PLANTS = {"none": 0, "rayleigh_energy": 1, "phi0_sign": 2, "compton_full_angle": 3, "doppler_one_sided": 4,
          "ck_l1_swap": 5, "line_shift": 6, "elecv_stale": 7, "dirv_phi_sign": 8, "slit_full": 9, "gauss_xy": 10,
          "compton_t_rayleigh": 11, "shell_next": 12, "qimax_ignored": 13, "doppler_no_negative": 14,
          "brute_blbs": 15, "auger_one_vacancy": 16, "radiative_shell": 17, "no_collimator": 18}


class planted:
    """Context manager: runs the oracle with one planted mistake."""

    def __init__(self, name: str):
        self.id = PLANTS[name]

    def __enter__(self):
        load().xmi_oracle_set_plant(self.id)
        return self

    def __exit__(self, *exc):
        load().xmi_oracle_set_plant(0)
        return False


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def probe_scatter(prepared, kind: str, element_index: int, energy: float, dirv, elecv, n: int, seed: int):
    """Rayleigh ("rayl") or Compton ("compt") interaction n times from one
    state; returns (new dirv [n,3], new elecv [n,3], energy [n])."""
    d = np.ascontiguousarray(dirv, np.float64)
    e = np.ascontiguousarray(elecv, np.float64)
    out = np.zeros((n, 7))
    ok = load().xmi_oracle_probe_scatter(prepared.inputF, prepared.hdf5F, C.byref(prepared.options),
                                         {"rayl": 1, "compt": 2}[kind], element_index, energy, _d(d), _d(e), n, seed,
                                         _d(out))
    if not ok:
        raise RuntimeError("probe failed")
    return out[:, :3], out[:, 3:6], out[:, 6]


def probe_fluorescence(prepared, element_index: int, energy: float, n: int, seed: int, directions: bool = False):
    """Photo-absorption + fluorescence n times; returns (line [n] (0 = none),
    weight factor [n], and with directions: dirv [n,3], elecv [n,3])."""
    line = np.zeros(n, np.int32)
    w = np.zeros(n)
    dirs = np.zeros((n, 6)) if directions else None
    ok = load().xmi_oracle_probe_fluorescence(prepared.inputF, prepared.hdf5F, C.byref(prepared.options),
                                              element_index, energy, n, seed,
                                              line.ctypes.data_as(C.POINTER(C.c_int)), _d(w),
                                              _d(dirs) if directions else None)
    if not ok:
        raise RuntimeError("probe failed")
    if directions:
        return line, w, dirs[:, :3], dirs[:, 3:]
    return line, w


def probe_rotate(dirv, elecv, theta: float, phi: float):
    d = np.ascontiguousarray(dirv, np.float64)
    e = np.ascontiguousarray(elecv, np.float64)
    out = np.zeros(6)
    load().xmi_oracle_probe_rotate(_d(d), _d(e), theta, phi, _d(out))
    return out[:3], out[3:]


def probe_phi0(dirv, elecv) -> float:
    d = np.ascontiguousarray(dirv, np.float64)
    e = np.ascontiguousarray(elecv, np.float64)
    return load().xmi_oracle_probe_phi0(_d(d), _d(e))


def probe_source(prepared, which: int, n: int, seed: int):
    """Primary photons of discrete line `which` (>= 0) or continuous interval
    -which-1; returns (entry point [n,3], dirv, elecv, energy [n])."""
    out = np.zeros((n, 10))
    ok = load().xmi_oracle_probe_source(prepared.inputF, prepared.hdf5F, C.byref(prepared.options), which, n, seed,
                                        _d(out))
    if not ok:
        raise RuntimeError("probe failed")
    return out[:, :3], out[:, 3:6], out[:, 6:9], out[:, 9]


def input_derived(prepared):
    """Derived fields of xmi_init_input as a dict."""
    nl = len(prepared.inp.layers)
    out = np.zeros(35 + 3 * nl)
    load().xmi_oracle_input_derived(prepared.inputF, _d(out))
    return {"detector_radius": out[0], "collimator_present": out[1], "collimator_radius": out[2],
            "half_apex": out[3], "vertex": out[4:7], "n_sample_orientation_det": out[7:10],
            "n_sample_orientation": out[10:13], "n_detector_orientation": out[13:16],
            "new": out[16:25].reshape(3, 3), "inverse": out[25:34].reshape(3, 3), "detector_solid_angle": out[34],
            "thickness_along_Z": out[35::3], "Z_coord_begin": out[36::3], "Z_coord_end": out[37::3]}


def probe_doppler(prepared, element_index: int, energy: float, theta: float, n: int, seed: int):
    """The estimator's Doppler-broadened Compton energies at (energy, theta)."""
    out = np.zeros(n)
    if not load().xmi_oracle_probe_doppler(prepared.inputF, prepared.hdf5F, C.byref(prepared.options), element_index,
                                           energy, theta, n, seed, _d(out)):
        raise RuntimeError("probe failed")
    return out


def probe_detector_hit(prepared, coords, dirv):
    """1 where a photon that left the sample at coords moving along dirv
    reaches the detector (lab frame, arrays of shape (n, 3))."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    dirv = np.ascontiguousarray(dirv, dtype=np.float64)
    out = np.zeros(len(coords), dtype=np.int32)
    load().xmi_oracle_probe_detector_hit(prepared.inputF, _d(coords), _d(dirv), len(coords),
                                         out.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def probe_detector_segment(prepared, begin, end):
    """Codes of the detector test of path segments begin -> end (lab frame):
    0 none, 1 collimator, 2 detector window, 3 window from behind."""
    begin = np.ascontiguousarray(begin, dtype=np.float64)
    end = np.ascontiguousarray(end, dtype=np.float64)
    out = np.zeros(len(begin), dtype=np.int32)
    load().xmi_oracle_probe_detector_segment(prepared.inputF, _d(begin), _d(end), len(begin),
                                             out.ctypes.data_as(C.POINTER(C.c_int)))
    return out
