"""Net-line surrogate: ctypes binding of the GPU MLP ensemble
(xmi_surrogate_* in include/xmimsim_amd.h) and the XRF encoding of its inputs
and outputs.

Inputs (85): weight fractions of Z = 11..92 on a logarithmic scale
(1 + log10(w) / 4 for present elements, 0 for absent ones: line intensities
go roughly as log w), log10 thickness [cm], density / 20 [g/cm3], excitation
energy / 40 [keV] of a single-layer sample.
Outputs (256): log10 of net line intensities summed over interaction orders
(the var_red_history rows of xmi_main_msim) for the lines in OUT_LINES;
absent elements and zero intensities are masked out of the loss.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence, Tuple

import numpy as np

from . import api

Z_MIN, Z_MAX = 11, 92
N_IN = (Z_MAX - Z_MIN + 1) + 3
# xraylib line numbers (var_red_history row = line - 1)
KL3, KM3, L3M5, M5N7 = 3, 15, 90, 216


def _out_lines() -> List[Tuple[int, int]]:
    lines = []
    for z in range(Z_MIN, Z_MAX + 1):
        lines += [(z, KL3), (z, L3M5), (z, M5N7)]
    lines += [(z, KM3) for z in range(Z_MIN, Z_MIN + 10)]
    return lines


OUT_LINES = _out_lines()
N_OUT = len(OUT_LINES)  # 256
DEFAULT_DIMS = (N_IN, 512, 512, 512, 512, N_OUT)


def encode(Z: Sequence[int], w: Sequence[float], thickness: float, density: float, energy: float) -> np.ndarray:
    x = np.zeros(N_IN, np.float32)
    for z, f in zip(Z, w):
        x[z - Z_MIN] = 1.0 + np.log10(max(float(f), 1e-4)) / 4.0
    x[-3] = np.log10(thickness)
    x[-2] = density / 20.0
    x[-1] = energy / 40.0
    return x


_OUT_Z = np.array([z for z, _ in OUT_LINES])
_OUT_ROW = np.array([line - 1 for _, line in OUT_LINES])


def targets(var_red_history: np.ndarray, Z: Sequence[int]) -> Tuple[np.ndarray, np.ndarray]:
    """log10 net-line intensities (summed over orders) and the loss mask."""
    v = var_red_history[_OUT_Z - 1, _OUT_ROW, :].sum(axis=-1)
    on = np.isin(_OUT_Z, np.asarray(Z)) & (v > 0.0)
    t = np.zeros(N_OUT, np.float32)
    t[on] = np.log10(v[on])
    return t, on.astype(np.float32)


# shell of each output line's initial vacancy (0 K .. 8 M5)
_LINE_SHELL = {KL3: 0, KM3: 0, L3M5: 3, M5N7: 8}


class LineBaseline:
    """First-order fundamental-parameter estimate of every output line from
    the encoded inputs: log10 of w_Z * P_s(E0) * omega_s * R_line * (1 -
    exp(-chi rho d)) / chi with chi = sqrt(2) (mu(E0) + mu(E_line)), the
    primary fluorescence of a homogeneous layer in the 45/45 degree geometry
    (P_s: the provider's cascade-enhanced shell cross section; mu: the
    mixture's mass attenuation).  The network learns the rest (secondary
    fluorescence, scattering, detector and geometry factors) as a correction
    to it; 0 where the line cannot be excited.  Tables on an excitation grid of
    DE keV, interpolated linearly."""
    E_LO, E_HI, DE = 9.5, 40.5, 0.005

    def __init__(self):
        from . import lines as L
        lib = api.load_library()
        self.E = np.arange(self.E_LO, self.E_HI + self.DE / 2, self.DE)
        Zs = range(Z_MIN, Z_MAX + 1)
        self.sig = np.array([[lib.xmi_provider_cs(0, z, e) for e in self.E] for z in Zs])
        e_line = np.array([L.line_energy(z, line) for z, line in OUT_LINES])
        self.sig_line = np.array([[lib.xmi_provider_cs(0, z, e) if e > 0.0 else 0.0 for e in e_line] for z in Zs])
        self.wr = np.array([lib.xmi_provider_line_data(z, 0, _LINE_SHELL[line]) * lib.xmi_provider_line_data(z, 1, line)
                            for z, line in OUT_LINES])
        ps = {}
        for z in sorted({z for z, _ in OUT_LINES}):
            ps[z] = np.array([L.shell_cross_sections(z, e, 4) for e in self.E])  # [nE][9]
        self.ps = np.array([ps[z][:, _LINE_SHELL[line]] for z, line in OUT_LINES])  # [256][nE]
        self.ok = (e_line > 0.0) & (self.wr > 0.0)
        # output lines of each element that the estimate covers, padded with -1
        cols = [np.flatnonzero((_OUT_Z == z) & self.ok) for z in Zs]
        self.cols = np.full((len(cols), max(len(c) for c in cols)), -1)
        for k, c in enumerate(cols):
            self.cols[k, :len(c)] = c

    def __call__(self, X: np.ndarray, chunk: int = 16384) -> np.ndarray:
        X = np.asarray(X)
        out = np.zeros((X.shape[0], N_OUT), np.float32)
        for s in range(0, X.shape[0], chunk):
            out[s:s + chunk] = self._block(np.asarray(X[s:s + chunk], np.float64))
        return out

    def _block(self, X: np.ndarray) -> np.ndarray:
        # sparse over the rows' own elements (a dozen of 82) and their lines
        n_, z_ = np.nonzero(X[:, :Z_MAX - Z_MIN + 1] > 0.0)
        wz = 10.0 ** (4.0 * (X[n_, z_] - 1.0))
        d, rho, E0 = 10.0 ** X[:, -3], 20.0 * X[:, -2], 40.0 * X[:, -1]
        t = np.clip((E0 - self.E_LO) / self.DE, 0.0, len(self.E) - 1.000001)
        i0 = t.astype(np.int64)
        f = t - i0
        mu0 = np.bincount(n_, wz * (self.sig[z_, i0[n_]] * (1.0 - f[n_]) + self.sig[z_, i0[n_] + 1] * f[n_]),
                          minlength=X.shape[0])
        w = np.zeros((X.shape[0], Z_MAX - Z_MIN + 1))
        w[n_, z_] = wz
        cz = self.cols[z_]
        e, k = np.nonzero(cz >= 0)
        r, c = n_[e], cz[e, k]
        chi = np.sqrt(2.0) * (mu0[r] + (w @ self.sig_line)[r, c])
        P = self.ps[c, i0[r]] * (1.0 - f[r]) + self.ps[c, i0[r] + 1] * f[r]
        x = chi * (rho * d)[r]
        out = np.zeros((X.shape[0], N_OUT))
        with np.errstate(divide="ignore", invalid="ignore"):
            A = np.where(chi > 0.0, -np.expm1(-x) / chi, 0.0)
            v = wz[e] * P * self.wr[c] * A
            out[r, c] = np.where(v > 0.0, np.log10(v), 0.0)
        return out


class NetLineModel:
    """The surrogate of the quantification: the MLP ensemble's output is a
    correction added to LineBaseline, so that predict(X) gives log10 net-line
    intensities; train on targets(...) minus baseline(X) (quant.train does)."""

    def __init__(self, net: "Surrogate", baseline: "LineBaseline" = None):
        self.net = net
        self.baseline = baseline or LineBaseline()
        self.n_nets = net.n_nets

    def predict(self, X: np.ndarray) -> np.ndarray:
        return self.net.predict(X) + self.baseline(X)[None]

    def close(self):
        self.net.close()


def _lib():
    lib = api.load_library()
    if not getattr(lib, "_sg_ready", False):
        P = C.POINTER
        fp = P(C.c_float)
        lib.xmi_surrogate_create.argtypes = [C.c_int, C.c_int, P(C.c_int), C.c_int, C.c_uint64]
        lib.xmi_surrogate_create.restype = C.c_void_p
        lib.xmi_surrogate_free.argtypes = [C.c_void_p]
        lib.xmi_surrogate_param_count.argtypes = [C.c_void_p]
        lib.xmi_surrogate_param_count.restype = C.c_long
        lib.xmi_surrogate_padded_width.argtypes = [C.c_void_p, C.c_int]
        lib.xmi_surrogate_padded_width.restype = C.c_int
        lib.xmi_surrogate_train_step.argtypes = [C.c_void_p, fp, fp, fp, C.c_int, C.c_float, fp]
        lib.xmi_surrogate_train_step.restype = C.c_int
        lib.xmi_surrogate_set_data.argtypes = [C.c_void_p, fp, fp, fp, C.c_long]
        lib.xmi_surrogate_set_data.restype = C.c_int
        lib.xmi_surrogate_train_resident.argtypes = [C.c_void_p, P(C.c_int32), C.c_int, C.c_int, fp, fp]
        lib.xmi_surrogate_train_resident.restype = C.c_int
        lib.xmi_surrogate_gradients.argtypes = [C.c_void_p, fp, fp, fp, C.c_int, fp, fp]
        lib.xmi_surrogate_gradients.restype = C.c_int
        lib.xmi_surrogate_predict.argtypes = [C.c_void_p, fp, C.c_int, fp]
        lib.xmi_surrogate_predict.restype = C.c_int
        lib.xmi_surrogate_get_params.argtypes = [C.c_void_p, fp]
        lib.xmi_surrogate_get_params.restype = C.c_int
        lib.xmi_surrogate_set_params.argtypes = [C.c_void_p, fp]
        lib.xmi_surrogate_set_params.restype = C.c_int
        lib._sg_ready = True
    return lib


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Surrogate:
    """Ensemble of n_nets MLPs on the GPU (BF16 matrix-core products, FP32
    master weights, Adam)."""

    def __init__(self, n_nets: int = 4, dims: Sequence[int] = DEFAULT_DIMS, max_batch: int = 4096, seed: int = 0):
        self.lib = _lib()
        self.n_nets, self.dims, self.max_batch = n_nets, tuple(dims), max_batch
        d = (C.c_int * len(dims))(*dims)
        self.h = self.lib.xmi_surrogate_create(n_nets, len(dims) - 1, d, max_batch, seed)
        if not self.h:
            raise RuntimeError("xmi_surrogate_create failed: " + api.last_error())
        self.n_params = self.lib.xmi_surrogate_param_count(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.xmi_surrogate_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def layer_shapes(self) -> List[Tuple[int, int]]:
        """(N, K padded) of each weight matrix."""
        return [(self.dims[l + 1], self.lib.xmi_surrogate_padded_width(self.h, l)) for l in range(len(self.dims) - 1)]

    def _check(self, ok, what):
        if not ok:
            raise RuntimeError(f"{what} failed: " + api.last_error())

    def train_step(self, X: np.ndarray, T: np.ndarray, M: np.ndarray, lr: float) -> np.ndarray:
        X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, M))
        loss = np.zeros(self.n_nets, np.float32)
        self._check(self.lib.xmi_surrogate_train_step(self.h, _fp(X), _fp(T), _fp(M), X.shape[0], lr, _fp(loss)),
                    "train_step")
        return loss

    def set_data(self, X: np.ndarray, T: np.ndarray, M: np.ndarray):
        """Keeps the training rows on the GPU for train_resident."""
        X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, M))
        self._check(self.lib.xmi_surrogate_set_data(self.h, _fp(X), _fp(T), _fp(M), X.shape[0]), "set_data")
        self.n_data = X.shape[0]

    def train_resident(self, idx: np.ndarray, lr: np.ndarray) -> np.ndarray:
        """Adam steps on rows idx[steps][batch] of the resident data with
        learning rates lr[steps]; returns loss[steps][n_nets]."""
        idx = np.ascontiguousarray(idx, np.int32)
        lr = np.ascontiguousarray(lr, np.float32)
        steps, batch = idx.shape
        loss = np.zeros((steps, self.n_nets), np.float32)
        self._check(self.lib.xmi_surrogate_train_resident(self.h, idx.ctypes.data_as(C.POINTER(C.c_int32)), steps,
                                                           batch, _fp(lr), _fp(loss)), "train_resident")
        return loss

    def gradients(self, X: np.ndarray, T: np.ndarray, M: np.ndarray):
        X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, M))
        loss = np.zeros(self.n_nets, np.float32)
        g = np.zeros((self.n_nets, self.n_params), np.float32)
        self._check(self.lib.xmi_surrogate_gradients(self.h, _fp(X), _fp(T), _fp(M), X.shape[0], _fp(loss), _fp(g)),
                    "gradients")
        return loss, g

    def predict(self, X: np.ndarray) -> np.ndarray:
        """Y[n_nets][rows][out]; rows are padded to the kernels' multiple of 64."""
        X = np.ascontiguousarray(X, np.float32)
        n = X.shape[0]
        out = np.zeros((self.n_nets, n, self.dims[-1]), np.float32)
        for s in range(0, n, self.max_batch):
            xb = X[s:s + self.max_batch]
            b = xb.shape[0]
            bp = max(64, (b + 63) // 64 * 64)
            if bp != b:
                xb = np.concatenate([xb, np.zeros((bp - b, X.shape[1]), np.float32)])
            y = np.zeros((self.n_nets, bp, self.dims[-1]), np.float32)
            self._check(self.lib.xmi_surrogate_predict(self.h, _fp(np.ascontiguousarray(xb)), bp, _fp(y)), "predict")
            out[:, s:s + b] = y[:, :b]
        return out

    def get_params(self) -> np.ndarray:
        p = np.zeros((self.n_nets, self.n_params), np.float32)
        self._check(self.lib.xmi_surrogate_get_params(self.h, _fp(p)), "get_params")
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        self._check(self.lib.xmi_surrogate_set_params(self.h, _fp(p)), "set_params")

    def unpack(self, flat: np.ndarray):
        """[(W [N][K padded], b [N])] of one net's flat parameter vector."""
        out, o = [], 0
        for N, K in self.layer_shapes():
            W = flat[o:o + N * K].reshape(N, K)
            o += N * K
            b = flat[o:o + N]
            o += N
            out.append((W, b))
        return out
