"""The reference's FWI iteration on the GPU (compat mode).

One call of ``CompatFWI.iterate`` is one pass of the loop body of
satan_i/fwi3d_cpu_inv_fwi_v1.4.c:27-109:

1. every source runs forward with a volume DFT at the reference's frequency
   bins w in [start_f, upper_f) (omega_w = w 2 pi/(it dt)); the volume spectra
   become E_ff = (E/J) T (laplaceToFreq_2) and the receiver residuals
   dE = Eobs - E_ff (residualE);
2. every receiver x direction runs a reciprocal Green's-function simulation
   (backpropagation: source at the receiver, waveform signal_x) whose spectra
   G = E/J (laplaceToFreq_back) are streamed into every source's sensitivity
   and conj-residual correlation (sensitivity, conv_GdelE);
3. calc_gamma gives the gradient, calc_phi the test model
   sigma - scaler grad; every source runs again on the test model and
   calc_alpha forms the step length; update_para2 updates sigma.

The reference stores every Green's volume and every time step; here the
Green's spectra are consumed as each run finishes and the DFTs run during the
time loop, so memory is O(sources x bins x cells).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check
from .fwi3d import Grid, Model


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class IterationResult:
    grad: np.ndarray           # calc_gamma, [iz, iy, ix]
    scaler: float              # calc_phi
    sigma_test: np.ndarray     # sigma - scaler grad
    alpha: float               # update_para2
    sigma: np.ndarray          # updated model
    err_sum: float             # residualE: sum over sources, receivers, bins of |dE|
    residuals: List[np.ndarray]  # per source complex [3, nrec, upper_f]


class CompatFWI:
    """Reference-style FWI over one grid: sources (z-directed electric dipoles
    by default, as fwdpropagation's read_source_3d(EZ, ...)), receivers,
    fictitious waveform `signal_x` (source and Green's runs) and the true
    diffusive waveform `signal_true` (T = its FFTW_BACKWARD spectrum)."""

    def __init__(self, grid: Grid, sources, receivers, it: int, signal_x: np.ndarray,
                 signal_true: np.ndarray, start_f: int, upper_f: int, src_comp: int = 2,
                 minval: float = 0.01, maxval: float = 10.0, iseabed: int = 0):
        self.grid = grid
        self.sources = np.ascontiguousarray(np.asarray(sources, dtype=np.int32).reshape(-1, 3))
        self.receivers = np.ascontiguousarray(np.asarray(receivers, dtype=np.int32).reshape(-1, 3))
        self.it = int(it)
        self.signal_x = np.ascontiguousarray(signal_x, dtype=np.float64)
        self.signal_true = np.ascontiguousarray(signal_true, dtype=np.float64)
        self.start_f, self.upper_f = int(start_f), int(upper_f)
        self.src_comp = int(src_comp)
        self.minval, self.maxval, self.iseabed = float(minval), float(maxval), int(iseabed)
        self.model = Model(grid)
        self.test_model = Model(grid)
        self._h = C.c_void_p()
        check(_lib.load().fwi3d_compat_create(
            self.model._h, len(self.sources), len(self.receivers), _ptr(self.receivers),
            self.start_f, self.upper_f, self.it, _ptr(self.signal_x), _ptr(self.signal_true),
            C.byref(self._h)))
        self.omegas = np.zeros(self.upper_f - self.start_f)
        check(_lib.load().fwi3d_compat_omegas(self._h, _ptr(self.omegas)))

    def close(self):
        if getattr(self, "_h", None):
            check(_lib.load().fwi3d_compat_destroy(self._h))
            self._h = None
            self.model.close()
            self.test_model.close()

    __del__ = close

    def _run(self, model: Model, src, comp: int):
        sh = model.shot(self.it, self.signal_x, tuple(int(v) for v in src), comp, self.receivers,
                        self.omegas)
        sh.run()
        sh.sync()
        return sh

    def forward_receivers(self, model: Model, src) -> np.ndarray:
        """E_ff at the receivers (laplaceToFreq_3), complex [3, nrec, upper_f]."""
        sh = self._run(model, src, self.src_comp)
        out = np.zeros((3, len(self.receivers), self.upper_f), dtype=np.complex128)
        check(_lib.load().fwi3d_compat_receivers(self._h, sh._h, _ptr(out)))
        sh.close()
        return out

    def iterate(self, sigma: np.ndarray, eobs: Sequence[np.ndarray]) -> IterationResult:
        L = _lib.load()
        nrec = len(self.receivers)
        self.model.media_coeff_3d(sigma)
        check(L.fwi3d_compat_reset(self._h))
        err = C.c_double(0.0)
        residuals = []
        for s, src in enumerate(self.sources):
            ob = np.ascontiguousarray(eobs[s], dtype=np.complex128).reshape(3, nrec, self.upper_f)
            de = np.zeros_like(ob)
            sh = self._run(self.model, src, self.src_comp)
            check(L.fwi3d_compat_source(self._h, s, sh._h, _ptr(ob), _ptr(de), C.byref(err)))
            sh.close()
            residuals.append(de)
        for r in range(nrec):
            for d in range(3):
                g = self._run(self.model, self.receivers[r], d)
                check(L.fwi3d_compat_green(self._h, r, d, g._h))
                g.close()
        grad = np.zeros(self.grid.ncells)
        check(L.fwi3d_compat_gamma(self._h, _ptr(grad)))
        scaler = C.c_double(0.0)
        sig2 = np.zeros(self.grid.ncells)
        check(L.fwi3d_compat_phi(self._h, self.iseabed, C.byref(scaler), _ptr(sig2)))
        self.test_model.media_coeff_3d(sig2.reshape(self.grid.shape))
        k1, k2 = np.zeros(2), np.zeros(2)
        for s, src in enumerate(self.sources):
            eff2 = self.forward_receivers(self.test_model, src)
            check(L.fwi3d_compat_alpha(self._h, s, _ptr(eff2), _ptr(k1), _ptr(k2)))
        alpha = C.c_double(0.0)
        new = np.zeros(self.grid.ncells)
        check(L.fwi3d_compat_update(self._h, _ptr(k1), _ptr(k2), self.minval, self.maxval,
                                    self.iseabed, C.byref(alpha), _ptr(new)))
        shp = self.grid.shape
        return IterationResult(grad.reshape(shp), scaler.value, sig2.reshape(shp), alpha.value,
                               new.reshape(shp), err.value, residuals)
