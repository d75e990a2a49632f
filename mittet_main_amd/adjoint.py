"""Gradient of a receiver misfit with respect to log(sigma), by the exact
discrete adjoint of the forward scheme (libmittet_fdtd.so, fwi3d_grad_*).

Misfit (L2 on complex receiver E at a few frequencies):
    D[c, r, f] = sum_n T[c, r, n] * w_n(f),   w_n(f) = dt exp((i-1) sqrt(omega0 omega_f) n dt)
    J = 1/2 sum_{c,r,f} |D - d_obs|^2
T are the E traces (sample n after step n).  dJ/dT[c,r,n] = Re sum_f conj(D-d) w_n(f)
drives the adjoint sweep; the library returns dJ/dsigma per cell, here turned
into dJ/dlog(sigma) = sigma * dJ/dsigma.

``MisfitFunction`` exposes it to torch.autograd (torch only carries the
gradient; the work runs on the GPU through the C ABI).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import check
from .fwi3d import Grid, Model


def dft_weights(it: int, dt: float, omega0: float, omegas) -> np.ndarray:
    n = np.arange(it)[:, None]
    a = np.sqrt(omega0 * np.asarray(omegas, dtype=np.float64))[None, :]
    return dt * np.exp((-1.0 + 1.0j) * a * n * dt)


def receiver_spectra(traces_e: np.ndarray, w: np.ndarray) -> np.ndarray:
    """(3, nrec, it) real traces -> (3, nrec, nf) complex spectra."""
    return traces_e @ w


def misfit_terms(traces_e: np.ndarray, d_obs: np.ndarray, w: np.ndarray):
    """J and dJ/dT for one shot (compensated summation for J)."""
    res = receiver_spectra(traces_e, w) - d_obs
    J = 0.5 * math.fsum(np.concatenate([(res.real ** 2).ravel(), (res.imag ** 2).ravel()]))
    # dJ/dT[c,r,n] = Re sum_f conj(res[c,r,f]) w[n,f]
    g = (res.real @ w.real.T) + (res.imag @ w.imag.T)
    return J, g


class ShotSpec:
    def __init__(self, src, signal: np.ndarray, src_comp: int = 0):
        self.src = tuple(int(v) for v in src)
        self.signal = np.ascontiguousarray(signal, dtype=np.float64)
        self.src_comp = int(src_comp)


class GradientProblem:
    """Misfit J(sigma) over several shots sharing a grid, receivers and frequencies.

    The model and the shots stay resident between evaluations; each evaluation
    uploads the new sigma (media_coeff_3d).  Misfit-only evaluations run each
    shot's plain time loop.  A gradient evaluation takes the shots one at a
    time: it attaches checkpoint storage to the shot (one state every `segment`
    steps plus one replayed segment), runs the checkpointed forward pass and the
    adjoint sweep, and frees the storage, so the HBM held at once is one shot's
    checkpoints whatever the number of shots."""

    def __init__(self, grid: Grid, shots: Sequence[ShotSpec], receivers, it: int, omegas,
                 d_obs: Optional[List[np.ndarray]] = None, segment: int = 0):
        self.grid = grid
        self.it = int(it)
        self.segment = int(segment)
        self.omegas = np.asarray(omegas, dtype=np.float64)
        self.receivers = np.ascontiguousarray(np.asarray(receivers, dtype=np.int32).reshape(-1, 3))
        self.model = Model(grid)
        self.specs = list(shots)
        self.shots = [self.model.shot(self.it, sp.signal, sp.src, sp.src_comp, self.receivers)
                      for sp in self.specs]
        self.w = dft_weights(self.it, grid.dt, grid.omega0, self.omegas)
        self.d_obs = d_obs
        self.peak_storage = 0

    def close(self):
        for sh in self.shots:
            sh.close()
        self.shots = []
        self.model.close()

    def storage_bytes(self) -> int:
        """Largest checkpoint storage held during a gradient evaluation so far."""
        return self.peak_storage

    def forward_data(self, sigma: np.ndarray) -> List[np.ndarray]:
        """Complex receiver spectra (3, nrec, nf) of every shot for sigma."""
        self.model.media_coeff_3d(sigma)
        out = []
        for sh in self.shots:
            sh.set_zero_eh()
            sh.run()
            out.append(receiver_spectra(sh.traces()[:3], self.w))
        return out

    def misfit(self, sigma: np.ndarray) -> float:
        data = self.forward_data(sigma)
        return 0.5 * math.fsum(float(np.sum(np.abs(d - o) ** 2)) for d, o in zip(data, self.d_obs))

    def value_and_grad(self, log_sigma: np.ndarray) -> Tuple[float, np.ndarray]:
        """J and dJ/dlog(sigma) (same shape as log_sigma)."""
        sigma = np.exp(np.asarray(log_sigma, dtype=np.float64))
        self.model.media_coeff_3d(sigma)
        L = _lib.load()
        total, terms = np.zeros(sigma.size), []
        for sh, obs in zip(self.shots, self.d_obs):
            h = C.c_void_p()
            check(L.fwi3d_grad_create(sh._h, self.segment, C.byref(h)))
            try:
                self.peak_storage = max(self.peak_storage, int(L.fwi3d_grad_bytes(h)))
                check(L.fwi3d_grad_forward(h))
                J, g = misfit_terms(sh.traces()[:3], obs, self.w)
                terms.append(J)
                dj = np.zeros(sigma.size)
                gg = np.ascontiguousarray(g)
                check(L.fwi3d_grad_backward(h, gg.ctypes.data_as(C.c_void_p),
                                            dj.ctypes.data_as(C.c_void_p)))
            finally:
                check(L.fwi3d_grad_destroy(h))
            total += dj
        return math.fsum(terms), (total * sigma.reshape(-1)).reshape(np.shape(log_sigma))


def _torch_function():
    import torch

    class MisfitFunction(torch.autograd.Function):
        """J(log sigma) with an exact adjoint gradient.  Input: float64 CPU tensor
        of shape (iz, iy, ix); output: 0-d tensor."""

        @staticmethod
        def forward(ctx, log_sigma, problem: GradientProblem):
            J, g = problem.value_and_grad(log_sigma.detach().cpu().numpy())
            ctx.save_for_backward(torch.from_numpy(g))
            return log_sigma.new_tensor(J)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g.to(grad_out.dtype), None

    return MisfitFunction


def misfit(log_sigma, problem: GradientProblem):
    """torch entry point: misfit(log_sigma_tensor, problem) -> differentiable scalar."""
    return _torch_function().apply(log_sigma, problem)
