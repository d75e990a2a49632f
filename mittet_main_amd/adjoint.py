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


def grad_plan(nt: int, levels: int) -> Tuple[List[int], int, int]:
    """Checkpoint plan for nt steps: segment length per depth, top segments,
    stored states (fwi3d_grad_plan)."""
    sl = (C.c_int * levels)()
    ns, st = C.c_int(0), C.c_longlong(0)
    check(_lib.load().fwi3d_grad_plan(int(nt), int(levels), sl, C.byref(ns), C.byref(st)))
    return list(sl), ns.value, st.value


def levels_for_budget(shot, budget_bytes: float, max_levels: int = 8) -> int:
    """Fewest recursion levels whose gradient storage fits budget_bytes."""
    L = _lib.load()
    for lv in range(1, max_levels + 1):
        b = int(L.fwi3d_grad_plan_bytes(shot._h, lv))
        if 0 <= b <= budget_bytes:
            return lv
    raise ValueError(f"no checkpoint plan of <= {max_levels} levels fits "
                     f"{budget_bytes / 1e9:.1f} GB")


def shot_value_and_grad(shot, obs: np.ndarray, w: np.ndarray, levels: int = 1,
                        segment: int = 0) -> Tuple[float, np.ndarray, int]:
    """One shot's misfit term, dJ/dsigma (flat) and the checkpoint bytes it
    held, on the model the shot runs on.  levels = 1 checkpoints every
    `segment` steps (0: sqrt(Nt)); levels > 1 recomputes recursively."""
    L = _lib.load()
    h = C.c_void_p()
    if levels == 1:
        check(L.fwi3d_grad_create(shot._h, int(segment), C.byref(h)))
    else:
        check(L.fwi3d_grad_create_levels(shot._h, int(levels), C.byref(h)))
    try:
        nbytes = int(L.fwi3d_grad_bytes(h))
        check(L.fwi3d_grad_forward(h))
        J, g = misfit_terms(shot.traces()[:3], obs, w)
        dj = np.zeros(shot.model.grid.ncells)
        gg = np.ascontiguousarray(g)
        check(L.fwi3d_grad_backward(h, gg.ctypes.data_as(C.c_void_p),
                                    dj.ctypes.data_as(C.c_void_p)))
    finally:
        check(L.fwi3d_grad_destroy(h))
    return J, dj, nbytes


class GradientProblem:
    """Misfit J(sigma) over several shots sharing a grid, receivers and frequencies.

    The model and the shots stay resident between evaluations; each evaluation
    uploads the new sigma (media_coeff_3d).  Misfit-only evaluations run each
    shot's plain time loop.  A gradient evaluation takes the shots one at a
    time: it attaches checkpoint storage to the shot (one state every `segment`
    steps plus one replayed segment), runs the checkpointed forward pass and the
    adjoint sweep, and frees the storage, so the HBM held at once is one shot's
    checkpoints whatever the number of shots.

    `levels` > 1 selects recursive recomputation (fwi3d_grad_create_levels):
    about (levels + 1) Nt^(1/(levels+1)) stored states for levels + 1 forward
    sweeps.  With `budget_bytes` the fewest levels whose storage fits are
    chosen per evaluation (ValueError if none does).  The gradient is bitwise
    the same for every choice."""

    def __init__(self, grid: Grid, shots: Sequence[ShotSpec], receivers, it: int, omegas,
                 d_obs: Optional[List[np.ndarray]] = None, segment: int = 0, levels: int = 1,
                 budget_bytes: Optional[float] = None):
        self.grid = grid
        self.it = int(it)
        self.segment = int(segment)
        self.levels = int(levels)
        self.budget_bytes = budget_bytes
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
        total, terms = np.zeros(sigma.size), []
        for sh, obs in zip(self.shots, self.d_obs):
            levels = self.levels if self.budget_bytes is None else \
                levels_for_budget(sh, self.budget_bytes)
            J, dj, nbytes = shot_value_and_grad(sh, obs, self.w, levels, self.segment)
            self.last_levels = levels
            self.peak_storage = max(self.peak_storage, nbytes)
            terms.append(J)
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
