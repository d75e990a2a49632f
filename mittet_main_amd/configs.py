"""Seeded synthetic workloads of the forward/adjoint benchmarks.

Conductivity model (all sizes): three horizontal layers (interfaces at 1/3 and
2/3 of the interior depth, z = k increasing downward) with log-uniform sigma in
[0.01, 3.3] S/m, plus four resistive boxes (sigma = 0.01 S/m) whose sides are
drawn in [4, 12] cells at 64^3 and scale with the interior size.  The CPML
shells continue the edge values.  Fictitious permittivity eps' = sigma/(2 omega0)
with omega0 = 2 pi rad/s, mu = mu0.  dt = 0.99 x the 4th-order CFL limit
(courant*6/7, satan_i/fwi3d_cpu_lattice_v1.4.c:38-39) at sigma_min = 0.01 S/m.

The x-directed electric dipole sits at the horizontal centre, 2 cells below the
top layer interface, driven by the Gaussian-derivative waveform (fmax 3 Hz,
satan_i/fwi3d_cpu_waveform_v1.4.c:132-143).  Receivers sit on the top
interface: a line along x (64^3: 32 receivers) or a square surface grid.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from .fwi3d import Grid, firstderiv_gauss

OMEGA0 = 2.0 * np.pi
SIG_MIN, SIG_MAX = 0.01, 3.3
FMAX_W = 3.0
DT_FACTOR = 0.99


def dt_for(dx: float, dy: float, dz: float, sig_min: float = SIG_MIN, omega0: float = OMEGA0,
           mu0: float = 12.5663706e-7, factor: float = DT_FACTOR):
    """cmax = sqrt(2 omega0/(mu0 sigma_min)); dt = courant*6/7*factor."""
    cmax = np.sqrt(2.0 * omega0 / mu0 / sig_min)
    courant = 1.0 / cmax / np.sqrt(1.0 / dx ** 2 + 1.0 / dy ** 2 + 1.0 / dz ** 2)
    return courant * 6.0 / 7.0 * factor, cmax


def layered_box_sigma(n_int: int, ncpml: int, seed: int) -> Tuple[np.ndarray, int]:
    """(sigma[iz,iy,ix], k_top) for an n_int^3 interior padded by ncpml cells."""
    rng = np.random.default_rng(seed)
    n = n_int + 2 * ncpml
    layers = np.exp(rng.uniform(np.log(SIG_MIN), np.log(SIG_MAX), 3))
    k1 = ncpml + n_int // 3
    k2 = ncpml + (2 * n_int) // 3
    sig = np.empty((n, n, n))
    sig[:k1] = layers[0]
    sig[k1:k2] = layers[1]
    sig[k2:] = layers[2]
    scale = n_int / 64.0
    lo, hi = max(1, int(round(4 * scale))), max(2, int(round(12 * scale)))
    for _ in range(4):
        side = rng.integers(lo, hi + 1, size=3)
        org = [int(rng.integers(ncpml, ncpml + n_int - s + 1)) for s in side]
        sig[org[2]:org[2] + side[2], org[1]:org[1] + side[1], org[0]:org[0] + side[0]] = SIG_MIN
    return sig, k1


def sigma_digest(sig: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(sig, dtype=np.float64).tobytes()).hexdigest()[:16]


@dataclass
class Workload:
    name: str
    grid: Grid
    sigma: np.ndarray
    it: int
    src: Tuple[int, int, int]
    src_comp: int
    receivers: np.ndarray          # (nrec, 3) 0-based
    freqs_hz: List[float]
    signal: np.ndarray
    vol_dft: bool = True

    @property
    def omegas(self) -> np.ndarray:
        return 2.0 * np.pi * np.asarray(self.freqs_hz)

    @property
    def steps(self) -> int:
        return self.it - 1


def make(name: str, n_int: int, ncpml: int, it: int, freqs_hz, nrec_side: int = 0,
         nrec_line: int = 0, seed: int = 0, dx: float = 50.0, vol_dft: bool = True) -> Workload:
    sig, k_top = layered_box_sigma(n_int, ncpml, seed)
    n = n_int + 2 * ncpml
    dt, cmax = dt_for(dx, dx, dx)
    grid = Grid(n, n, n, ncpml, dx, dx, dx, dt, OMEGA0, cmax)
    src = (n // 2, n // 2, k_top + 2)
    margin = 2
    lo, hi = ncpml + margin, ncpml + n_int - 1 - margin
    if nrec_line:
        xs = np.linspace(lo, hi, nrec_line).round().astype(int)
        rec = np.stack([xs, np.full_like(xs, n // 2), np.full_like(xs, k_top)], axis=1)
    else:
        g = np.linspace(lo, hi, nrec_side).round().astype(int)
        X, Y = np.meshgrid(g, g, indexing="xy")
        rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, k_top)], axis=1)
    signal = firstderiv_gauss(it, dt, FMAX_W)
    return Workload(name, grid, sig, it, src, 0, rec.astype(np.int32), list(freqs_hz), signal,
                    vol_dft)


def plain_64(seed: int = 1) -> Workload:
    """64^3 interior + 10-cell CPML, Nt=2000, DFT at 0.25/0.5/1/2 Hz, 32 receivers on x."""
    return make("forward_64cube_fp64", 64, 10, 2000, [0.25, 0.5, 1.0, 2.0], nrec_line=32, seed=seed)


def roofline_256(seed: int = 2, it: int = 8000) -> Workload:
    """256^3 interior + 16-cell CPML, Nt=8000, 8 log-spaced DFT frequencies 0.1-4 Hz,
    full-volume E DFT, 256 receivers on a 16x16 surface grid."""
    freqs = list(np.geomspace(0.1, 4.0, 8))
    return make("forward_256cube_fp64", 256, 16, it, freqs, nrec_side=16, seed=seed)


def scaled(n_int: int, ncpml: int, it: int, nf: int = 4, seed: int = 3, nrec_side: int = 4,
           vol_dft: bool = True) -> Workload:
    freqs = list(np.geomspace(0.25, 2.0, nf))
    return make(f"forward_{n_int}cube_fp64", n_int, ncpml, it, freqs, nrec_side=nrec_side,
                seed=seed, vol_dft=vol_dft)


def smoothed_background(n_int: int, ncpml: int, seed: int, half_width: int = 0) -> np.ndarray:
    """Starting model of the gradient workload: the three layers of
    layered_box_sigma (same seed, so the same layer values) without the boxes,
    log(sigma) smoothed along z by a moving average over 2*half_width+1 cells
    (default half_width = n_int // 16)."""
    rng = np.random.default_rng(seed)
    n = n_int + 2 * ncpml
    layers = np.exp(rng.uniform(np.log(SIG_MIN), np.log(SIG_MAX), 3))
    k1 = ncpml + n_int // 3
    k2 = ncpml + (2 * n_int) // 3
    col = np.empty(n)
    col[:k1], col[k1:k2], col[k2:] = np.log(layers[0]), np.log(layers[1]), np.log(layers[2])
    hw = half_width or max(1, n_int // 16)
    pad = np.concatenate([np.full(hw, col[0]), col, np.full(hw, col[-1])])
    smooth = np.convolve(pad, np.full(2 * hw + 1, 1.0 / (2 * hw + 1)), mode="valid")
    return np.broadcast_to(np.exp(smooth)[:, None, None], (n, n, n)).copy()


@dataclass
class GradientWorkload:
    name: str
    grid: Grid
    sigma_true: np.ndarray
    sigma_start: np.ndarray
    it: int
    sources: np.ndarray            # (nshots, 3) 0-based x-dipole cells
    receivers: np.ndarray          # (nrec, 3)
    freqs_hz: List[float]
    signal: np.ndarray

    @property
    def omegas(self) -> np.ndarray:
        return 2.0 * np.pi * np.asarray(self.freqs_hz)


def gradient_128(seed: int = 4, it: int = 0, nshots: int = 8) -> GradientWorkload:
    """128^3 interior + 12-cell CPML; true model = layered_box_sigma, start =
    smoothed_background (same seed); 8 x-dipoles 2 cells below the top
    interface at seeded positions; 64 receivers on an 8x8 grid on the
    interface; L2 misfit at 4 frequencies (0.25-2 Hz); Nt = 30 x 128."""
    n_int, nc = 128, 12
    it = it or 30 * n_int
    sig, k_top = layered_box_sigma(n_int, nc, seed)
    sig0 = smoothed_background(n_int, nc, seed)
    n = n_int + 2 * nc
    dt, cmax = dt_for(50.0, 50.0, 50.0)
    grid = Grid(n, n, n, nc, 50.0, 50.0, 50.0, dt, OMEGA0, cmax)
    rng = np.random.default_rng(seed + 1000)
    xy = rng.integers(nc + 8, n - nc - 8, size=(nshots, 2))
    src = np.concatenate([xy, np.full((nshots, 1), k_top + 2)], axis=1).astype(np.int32)
    g = np.linspace(nc + 4, n - nc - 5, 8).round().astype(int)
    X, Y = np.meshgrid(g, g, indexing="xy")
    rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, k_top)], axis=1).astype(np.int32)
    return GradientWorkload("gradient_128cube_fp64", grid, sig, sig0, it, src, rec,
                            list(np.geomspace(0.25, 2.0, 4)), firstderiv_gauss(it, dt, FMAX_W))


def fwi_384(seed: int = 5, it: int = 1000, nshots: int = 2) -> GradientWorkload:
    """The served-FWI grid: 384^3 interior + 16-cell CPML (416^3 cells); true
    model = layered_box_sigma, start = smoothed_background (same seed);
    x-dipoles 2 cells below the top interface at seeded positions; 64
    receivers on an 8x8 grid on the interface; 8 frequencies log-spaced in
    0.1-4 Hz.  `it` is the time-sample count (bounded: a full-length run at
    this size takes minutes per shot)."""
    n_int, nc = 384, 16
    sig, k_top = layered_box_sigma(n_int, nc, seed)
    sig0 = smoothed_background(n_int, nc, seed)
    n = n_int + 2 * nc
    dt, cmax = dt_for(50.0, 50.0, 50.0)
    grid = Grid(n, n, n, nc, 50.0, 50.0, 50.0, dt, OMEGA0, cmax)
    rng = np.random.default_rng(seed + 1000)
    xy = rng.integers(nc + 8, n - nc - 8, size=(nshots, 2))
    src = np.concatenate([xy, np.full((nshots, 1), k_top + 2)], axis=1).astype(np.int32)
    g = np.linspace(nc + 4, n - nc - 5, 8).round().astype(int)
    X, Y = np.meshgrid(g, g, indexing="xy")
    rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, k_top)], axis=1).astype(np.int32)
    return GradientWorkload("fwi_384cube_fp64", grid, sig, sig0, it, src, rec,
                            list(np.geomspace(0.1, 4.0, 8)), firstderiv_gauss(it, dt, FMAX_W))
