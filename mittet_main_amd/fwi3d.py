"""Python host layer over libmittet_fdtd.so, mirroring the reference's
operator surface (satan_i/fwi3d_cpu_function_v1.4.h, fwi3d_cpu_cpmlabc_v1.4.h,
fwi3d_cpu_lattice_v1.4.h) with the reference's globals held in objects:

    model = Model(grid)                  # read_shotrec + init_cpml (cpmlabc.c:7-175)
    model.media_coeff_3d(sigma)          # function.c:325-435
    shot = model.shot(it, signal, src, receivers, ...)
    shot.set_zero_eh()                   # function.c:964-986
    for step in range(it - 1):           # rec2rec.c:48-57
        shot.e_field_cpml42(); shot.epml_abcs4(); shot.read_source_3d(step)
        shot.h_field_cpml42(); shot.hpml_abcs4(); shot.output(step)
    # or, fused and graph-captured:  shot.run()

Coordinates are 0-based (i, j, k); ``from_1based`` converts the reference's
config coordinates.  All numbers are float64; complex outputs are complex128.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import GridParams, ShotParams, check

# FP64 values of the reference's constants (float literals widened to double)
C1 = float(np.float32(1.125))             # satan_i/fwi3d_cpu_function_v1.4.c:1187
C2 = float(np.float32(-0.04167))          # same line: (double)-0.04167f
MU0 = 12.5663706e-7                       # satan_i/fwi3d_cpu_alloex1_v1.4.h:3
RCOEF = float(np.float32(0.01))           # satan_i/fwi3d_cpu_cpmlabc_v1.4.c:13
ALPHA_MAX = float(np.float32(0.2))        # satan_i/fwi3d_cpu_cpmlabc_v1.4.c:21
DT_FACTOR_REF = float(np.float32(0.999))  # satan_i/fwi3d_cpu_lattice_v1.4.c:39

# The Fortran lineage (model variant 1): constants of ver0040/cpu_parameter_wafe.f90
SRC_HZ = 5                                # magnetic source on Hz (FWI3D_SRC_HZ)
F_C2 = -0.04167                           # :37, a double literal there
F_MU0 = 1.2566370614e-6                   # :86
F_PI = 3.14159265358979                   # :32, the pi of every Fortran formula
F_RCOEF = 0.01                            # :131
F_ALPHA_MAX = 0.2                         # :127

EX, EY, EZ, HX, HY, HZ = range(6)
PROFILES = ("ESIG", "MSIG", "EKAP", "MKAP", "AE", "AM", "BE", "BH", "CE", "CH", "KED", "KHD")


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def from_1based(ijk1) -> tuple:
    """Reference config coordinates (jx, jy, jz), 1-based -> 0-based (function.c:524)."""
    return tuple(int(v) - 1 for v in ijk1)


def lattice_time_3d(minval: float, omega0: float, dx: float, dy: float, dz: float,
                    factor: float = DT_FACTOR_REF, mu0: float = MU0):
    """dt, cmax of satan_i/fwi3d_cpu_lattice_v1.4.c:35-39 (dt = courant*6/7*factor)."""
    cm = C.c_double(0.0)
    dt = _lib.load().fwi3d_lattice_time_3d(minval, omega0, mu0, dx, dy, dz, factor, C.byref(cm))
    return dt, cm.value


def firstderiv_gauss(nt: int, dt: float, fmax_w: float) -> np.ndarray:
    """Gaussian-derivative source of satan_i/fwi3d_cpu_waveform_v1.4.c:132-143:
    s(t) = -2*beta*(t-t0)*sqrt(beta/pi)*exp(-beta*(t-t0)^2), beta = pi*fmax^2, t0 = pi/fmax."""
    beta = np.pi * fmax_w ** 2
    t0 = np.pi / fmax_w
    t = np.arange(nt) * dt
    return -2.0 * beta * (t - t0) * np.sqrt(beta / np.pi) * np.exp(-beta * (t - t0) ** 2)


@dataclass
class Grid:
    ix: int
    iy: int
    iz: int
    ncpml: int
    dx: float
    dy: float
    dz: float
    dt: float
    omega0: float
    cmax: float
    mu0: float = MU0
    c1: float = C1
    c2: float = C2
    rcoef: float = RCOEF
    nn: float = 3.0
    order: float = 0.0
    ma: float = 1.0
    kappa_max: float = 1.0
    alpha_max: float = ALPHA_MAX
    opt_to_max: tuple = (10.0, 1.0, 1.0)

    @property
    def shape(self):
        return (self.iz, self.iy, self.ix)

    @property
    def ncells(self) -> int:
        return self.ix * self.iy * self.iz

    def to_c(self) -> GridParams:
        return GridParams(self.ix, self.iy, self.iz, self.ncpml, self.dx, self.dy, self.dz,
                          self.dt, self.omega0, self.mu0, self.c1, self.c2, self.cmax,
                          self.rcoef, self.nn, self.order, self.ma, self.kappa_max,
                          self.alpha_max, (C.c_double * 3)(*self.opt_to_max))


def fortran_grid(ix: int, iy: int, iz: int, ncpml: int, dx: float, dy: float, dz: float,
                 dt: float, omega0: float, cmax: float) -> "Grid":
    """Grid with the Fortran lineage's constants: c2 = -0.04167, mu0 =
    1.2566370614e-6, Rcoef 0.01, alpha_max 0.2, no sigma_max factor on x
    (ver0040/cpu_parameter_wafe.f90:35-37, 86, 127-131; cpu_cpml4.f90:21)."""
    return Grid(ix, iy, iz, ncpml, dx, dy, dz, dt, omega0, cmax, mu0=F_MU0, c1=1.125, c2=F_C2,
                rcoef=F_RCOEF, alpha_max=F_ALPHA_MAX, opt_to_max=(1.0, 1.0, 1.0))


def fortran_dt(dx: float, cmax: float) -> float:
    """dt = 0.999 dx/cmax/sqrt(3)/1.16667 (ver0040/cpu_parameter_wafe.f90:118)."""
    return 0.999 * dx / cmax / np.sqrt(3.0) / 1.16667


def fortran_waveform(nstep: int, dt: float, fmax: float) -> np.ndarray:
    """The Fortran source s(istep), istep = 1..nstep, as written in
    ver0040/cpu_waveform.f90:29: -(2 beta (t - t0) sqrt(beta/pi)) exp(-beta (t - t0)**2),
    t = istep dt, t0 = pi/fmax, beta = pi fmax**2 with the Fortran pi.  Element n
    drives step n of a variant-1 shot."""
    t0 = F_PI / fmax
    beta = F_PI * fmax ** 2.0
    t = np.arange(1, nstep + 1) * dt
    return -(2.0 * beta * (t - t0) * np.sqrt(beta / F_PI)) * np.exp(-beta * (t - t0) ** 2.0)


def time_reversed_residual(observed: np.ndarray, computed: np.ndarray) -> np.ndarray:
    """Drive signal of the residual back-injection (read_source_3d_res,
    cw_19/ver0030/cpu_waveform.f90:62-85): the receiver residual
    observed - computed, reversed in time.  A variant-1 shot with this signal
    at the receiver's Hz points is that injection."""
    o = np.asarray(observed, dtype=np.float64)
    c = np.asarray(computed, dtype=np.float64)
    if o.shape != c.shape:
        raise ValueError("observed and computed traces differ in length")
    return np.ascontiguousarray((o - c)[::-1])


class Model:
    """Grid + conductivity + CPML profiles resident in HBM.  variant 1 selects
    the Fortran lineage's operators (fwi3d_model_create_variant)."""

    def __init__(self, grid: Grid, stream: Optional[int] = None, variant: int = 0):
        self.grid = grid
        self.variant = int(variant)
        self._p = grid.to_c()
        h = C.c_void_p()
        check(_lib.load().fwi3d_model_create_variant(C.byref(self._p), self.variant, stream,
                                                     C.byref(h)))
        self._h = h
        self._shots = []

    def set_mu(self, mu: np.ndarray) -> None:
        """Per-cell permeability (variant 1), [iz, iy, ix] H/m."""
        m = np.ascontiguousarray(mu, dtype=np.float64).reshape(-1)
        if m.size != self.grid.ncells:
            raise ValueError(f"mu has {m.size} cells, grid has {self.grid.ncells}")
        check(_lib.load().fwi3d_model_set_mu(self._h, _dp(m)))

    def close(self):
        if getattr(self, "_h", None):
            for s in list(self._shots):
                s.close()
            check(_lib.load().fwi3d_model_destroy(self._h))
            self._h = None

    __del__ = close

    def profile(self, axis: int, name: str) -> np.ndarray:
        n = (self.grid.ix, self.grid.iy, self.grid.iz)[axis]
        out = np.zeros(n)
        check(_lib.load().fwi3d_model_get_profile(self._h, axis, PROFILES.index(name), _dp(out)))
        return out

    def media_coeff_3d(self, sigma: np.ndarray) -> None:
        sig = np.ascontiguousarray(sigma, dtype=np.float64).reshape(-1)
        if sig.size != self.grid.ncells:
            raise ValueError(f"sigma has {sig.size} cells, grid has {self.grid.ncells}")
        check(_lib.load().fwi3d_media_coeff_3d(self._h, _dp(sig)))

    def shot(self, it: int, signal: Optional[np.ndarray] = None, src=(0, 0, 0), src_comp: int = 0,
             receivers: Optional[Sequence] = None, omega_vol: Optional[Sequence[float]] = None,
             ring: int = 0, zchunk: int = 0) -> "Shot":
        s = Shot(self, it, signal, src, src_comp, receivers, omega_vol, ring, zchunk)
        self._shots.append(s)
        return s


def shot_params(it, signal, src, src_comp, receivers, omega_vol, ring, keep, zchunk=0):
    sig = None if signal is None else np.ascontiguousarray(signal, dtype=np.float64)
    if sig is not None and sig.size < it:
        raise ValueError(f"signal has {sig.size} samples, it={it}")
    rec = (np.zeros((0, 3), np.int32) if receivers is None
           else np.ascontiguousarray(np.asarray(receivers, dtype=np.int32).reshape(-1, 3)))
    om = (np.zeros(0) if omega_vol is None
          else np.ascontiguousarray(np.asarray(omega_vol, dtype=np.float64)))
    keep.extend([sig, rec, om])
    return ShotParams(int(it), int(src_comp), int(src[0]), int(src[1]), int(src[2]),
                      None if sig is None else sig.ctypes.data_as(C.POINTER(C.c_double)),
                      rec.shape[0], rec.ctypes.data_as(C.POINTER(C.c_int)),
                      om.size, om.ctypes.data_as(C.POINTER(C.c_double)), int(ring), int(zchunk))


def shot_bytes(grid: Grid, it: int, nrec: int = 0, nf_vol: int = 0, ring: int = 0,
               volume_dft_h: bool = False) -> int:
    """Device bytes of one shot (fwi3d_shot_bytes), plus the H ring and H
    accumulators of fwi3d_enable_volume_dft_h when volume_dft_h is set."""
    keep = []
    sp = shot_params(it, None, (0, 0, 0), 0, np.zeros((nrec, 3)), np.zeros(nf_vol), ring, keep)
    gp = grid.to_c()
    L = _lib.load()
    b = int(L.fwi3d_shot_bytes(C.byref(gp), C.byref(sp)))
    if volume_dft_h:
        b += int(L.fwi3d_volume_dft_h_bytes(C.byref(gp), C.byref(sp)))
    return b


class Shot:
    """Fields, psi shells, receivers and DFT accumulators of one source."""

    def __init__(self, model: Model, it, signal, src, src_comp, receivers, omega_vol, ring,
                 zchunk=0):
        self.model = model
        self.it = int(it)
        keep = []
        sp = shot_params(it, signal, src, src_comp, receivers, omega_vol, ring, keep, zchunk)
        self.nrec = sp.nrec
        self.nf = sp.nf_vol
        self.omega_vol = np.array(keep[2])
        self.signal = None if keep[0] is None else np.array(keep[0][:it])
        h = C.c_void_p()
        check(_lib.load().fwi3d_shot_create(model._h, C.byref(sp), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            check(_lib.load().fwi3d_shot_destroy(self._h))
            self._h = None
            if self in self.model._shots:
                self.model._shots.remove(self)

    __del__ = close

    # ---- per-step operators, reference names
    def add_source(self, i: int, j: int, k: int) -> None:
        """Another Hz source point with this shot's waveform (variant 1)."""
        check(_lib.load().fwi3d_shot_add_source(self._h, int(i), int(j), int(k)))

    def set_zero_eh(self):
        check(_lib.load().fwi3d_set_zero_eh(self._h))

    def e_field_cpml42(self, sign: int = 1):
        check(_lib.load().fwi3d_e_field_cpml42(self._h, sign))

    def epml_abcs4(self):
        check(_lib.load().fwi3d_epml_abcs4(self._h))

    def read_source_3d(self, step: int):
        check(_lib.load().fwi3d_read_source_3d(self._h, step))

    def h_field_cpml42(self, sign: int = 1):
        check(_lib.load().fwi3d_h_field_cpml42(self._h, sign))

    def hpml_abcs4(self):
        check(_lib.load().fwi3d_hpml_abcs4(self._h))

    def output(self, step: int):
        check(_lib.load().fwi3d_output(self._h, step))

    # ---- fused
    def step_n(self, nsteps: int):
        check(_lib.load().fwi3d_step_n(self._h, nsteps))

    def run(self):
        check(_lib.load().fwi3d_run(self._h))

    def sync(self):
        check(_lib.load().fwi3d_shot_sync(self._h))

    def profile_steps(self, nsteps: int) -> dict:
        """Mean kernel durations (ms) over nsteps event-bracketed fused steps."""
        ms = np.zeros(9)
        check(_lib.load().fwi3d_profile_steps(self._h, nsteps, _dp(ms)))
        return dict(zip(("efield", "hfield", "receivers", "flush", "step", "hstore", "hflush",
                         "ehfield", "fused"), ms.tolist()))

    @property
    def step(self) -> int:
        return _lib.load().fwi3d_shot_step(self._h)

    @property
    def diverged(self) -> bool:
        """A receiver sample or the source cell's field turned NaN or infinite."""
        rc = _lib.load().fwi3d_shot_diverged(self._h)
        check(min(rc, 0))
        return rc == 1

    # ---- readback
    def field(self, which: int) -> np.ndarray:
        out = np.zeros(self.model.grid.shape)
        check(_lib.load().fwi3d_get_field(self._h, which, _dp(out)))
        return out

    def set_field(self, which: int, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(self.model.grid.shape)
        check(_lib.load().fwi3d_set_field(self._h, which, _dp(v)))

    def traces(self) -> np.ndarray:
        """(6, nrec, it): Ex, Ey, Ez, Hx, Hy, Hz at the receivers, sample n after step n."""
        out = np.zeros((6, self.nrec, self.it))
        check(_lib.load().fwi3d_get_traces(self._h, _dp(out)))
        return out

    def volume_dft(self) -> np.ndarray:
        """(3, nf_vol, iz, iy, ix) complex: sum_n E[n] dt exp((i-1) sqrt(omega0 omega) n dt)."""
        g = self.model.grid
        out = np.zeros((3, self.nf, g.ncells), dtype=np.complex128)
        check(_lib.load().fwi3d_get_volume_dft(self._h, _dp(out)))
        return out.reshape(3, self.nf, *g.shape)

    def enable_volume_dft_h(self) -> None:
        """Also accumulate H over the volume at the shot's volume-DFT
        frequencies (sum_n H[n] w_n, H after step n); before the first step.
        Allocates an H ring of the shot's ring slots (ring * 24 B per cell)."""
        check(_lib.load().fwi3d_enable_volume_dft_h(self._h))

    def volume_dft_h(self) -> np.ndarray:
        """(3, nf_vol, iz, iy, ix) complex: sum_n H[n] dt exp((i-1) sqrt(omega0 omega) n dt)."""
        g = self.model.grid
        out = np.zeros((3, self.nf, g.ncells), dtype=np.complex128)
        check(_lib.load().fwi3d_get_volume_dft_h(self._h, _dp(out)))
        return out.reshape(3, self.nf, *g.shape)

    def receiver_dft(self, omegas: Sequence[float], h_factor: bool = True) -> np.ndarray:
        """(6, nrec, nfreq) complex receiver spectra (f_to_d_e / f_to_d_h)."""
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        out = np.zeros((6, self.nrec, om.size), dtype=np.complex128)
        check(_lib.load().fwi3d_receiver_dft(self._h, om.size, _dp(om), int(h_factor), _dp(out)))
        return out

    def laplace_to_freq(self, nbins: int, strue: np.ndarray) -> dict:
        """satan_i/fwi3d_cpu_function_v1.4.c:1976-2089 at bins n < nbins: E, J, G=E/J, OBS."""
        st = np.ascontiguousarray(strue, dtype=np.float64)
        if st.size < self.it:
            raise ValueError("true waveform shorter than it")
        e = np.zeros((3, self.nrec, nbins), dtype=np.complex128)
        j = np.zeros(nbins, dtype=np.complex128)
        g = np.zeros_like(e)
        o = np.zeros_like(e)
        check(_lib.load().fwi3d_laplace_to_freq(self._h, nbins, _dp(st), _dp(e), _dp(j), _dp(g),
                                                _dp(o)))
        return {"E": e, "J": j, "G": g, "OBS": o}

    def laplace_fft(self) -> dict:
        """satan_i/fwi3d_cpu_function_v1.4.c:1430-1564: E, J, G at every bin,
        back-transformed to diffusive time.  E_t, G_t complex (3, nrec, it);
        J_t complex (it,)."""
        e = np.zeros((3, self.nrec, self.it), dtype=np.complex128)
        g = np.zeros_like(e)
        j = np.zeros(self.it, dtype=np.complex128)
        check(_lib.load().fwi3d_laplace_fft(self._h, _dp(e), _dp(j), _dp(g)))
        return {"E_t": e, "J_t": j, "G_t": g}


def dft_rows(x: np.ndarray, sign: int, scale: float = 1.0) -> np.ndarray:
    """scale * sum_n x[..., n] exp(sign 2 pi i k n / nd) over the last axis, on
    the GPU (fwi3d_dft_rows).  sign -1 is FFTW_FORWARD, +1 FFTW_BACKWARD."""
    a = np.ascontiguousarray(x, dtype=np.complex128)
    nd = a.shape[-1]
    out = np.zeros_like(a)
    check(_lib.load().fwi3d_dft_rows(_dp(a), a.size // nd, nd, int(sign), float(scale), _dp(out)))
    return out


def to_diffusive_time(spec: np.ndarray, dt: float, sign: int = -1) -> np.ndarray:
    """Diffusive-time signal from spectra at the bins n 2 pi/(nd dt), n < nd:
    x_t[k] = 2/(nd dt) sum_n X[n] exp(sign 2 pi i k n / nd).  sign -1 is the C
    code's FFTW_FORWARD (satan_i/fwi3d_cpu_function_v1.4.c:1507-1527), +1 the
    Fortran f_to_d_e's FFTW_BACKWARD (ver0040/cpu_f_to_d_e.f90:281-320)."""
    nd = np.shape(spec)[-1]
    return dft_rows(spec, sign, 2.0 / nd / dt)


def convolve_gj(g_t: np.ndarray, j_t: np.ndarray) -> np.ndarray:
    """convolution_GJ_to_E (satan_i/fwi3d_cpu_function_v1.4.c:1566-1635):
    FFTW_BACKWARD(G_t) * FFTW_BACKWARD(J_t) per row; g_t complex (..., nd),
    j_t real (nd,)."""
    g = np.ascontiguousarray(g_t, dtype=np.complex128)
    j = np.ascontiguousarray(j_t, dtype=np.float64)
    nd = g.shape[-1]
    if j.shape != (nd,):
        raise ValueError(f"j_t must have {nd} samples")
    out = np.zeros_like(g)
    check(_lib.load().fwi3d_convolve_gj(_dp(g), g.size // nd, nd, _dp(j), _dp(out)))
    return out


def source_spectrum(signal: np.ndarray, dt: float, omega0: float, omegas) -> np.ndarray:
    """J(omega) = csqrt(-2 omega0/(i omega)) sum J[n] w_n, J(0) := 2 omega0
    (satan_i/fwi3d_cpu_function_v1.4.c:2044-2050)."""
    sig = np.ascontiguousarray(signal, dtype=np.float64)
    om = np.ascontiguousarray(omegas, dtype=np.float64)
    out = np.zeros(om.size, dtype=np.complex128)
    check(_lib.load().fwi3d_source_spectrum(dt, omega0, sig.size, _dp(sig), om.size, _dp(om),
                                            _dp(out)))
    return out


def f_to_d_h(spectra: np.ndarray, omegas, omega0: float) -> np.ndarray:
    """H(omega) = sqrt(-2 omega0/(i omega)) * H'(omega) (ver0040/cpu_f_to_d_h.f90:146-148);
    spectra (..., nfreq) complex."""
    om = np.ascontiguousarray(omegas, dtype=np.float64)
    sp = np.ascontiguousarray(spectra, dtype=np.complex128).copy()
    count = sp.size // om.size
    check(_lib.load().fwi3d_f_to_d_h(omega0, om.size, _dp(om), count, _dp(sp)))
    return sp


class Batch:
    """Persistent multi-shot launches: every member shot advances in the same
    kernel launches, whatever its grid size."""

    def __init__(self, max_shots: int, chunk: int = 0, stream: Optional[int] = None):
        h = C.c_void_p()
        check(_lib.load().fwi3d_batch_create(max_shots, chunk, stream, C.byref(h)))
        self._h = h
        self.members = []

    def close(self):
        if getattr(self, "_h", None):
            check(_lib.load().fwi3d_batch_destroy(self._h))
            self._h = None
            self.members = []

    __del__ = close

    def add(self, shot: Shot):
        check(_lib.load().fwi3d_batch_add(self._h, shot._h))
        self.members.append(shot)

    def remove(self, shot: Shot):
        check(_lib.load().fwi3d_batch_remove(self._h, shot._h))
        self.members.remove(shot)

    def step(self, nsteps: int):
        check(_lib.load().fwi3d_batch_step(self._h, nsteps))

    def sync(self):
        check(_lib.load().fwi3d_batch_sync(self._h))

    def __len__(self):
        return _lib.load().fwi3d_batch_size(self._h)
