"""CPU oracle for the fictitious-wave-domain FDTD hot path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker / the timed CPU
baseline.  The product package ``mittet_main_amd`` never imports it.

Two native pieces are wrapped here:

* ``liboracle.so`` -- the C restatement of the reference's C v1.4 forward path
  (``oracle/fwi3d_oracle.c``), each function citing the reference lines it
  follows.
* ``_ref/libref_cpml.so`` -- the reference's own ``fwi3d_cpu_cpmlabc_v1.4.c`` and
  ``fwi3d_cpu_lattice_v1.4.c`` compiled from the reference tree (only where it
  is mounted), driven through ``oracle/ref_harness.c``.  The reference's
  ``fwi3d_cpu_function_v1.4.c`` includes ``fftw3.h``, absent from this image,
  so the E/H update, source, receiver and transform functions are not built;
  they are pinned by the invariant and closed-form tests instead.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

# Exact FP64 values of the reference's constants (float literals widened).
C1 = float(np.float32(1.125))
C2 = float(np.float32(-0.04167))          # satan_i/fwi3d_cpu_function_v1.4.c:1187
MU0_C = 12.5663706e-7                     # satan_i/fwi3d_cpu_alloex1_v1.4.h:3
RCOEF = float(np.float32(0.01))           # satan_i/fwi3d_cpu_cpmlabc_v1.4.c:13
ALPHA_MAX = float(np.float32(0.2))        # satan_i/fwi3d_cpu_cpmlabc_v1.4.c:21
DT_FACTOR_REF = float(np.float32(0.999))  # satan_i/fwi3d_cpu_lattice_v1.4.c:39

ARRAY_IDS = {
    "EX": 0, "EY": 1, "EZ": 2, "HX": 3, "HY": 4, "HZ": 5, "SIG": 6,
    "CA_X": 7, "CA_Y": 8, "CA_Z": 9, "CB_X": 10, "CB_Y": 11, "CB_Z": 12,
    "DA_X": 13, "DA_Y": 14, "DA_Z": 15, "DB_X": 16, "DB_Y": 17, "DB_Z": 18,
    "PSI_EZX": 19, "PSI_EYX": 20, "PSI_EXY": 21, "PSI_EZY": 22, "PSI_EYZ": 23, "PSI_EXZ": 24,
    "PSI_HYX": 25, "PSI_HZX": 26, "PSI_HZY": 27, "PSI_HXY": 28, "PSI_HXZ": 29, "PSI_HYZ": 30,
}
PROFILE_IDS = {"ESIG": 0, "MSIG": 1, "EKAP": 2, "MKAP": 3, "AE": 4, "AM": 5,
               "BE": 6, "BH": 7, "CE": 8, "CH": 9, "KED": 10, "KHD": 11}
PSI_NAMES = [k for k in ARRAY_IDS if k.startswith("PSI_")]
FIELD_NAMES = ["EX", "EY", "EZ", "HX", "HY", "HZ"]


class OrParams(C.Structure):
    _fields_ = [("ix", C.c_int), ("iy", C.c_int), ("iz", C.c_int), ("ncpml", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dt", C.c_double), ("omega0", C.c_double), ("mu0", C.c_double),
                ("c1", C.c_double), ("c2", C.c_double), ("cmax", C.c_double),
                ("rcoef", C.c_double), ("nn", C.c_double), ("order", C.c_double),
                ("ma", C.c_double), ("kappa_max", C.c_double), ("alpha_max", C.c_double),
                ("opt_to_max", C.c_double * 3)]


def build():
    """Compile liboracle.so (and _ref/ where the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, dp, ip, i, d = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int, C.c_double
        L.or_lattice_time_3d.restype = d
        L.or_lattice_time_3d.argtypes = [d, d, d, d, d, d, d, dp]
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(OrParams)]
        L.or_destroy.argtypes = [vp]
        L.or_array.restype = dp
        L.or_array.argtypes = [vp, i]
        L.or_profile.restype = dp
        L.or_profile.argtypes = [vp, i, i]
        for nm in ("or_init_cpml", "or_media_coeff_3d", "or_set_zero_eh",
                   "or_epml_abcs4", "or_hpml_abcs4"):
            getattr(L, nm).argtypes = [vp]
        L.or_e_field_cpml42.argtypes = [vp, i]
        L.or_h_field_cpml42.argtypes = [vp, i]
        L.or_read_source_3d.argtypes = [vp, i, dp, i, i, i, i]
        L.or_output.argtypes = [vp, i, ip, i, i, dp, dp, dp]
        L.or_forward.argtypes = [vp, i, dp, i, i, i, i, i, i, ip, i, dp, dp, dp, i, dp, dp]
        L.or_firstderiv_gauss.argtypes = [i, d, d, dp]
        L.or_ricker.argtypes = [i, d, dp]
        L.or_dft_weights.argtypes = [i, d, d, i, dp, dp]
        L.or_fft_backward_real.argtypes = [i, dp, dp]
        L.or_laplace_to_freq.argtypes = [i, d, d, i, dp, dp, dp, dp, dp, dp, dp]
        L.or_num_threads.restype = i
        L.or_set_num_threads.argtypes = [i]
        L.or_stream_triad.restype = d
        L.or_stream_triad.argtypes = [C.c_long, i]
        for nm in ("orf_init_cpml", "orf_e_field_cpml4", "orf_cpml_e4", "orf_h_field_cpml4",
                   "orf_cpml_h4"):
            getattr(L, nm).argtypes = [vp]
        L.orf_media_coeff.argtypes = [vp, dp]
        L.orf_waveform.argtypes = [i, d, d, d, dp]
        L.orf_source_hz.argtypes = [vp, i, ip, dp, d]
        L.orf_forward.argtypes = [vp, i, ip, dp, dp, i, i, ip, i, dp]
        L.or_laplace_volume.argtypes = [i, i, i, dp, dp, dp, dp]
        L.or_residual_e.argtypes = [i, i, dp, dp, dp, dp]
        L.or_sensitivity.argtypes = [i, i, i, dp, dp, dp]
        L.or_conv_gdele.argtypes = [i, i, i, dp, dp, dp]
        L.or_calc_gamma.argtypes = [i, i, i, i, i, dp, dp, dp, dp]
        L.or_calc_phi.argtypes = [i, i, i, i, dp, dp, dp, dp]
        L.or_calc_alpha.argtypes = [i, i, dp, dp, dp, dp, dp]
        L.or_update_para2.argtypes = [i, i, i, i, i, dp, dp, d, d, d, dp, dp, dp]
        _LIB = L
    return _LIB


def num_threads() -> int:
    return lib().or_num_threads()


def set_num_threads(n: int) -> None:
    lib().or_set_num_threads(int(n))


def stream_triad(n: int = 40_000_000, reps: int = 5) -> float:
    """Best STREAM-triad GB/s of the host with the current thread count."""
    return lib().or_stream_triad(int(n), int(reps))


def lattice_time_3d(minval, omega0, dx, dy, dz, factor=DT_FACTOR_REF, mu0=MU0_C):
    """dt and cmax as satan_i/fwi3d_cpu_lattice_v1.4.c:35-39 compute them."""
    cm = C.c_double(0.0)
    dt = lib().or_lattice_time_3d(minval, omega0, mu0, dx, dy, dz, factor, C.byref(cm))
    return dt, cm.value


def firstderiv_gauss(nt, dt, fmax_w):
    out = np.zeros(nt)
    lib().or_firstderiv_gauss(nt, dt, fmax_w, _dp(out))
    return out


def fortran_waveform(nstep, dt, fmax, pi=3.14159265358979):
    """s(istep), istep = 1..nstep, of ver0040/cpu_waveform.f90:20-26 (Fortran pi)."""
    out = np.zeros(nstep)
    lib().orf_waveform(nstep, dt, fmax, pi, _dp(out))
    return out


def ricker(nt, dt):
    out = np.zeros(nt)
    lib().or_ricker(nt, dt, _dp(out))
    return out


def dft_weights(it, dt, omega0, omegas):
    om = np.ascontiguousarray(omegas, dtype=np.float64)
    w = np.zeros((it, len(om)), dtype=np.complex128)
    lib().or_dft_weights(it, dt, omega0, len(om), _dp(om), _dp(w.view(np.float64)))
    return w


def fft_backward_real(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x), dtype=np.complex128)
    lib().or_fft_backward_real(len(x), _dp(x), _dp(out.view(np.float64)))
    return out


def laplace_fft(trace, jsig, dt, omega0):
    """laplace_fft for one trace (satan_i/fwi3d_cpu_function_v1.4.c:1430-1564):
    E, J, G at every bin n < it (laplace_to_freq), then the FFTW_FORWARD
    back-transform scaled by 2/(it dt), as numpy's forward FFT."""
    it = len(trace)
    lap = laplace_to_freq(trace, jsig, np.zeros(it), dt, omega0, it)
    back = lambda x: np.fft.fft(x) * 2.0 / it / dt
    return {"E_t": back(lap["E"]), "J_t": back(lap["J"]), "G_t": back(lap["G"])}


def efld_fic(rel, t, cval, fmax_w, mu0=MU0_C):
    """E_x of a unit x-directed dipole in a homogeneous whole space, fictitious
    (wave) domain, at offset rel = (x, y, z) m and times t: the retarded dipole
    field of satan_i/fwi3d_cpu_function_v1.4.c:1776-1818 (fwi3d_cpu_convert_v1.4.c:
    207-240), E_x = mu0/(4 pi r) [(nx^2 - 1) p'' + (3 nx^2 - 1)(c/r p' + (c/r)^2 p)],
    p(t) = sqrt(beta/pi) exp(-beta (t - r/c - t0)^2), beta = pi fmax^2, t0 = pi/fmax."""
    x, y, z = (float(v) for v in rel)
    r = np.sqrt(x * x + y * y + z * z)
    beta, t0 = np.pi * fmax_w ** 2, np.pi / fmax_w
    u = np.asarray(t, dtype=np.float64) - r / cval - t0
    g = np.sqrt(beta / np.pi) * np.exp(-beta * u ** 2)
    g1 = -2 * beta * u * g
    g2 = g * (4 * beta ** 2 * u ** 2 - 2 * beta)
    nx2 = (x / r) ** 2
    return mu0 / (4 * np.pi * r) * ((nx2 - 1) * g2 + (3 * nx2 - 1) * (cval / r * g1 + (cval / r) ** 2 * g))


def efld_dip(rel, t, sig=1.0, mu0=MU0_C):
    """E_x(t) of a unit x-directed dipole in a whole space of conductivity sig,
    diffusive domain (satan_i/fwi3d_cpu_function_v1.4.c:1726-1774,
    fwi3d_cpu_convert_v1.4.c:164-201): with theta = sqrt(mu0 sig/(4 t)) and
    u = theta r, E_x = A nx^2 - B, A = [(4 u^3 + 6 u) e^{-u^2}/sqrt(pi) +
    3 erfc u]/(4 pi sig r^3), B = [(4 u^3 + 2 u) e^{-u^2}/sqrt(pi) +
    erfc u]/(4 pi sig r^3); E_x(0) = 0.  Returns (E_x, its forward
    difference dE_x/dt on the same samples, 0 at t[0])."""
    from math import erfc
    x, y, z = (float(v) for v in rel)
    r = np.sqrt(x * x + y * y + z * z)
    t = np.asarray(t, dtype=np.float64)
    ex = np.zeros(t.shape)
    for q, tq in enumerate(t):
        if tq <= 0:
            continue
        u = np.sqrt(mu0 * sig / 4.0 / tq) * r
        e = np.exp(-u * u)
        a = ((4 / np.sqrt(np.pi) * u ** 3 + 6 / np.sqrt(np.pi) * u) * e + 3 * erfc(u)) / (4 * np.pi * sig * r ** 3)
        b = ((4 / np.sqrt(np.pi) * u ** 3 + 2 / np.sqrt(np.pi) * u) * e + erfc(u)) / (4 * np.pi * sig * r ** 3)
        ex[q] = a * (x / r) ** 2 - b
    dex = np.zeros(t.shape)
    dex[1:] = np.diff(ex) / np.diff(t)
    return ex, dex


def laplace_to_freq(trace, jsig, strue, dt, omega0, nbins):
    """Receiver spectra at bins n < nbins (satan_i/fwi3d_cpu_function_v1.4.c:1976-2089)."""
    it = len(trace)
    tr = np.ascontiguousarray(trace, dtype=np.float64)
    js = np.ascontiguousarray(jsig, dtype=np.float64)
    st = np.ascontiguousarray(strue, dtype=np.float64)
    outs = [np.zeros(nbins, dtype=np.complex128) for _ in range(4)]
    with np.errstate(all="ignore"):
        lib().or_laplace_to_freq(it, dt, omega0, nbins, _dp(tr), _dp(js), _dp(st),
                                 *[_dp(o.view(np.float64)) for o in outs])
    return dict(zip(("E", "J", "G", "OBS"), outs))


class OracleSim:
    """One simulation state of the oracle (the reference's global state, owned)."""

    def __init__(self, ix, iy, iz, ncpml, dx, dy, dz, dt, omega0, cmax,
                 mu0=MU0_C, c1=C1, c2=C2, rcoef=RCOEF, nn=3.0, order=0.0, ma=1.0,
                 kappa_max=1.0, alpha_max=ALPHA_MAX, opt_to_max=(10.0, 1.0, 1.0)):
        self.shape = (iz, iy, ix)
        self.n = ix * iy * iz
        self.p = OrParams(ix, iy, iz, ncpml, dx, dy, dz, dt, omega0, mu0, c1, c2, cmax,
                          rcoef, nn, order, ma, kappa_max, alpha_max,
                          (C.c_double * 3)(*opt_to_max))
        self._s = lib().or_create(C.byref(self.p))
        if not self._s:
            raise MemoryError("oracle allocation failed")

    def __del__(self):
        if getattr(self, "_s", None):
            lib().or_destroy(self._s)
            self._s = None

    def arr(self, name):
        ptr = lib().or_array(self._s, ARRAY_IDS[name])
        return np.ctypeslib.as_array(ptr, shape=(self.n,)).reshape(self.shape)

    def prof(self, axis, name):
        ln = (self.p.ix, self.p.iy, self.p.iz)[axis]
        ptr = lib().or_profile(self._s, axis, PROFILE_IDS[name])
        return np.ctypeslib.as_array(ptr, shape=(ln,))

    def set_sigma(self, sig):
        self.arr("SIG")[...] = np.asarray(sig, dtype=np.float64).reshape(self.shape)

    def init_cpml(self):
        lib().or_init_cpml(self._s)

    def media_coeff_3d(self):
        lib().or_media_coeff_3d(self._s)

    def set_zero_eh(self):
        lib().or_set_zero_eh(self._s)

    # ---- Fortran lineage (ver0040), see fwi3d_oracle.h
    def f_init_cpml(self):
        lib().orf_init_cpml(self._s)

    def f_media_coeff(self, mu):
        self._mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(-1)
        lib().orf_media_coeff(self._s, _dp(self._mu))

    def f_source_hz(self, cells, sval):
        c = np.ascontiguousarray(cells, dtype=np.int32)
        lib().orf_source_hz(self._s, c.size, _ip(c), _dp(self._mu), float(sval))

    def f_e_field_cpml4(self):
        lib().orf_e_field_cpml4(self._s)

    def f_cpml_e4(self):
        lib().orf_cpml_e4(self._s)

    def f_h_field_cpml4(self):
        lib().orf_h_field_cpml4(self._s)

    def f_cpml_h4(self):
        lib().orf_cpml_h4(self._s)

    def f_forward(self, cells, signal, nsteps, rec_cells, it):
        """Fortran-lineage steps (source before E); traces (6, nrec, it) of E
        and H after each step.  cells / rec_cells: 0-based flat indices."""
        c = np.ascontiguousarray(cells, dtype=np.int32)
        r = np.ascontiguousarray(rec_cells, dtype=np.int32)
        sig = None if signal is None else np.ascontiguousarray(signal, dtype=np.float64)
        tr = np.zeros((6, max(r.size, 1), it))
        lib().orf_forward(self._s, c.size, _ip(c), _dp(self._mu),
                          _dp(sig) if sig is not None else None, nsteps, r.size,
                          _ip(r) if r.size else None, it, _dp(tr))
        return tr[:, :r.size]

    def e_field_cpml42(self, sign=1):
        lib().or_e_field_cpml42(self._s, sign)

    def h_field_cpml42(self, sign=1):
        lib().or_h_field_cpml42(self._s, sign)

    def epml_abcs4(self):
        lib().or_epml_abcs4(self._s)

    def hpml_abcs4(self):
        lib().or_hpml_abcs4(self._s)

    def read_source_3d(self, comp, signal, src1, step):
        sig = np.ascontiguousarray(signal, dtype=np.float64)
        lib().or_read_source_3d(self._s, comp, _dp(sig), int(src1[0]), int(src1[1]),
                                int(src1[2]), step)

    def forward(self, comp, signal, src1, nsteps, rec1=None, it=None, step0=0,
                weights=None, acc=None, traces=None):
        """Steps [step0, step0+nsteps) of the rec2rec loop.  Returns (traces, acc).

        rec1: (nrec,3) 1-based receiver cells; traces: (3,nrec,it) float64;
        weights: (it,nf) complex128; acc: (3,nf,N) complex128 accumulated in place.
        """
        it = it if it is not None else step0 + nsteps + 1
        sig = None if signal is None else np.ascontiguousarray(signal, dtype=np.float64)
        if rec1 is not None and len(rec1):
            r = np.ascontiguousarray(rec1, dtype=np.int32).reshape(-1, 3)
            nrec = r.shape[0]
            if traces is None:
                traces = np.zeros((3, nrec, it))
        else:
            r = np.zeros((1, 3), dtype=np.int32)
            nrec = 0
            traces = np.zeros((3, 1, 1)) if traces is None else traces
        nf = 0
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.complex128)
            nf = w.shape[1]
            if acc is None:
                acc = np.zeros((3, nf, self.n), dtype=np.complex128)
            wp, ap = _dp(w.view(np.float64)), _dp(acc.view(np.float64))
        else:
            wp = ap = None
        # steps n in [step0, step0 + nsteps) read signal[n + 1], write trace
        # sample n and read weight row n: refuse what does not fit
        end = step0 + nsteps
        if step0 < 0 or nsteps < 0:
            raise ValueError(f"step0 {step0} and nsteps {nsteps} must be >= 0")
        if sig is not None and end + 1 > sig.size:
            raise ValueError(f"steps up to {end - 1} need {end + 1} signal samples, got {sig.size}")
        if nrec and (traces.shape != (3, nrec, it) or end > it):
            raise ValueError(f"traces {traces.shape} cannot hold samples up to {end - 1} "
                             f"of (3, {nrec}, {it})")
        if nf and (end > w.shape[0] or acc.shape != (3, nf, self.n)):
            raise ValueError(f"weights of {w.shape[0]} steps / accumulators {acc.shape} do not fit "
                             f"steps up to {end - 1}")
        src = src1 if src1 is not None else (1, 1, 1)
        lib().or_forward(self._s, comp, _dp(sig) if sig is not None else None,
                         int(src[0]), int(src[1]), int(src[2]), step0, nsteps,
                         nrec, _ip(r), it, _dp(traces[0]), _dp(traces[1]), _dp(traces[2]),
                         nf, wp, ap)
        return traces, acc


class RefCpml:
    """The reference's own CPML/lattice code (oracle/_ref/libref_cpml.so).

    Available only where the reference tree was mounted at build time; the
    reference keeps process-global state, so one instance at a time.
    """

    @staticmethod
    def available():
        return os.path.exists(os.path.join(_HERE, "_ref", "libref_cpml.so"))

    def __init__(self, ix, iy, iz, ncpml, dx, dy, dz, omega0, minval, maxval):
        global _REF
        if _REF is None:
            L = C.CDLL(os.path.join(_HERE, "_ref", "libref_cpml.so"))
            d, i, dp = C.c_double, C.c_int, C.POINTER(C.c_double)
            L.ref_setup.argtypes = [i, i, i, i, d, d, d, d, d, d]
            L.ref_lattice.restype = d
            L.ref_lattice.argtypes = [dp]
            L.ref_set_dt.argtypes = [d]
            L.ref_set_cmax.argtypes = [d]
            L.ref_init_cpml.argtypes = [C.c_char_p]
            L.ref_put.argtypes = [i, dp]
            L.ref_get.argtypes = [i, dp]
            L.ref_get_profile.argtypes = [i, i, dp]
            _REF = L
        self.L = _REF
        self.shape = (iz, iy, ix)
        self.dims = (ix, iy, iz)
        if self.L.ref_setup(ix, iy, iz, ncpml, dx, dy, dz, omega0, minval, maxval) != 0:
            raise MemoryError("reference setup failed")

    def put(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.L.ref_put(ARRAY_IDS[name], _dp(a)) == 0

    def get(self, name):
        out = np.zeros(self.shape)
        assert self.L.ref_get(ARRAY_IDS[name], _dp(out)) == 0
        return out

    def lattice(self):
        cm = C.c_double(0.0)
        dt = self.L.ref_lattice(C.byref(cm))
        return dt, cm.value

    def set_dt(self, dt):
        self.L.ref_set_dt(dt)

    def set_cmax(self, cmax):
        self.L.ref_set_cmax(cmax)

    def init_cpml(self):
        with tempfile.TemporaryDirectory() as td:
            os.makedirs(os.path.join(td, "data"))
            assert self.L.ref_init_cpml(td.encode()) == 0

    def profile(self, axis, name):
        out = np.zeros(self.dims[axis])
        assert self.L.ref_get_profile(axis, PROFILE_IDS[name], _dp(out)) == 0
        return out

    def epml(self):
        self.L.ref_epml()

    def hpml(self):
        self.L.ref_hpml()


# ---- the reference's FWI gradient chain (compat mode) -------------------------
# Thin wrappers over the literal C loops; complex arrays are complex128.

def _cd(a):
    return _dp(np.ascontiguousarray(a, dtype=np.complex128).view(np.float64))


def laplace_volume(acc_bins, jw, tw, upperf, startf):
    """G = E/J (tw None) or E_ff = (E/J) T over the volume: [3, upperf, N], zero
    below startf (satan_i/fwi3d_cpu_function_v1.4.c:2166-2189, 1883-1894)."""
    acc = np.ascontiguousarray(acc_bins, dtype=np.complex128)
    n = acc.shape[-1]
    out = np.zeros((3, upperf, n), dtype=np.complex128)
    jwc = np.ascontiguousarray(jw, dtype=np.complex128)
    twc = None if tw is None else np.ascontiguousarray(tw, dtype=np.complex128)
    lib().or_laplace_volume(n, upperf, startf, _cd(acc), _cd(jwc),
                            None if twc is None else _cd(twc), _cd(out))
    return out


def residual_e(eobs, effrec):
    """dE = Eobs - E_ff at the receivers and sum |dE| (function.c:2202-2244)."""
    ob = np.ascontiguousarray(eobs, dtype=np.complex128)
    ef = np.ascontiguousarray(effrec, dtype=np.complex128)
    de = np.zeros_like(ob)
    err = np.zeros(1)
    lib().or_residual_e(ob.shape[1], ob.shape[2], _cd(ob), _cd(ef), _cd(de), _dp(err))
    return de, float(err[0])


def sensitivity(eff, ebb):
    """eff [3, upperf, N]; ebb [3, nrec, 3, upperf, N] (function.c:2246-2290)."""
    eff = np.ascontiguousarray(eff, dtype=np.complex128)
    ebb = np.ascontiguousarray(ebb, dtype=np.complex128)
    n, upperf, nrec = eff.shape[-1], eff.shape[1], ebb.shape[1]
    sense = np.zeros(n)
    lib().or_sensitivity(n, nrec, upperf, _cd(eff), _cd(ebb), _dp(sense))
    return sense


def conv_gdele(ebb, de):
    """E_back [3, upperf, N] (function.c:2292-2349)."""
    ebb = np.ascontiguousarray(ebb, dtype=np.complex128)
    de = np.ascontiguousarray(de, dtype=np.complex128)
    nrec, upperf, n = ebb.shape[1], ebb.shape[3], ebb.shape[4]
    out = np.zeros((3, upperf, n), dtype=np.complex128)
    lib().or_conv_gdele(n, nrec, upperf, _cd(ebb), _cd(de), _cd(out))
    return out


def calc_gamma(shape, nc, eff, eback, sense, grad=None):
    """grad += ... on the interior cells (function.c:2351-2384); shape (iz, iy, ix)."""
    iz, iy, ix = shape
    eff = np.ascontiguousarray(eff, dtype=np.complex128)
    g = np.zeros(ix * iy * iz) if grad is None else np.ascontiguousarray(grad, dtype=np.float64)
    lib().or_calc_gamma(ix, iy, iz, nc, eff.shape[1], _cd(eff), _cd(eback),
                        _dp(np.ascontiguousarray(sense)), _dp(g))
    return g


def calc_phi(shape, iseabed, sig, grad):
    """(sig2, scaler) (function.c:2565-2603)."""
    iz, iy, ix = shape
    sg = np.ascontiguousarray(sig, dtype=np.float64).ravel()
    gr = np.ascontiguousarray(grad, dtype=np.float64).ravel()
    sig2 = np.zeros_like(sg)
    sc = np.zeros(1)
    lib().or_calc_phi(ix, iy, iz, iseabed, _dp(sg), _dp(gr), _dp(sig2), _dp(sc))
    return sig2, float(sc[0])


def calc_alpha(de, eff2, effl, k1=0j, k2=0j):
    """(k1, k2) accumulated over one source (function.c:2757-2788)."""
    a = np.array([k1.real, k1.imag])
    b = np.array([k2.real, k2.imag])
    de = np.ascontiguousarray(de, dtype=np.complex128)
    lib().or_calc_alpha(de.shape[1], de.shape[2], _cd(de), _cd(eff2), _cd(effl), _dp(a), _dp(b))
    return complex(a[0], a[1]), complex(b[0], b[1])


def update_para2(shape, nc, iseabed, k1, k2, scaler, minval, maxval, sig, grad):
    """(alpha, new sig); grad is zeroed as the reference does (function.c:2790-2832)."""
    iz, iy, ix = shape
    sg = np.array(sig, dtype=np.float64).ravel()
    gr = np.array(grad, dtype=np.float64).ravel()
    a = np.array([k1.real, k1.imag])
    b = np.array([k2.real, k2.imag])
    al = np.zeros(1)
    lib().or_update_para2(ix, iy, iz, nc, iseabed, _dp(a), _dp(b), scaler, minval, maxval,
                          _dp(sg), _dp(gr), _dp(al))
    return float(al[0]), sg
