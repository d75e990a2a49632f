"""The reference's experiment directory (``data/``): readers and writers of its
text files, and the bridge that replays such a directory on the GPU.

Files, with the reference function that reads or writes each (satan_i/):

* ``model_env.dat``   -- grid, time and frequency set-up, transmitters and
  receivers (1-based).  Written by fwi3d_cpu_posiinf_v1.4.c:71-100, read by
  read_shotrec (fwi3d_cpu_function_v1.4.c:14-75).
* ``minmax.dat``      -- sigma bounds, ``%lf`` twice (sigmodel.c:70-71); they set
  cmax and hence dt (lattice.c:35-39) and clamp the model update.
* ``iseabed.dat``     -- one ``%12d`` (sigmodel.c:78).
* ``conductivity.dat``-- ``x y z sigma`` per cell, i fastest then j then k
  (sigmodel.c:57-68; read by true_model / init_model, function.c:268-323;
  init_model replaces sigma >= 4.0f by 0.45f).
* ``waveform{X,Y,Z}.dat``, ``true_waveformX.dat`` -- ``%e  %e`` (t, value)
  per step (waveform.c:154-170; read_trawave / read_waveform,
  function.c:78-117).
* ``spectrum_{e,j,g,c,ex,ey,ez}_SSS_RRR.dat`` -- per source SSS and receiver
  RRR, ``freq re im abs arg`` per bin (laplaceToFreq, function.c:2009-2083);
  ``e/j/g/c`` over all ``it`` bins for the x component, ``ex/ey/ez`` (the
  observed data read by read_Eobs, function.c:904-932) over bins < upperF.

Every writer uses the reference's printf format, so its files are what the
reference would write for the same values; the printed precision (``%e``: 7
significant digits, ``%lf``: 6 decimals) is the reference's own and is part of
what a replay reads back.  Readers take whitespace-separated tokens in the
reference's fscanf order and raise on a missing file, a short file or a token
that is not a number.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .fwi3d import Grid, from_1based, lattice_time_3d

# printf formats of the reference's writers
FMT_INT = "%d\n"                                  # posiinf.c:71-73, 78-80, 84, 90-92, 98
FMT_F = "%f\n"                                    # posiinf.c:74-77, 85-89
FMT_E = "%e\n"                                    # posiinf.c:81-83
FMT_POS = "%4d  %4d  %4d\n"                       # posiinf.c:94, 100
FMT_MINMAX = "%lf\n"                              # sigmodel.c:70-71
FMT_SEABED = "%12d\n"                             # sigmodel.c:78
FMT_SIGMA = "%12lf   %12lf   %12lf   %15.5e\n"    # sigmodel.c:66
FMT_WAVE = "%e  %e\n"                             # waveform.c:166-169
FMT_SPEC = "%10e %10e  %10e  %10e  %10e\n"        # function.c:2063-2075
FMT_SPEC_E = "%10e %10e  %10e  %10e   %10e\n"     # function.c:2061 (spectrum_e)

# init_model's replacement of conductors (function.c:308): sig >= 4.0f -> 0.45f
INIT_CUT = float(np.float32(4.0))
INIT_VALUE = float(np.float32(0.45))


class _Tokens:
    """Whitespace-separated tokens of a text file, consumed in fscanf order."""

    def __init__(self, path: str):
        self.path = path
        with open(path, "r") as f:
            self.tok = f.read().split()
        self.pos = 0

    def _next(self, what: str) -> str:
        if self.pos >= len(self.tok):
            raise ValueError(f"{self.path}: file ends before {what}")
        t = self.tok[self.pos]
        self.pos += 1
        return t

    def int(self, what: str) -> int:
        t = self._next(what)
        try:
            return int(t)
        except ValueError:
            raise ValueError(f"{self.path}: {what}: expected an integer, got {t!r}") from None

    def float(self, what: str) -> float:
        t = self._next(what)
        try:
            return float(t)
        except ValueError:
            raise ValueError(f"{self.path}: {what}: expected a number, got {t!r}") from None

    def floats(self, count: int, what: str) -> np.ndarray:
        if self.pos + count > len(self.tok):
            raise ValueError(f"{self.path}: {what}: need {count} numbers, "
                             f"{len(self.tok) - self.pos} left")
        try:
            out = np.array(self.tok[self.pos:self.pos + count], dtype=np.float64)
        except ValueError:
            raise ValueError(f"{self.path}: {what}: non-numeric token") from None
        self.pos += count
        return out


# ---------------------------------------------------------------- model_env.dat

@dataclass
class ModelEnv:
    """The fields of ``model_env.dat`` in file order (posiinf.c:71-100).
    ``shots`` and ``receivers`` are the file's 1-based (x, y, z) triplets;
    ``bandpass`` holds FL, FH, FS, AP, AS, which the solver ignores."""
    ix: int
    iy: int
    iz: int
    dx: float
    dy: float
    dz: float
    dt_ratio: float
    it: int
    mstep: int
    trstep: int
    omega0: float
    fmax_w: float
    glim: float
    ncpml: int
    bandpass: Tuple[float, float, float, float, float]
    start_f: int
    upper_f: int
    shots: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), dtype=np.int64))
    receivers: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), dtype=np.int64))

    def sources0(self) -> np.ndarray:
        """0-based transmitter cells (read_source_3d, function.c:520-524)."""
        return np.array([from_1based(s) for s in self.shots], dtype=np.int64).reshape(-1, 3)

    def receivers0(self) -> np.ndarray:
        """0-based receiver cells (residualE, function.c:2220-2224)."""
        return np.array([from_1based(r) for r in self.receivers], dtype=np.int64).reshape(-1, 3)


def read_model_env(path: str) -> ModelEnv:
    """read_shotrec's pass over model_env.dat (function.c:23-64)."""
    t = _Tokens(path)
    ix, iy, iz = t.int("ix"), t.int("iy"), t.int("iz")
    dx, dy, dz = t.float("dx"), t.float("dy"), t.float("dz")
    dt_ratio = t.float("dt_ratio")
    it, mstep, trstep = t.int("it"), t.int("mstep"), t.int("trstep")
    omega0, fmax_w, glim = t.float("omega_0"), t.float("fmax_w"), t.float("Glim")
    ncpml = t.int("ncpml")
    band = tuple(t.float(n) for n in ("FL", "FH", "FS", "AP", "AS"))
    start_f, upper_f = t.int("startF"), t.int("upperF")
    shots = [[t.int(f"shot {i} {a}") for a in "xyz"] for i in range(t.int("shot_num"))]
    recs = [[t.int(f"receiver {i} {a}") for a in "xyz"] for i in range(t.int("rec_num"))]
    return ModelEnv(ix, iy, iz, dx, dy, dz, dt_ratio, it, mstep, trstep, omega0, fmax_w, glim,
                    ncpml, band, start_f, upper_f,
                    np.array(shots, dtype=np.int64).reshape(-1, 3),
                    np.array(recs, dtype=np.int64).reshape(-1, 3))


def format_model_env(env: ModelEnv) -> str:
    """The text posiinf.c:71-100 writes for these values."""
    parts = [FMT_INT % v for v in (env.ix, env.iy, env.iz)]
    parts += [FMT_F % v for v in (env.dx, env.dy, env.dz, env.dt_ratio)]
    parts += [FMT_INT % v for v in (env.it, env.mstep, env.trstep)]
    parts += [FMT_E % v for v in (env.omega0, env.fmax_w, env.glim)]
    parts.append(FMT_INT % env.ncpml)
    parts += [FMT_F % v for v in env.bandpass]
    parts += [FMT_INT % v for v in (env.start_f, env.upper_f, len(env.shots))]
    parts += [FMT_POS % tuple(int(c) for c in s) for s in env.shots]
    parts.append(FMT_INT % len(env.receivers))
    parts += [FMT_POS % tuple(int(c) for c in r) for r in env.receivers]
    return "".join(parts)


def write_model_env(path: str, env: ModelEnv) -> None:
    with open(path, "w") as f:
        f.write(format_model_env(env))


# ---------------------------------------------------- minmax.dat, iseabed.dat

def read_minmax(path: str) -> Tuple[float, float]:
    """(minval, maxval) (function.c:66-67)."""
    t = _Tokens(path)
    return t.float("minval"), t.float("maxval")


def write_minmax(path: str, minval: float, maxval: float) -> None:
    with open(path, "w") as f:
        f.write(FMT_MINMAX % minval + FMT_MINMAX % maxval)


def read_iseabed(path: str) -> int:
    return _Tokens(path).int("iseabed")


def write_iseabed(path: str, iseabed: int) -> None:
    with open(path, "w") as f:
        f.write(FMT_SEABED % iseabed)


# ------------------------------------------------------------ conductivity.dat

def read_conductivity(path: str, shape: Sequence[int], initial: bool = False) -> np.ndarray:
    """sigma[iz, iy, ix] from the 4th column (true_model, function.c:268-287);
    ``initial`` applies init_model's replacement sig >= 4.0f -> 0.45f
    (function.c:296-310)."""
    iz, iy, ix = (int(v) for v in shape)
    n = ix * iy * iz
    vals = _Tokens(path).floats(4 * n, f"{n} cells of (x, y, z, sigma)")
    sig = vals.reshape(n, 4)[:, 3].copy()
    if initial:
        sig[sig >= INIT_CUT] = INIT_VALUE
    return sig.reshape(iz, iy, ix)


def format_conductivity(sigma: np.ndarray, dx: float, dy: float, dz: float) -> str:
    """sigmodel.c:57-68: one line per cell, i fastest."""
    iz, iy, ix = sigma.shape
    s = np.asarray(sigma, dtype=np.float64)
    return "".join(FMT_SIGMA % (i * dx, j * dy, k * dz, s[k, j, i])
                   for k in range(iz) for j in range(iy) for i in range(ix))


def write_conductivity(path: str, sigma: np.ndarray, dx: float, dy: float, dz: float) -> None:
    with open(path, "w") as f:
        f.write(format_conductivity(sigma, dx, dy, dz))


# -------------------------------------------------------------- waveform files

def read_waveform(path: str, it: int) -> np.ndarray:
    """The value column of the first ``it`` (t, value) pairs (function.c:91-96)."""
    vals = _Tokens(path).floats(2 * int(it), f"{it} (t, value) pairs")
    return vals.reshape(-1, 2)[:, 1].copy()


def format_waveform(signal: np.ndarray, dt: float) -> str:
    return "".join(FMT_WAVE % (i * dt, v) for i, v in enumerate(np.asarray(signal, np.float64)))


def write_waveform(path: str, signal: np.ndarray, dt: float) -> None:
    """waveform.c:165-170: t = i*dt."""
    with open(path, "w") as f:
        f.write(format_waveform(signal, dt))


# -------------------------------------------------------------------- spectra

SPECTRUM_KINDS = ("e", "j", "g", "c", "ex", "ey", "ez")


def spectrum_path(data_dir: str, kind: str, isource: int, irec: int) -> str:
    """./data/spectrum_<kind>_%03d_%03d.dat (function.c:2010-2016)."""
    if kind not in SPECTRUM_KINDS:
        raise ValueError(f"unknown spectrum kind {kind!r}")
    return os.path.join(data_dir, "spectrum_%s_%03d_%03d.dat" % (kind, isource, irec))


def format_spectrum(freqs: np.ndarray, values: np.ndarray, kind: str = "ex") -> str:
    """``freq re im |v| arg v`` per bin (function.c:2060-2076)."""
    fmt = FMT_SPEC_E if kind == "e" else FMT_SPEC
    v = np.asarray(values, dtype=np.complex128)
    return "".join(fmt % (f, z.real, z.imag, abs(z), np.angle(z))
                   for f, z in zip(np.asarray(freqs, np.float64), v))


def write_spectrum(path: str, freqs: np.ndarray, values: np.ndarray, kind: str = "ex") -> None:
    with open(path, "w") as f:
        f.write(format_spectrum(freqs, values, kind))


def read_spectrum(path: str, nbins: int) -> np.ndarray:
    """re + i im of the first ``nbins`` lines (function.c:920-927)."""
    vals = _Tokens(path).floats(5 * int(nbins), f"{nbins} spectrum lines").reshape(-1, 5)
    return vals[:, 1] + 1j * vals[:, 2]


def read_eobs(data_dir: str, isource: int, nrec: int, upper_f: int) -> np.ndarray:
    """read_Eobs (function.c:904-932): complex [3, nrec, upper_f], components
    x, y, z, bins 0 .. upper_f-1 -- the layout ``CompatFWI.iterate`` takes."""
    out = np.zeros((3, nrec, upper_f), dtype=np.complex128)
    for r in range(nrec):
        for c, kind in enumerate(("ex", "ey", "ez")):
            out[c, r] = read_spectrum(spectrum_path(data_dir, kind, isource, r), upper_f)
    return out


def bin_frequencies(it: int, dt: float, nbins: int) -> np.ndarray:
    """The frequency column n*om/2/pi, om = 2 pi/(it dt) (function.c:1990, 2061)."""
    om = 2.0 * np.pi / it / dt
    return np.arange(nbins) * om / 2.0 / np.pi


def write_laplace_spectra(data_dir: str, isource: int, lap: dict, it: int, dt: float,
                          upper_f: int) -> None:
    """The seven files laplaceToFreq writes per receiver (function.c:2009-2083)
    from a ``Shot.laplace_to_freq(it, ...)`` result: e/j/g/c for the x component
    over every bin present, ex/ey/ez (observed data) over bins < upper_f."""
    e, j, g, obs = lap["E"], lap["J"], lap["G"], lap["OBS"]
    nrec, nb = e.shape[1], e.shape[2]
    if nb < upper_f:
        raise ValueError(f"spectra hold {nb} bins, upper_f is {upper_f}")
    freqs = bin_frequencies(it, dt, nb)
    for r in range(nrec):
        for kind, vals in (("e", e[0, r]), ("j", j), ("g", g[0, r]), ("c", obs[0, r])):
            write_spectrum(spectrum_path(data_dir, kind, isource, r), freqs, vals, kind)
        for c, kind in enumerate(("ex", "ey", "ez")):
            write_spectrum(spectrum_path(data_dir, kind, isource, r), freqs[:upper_f],
                           obs[c, r, :upper_f], kind)


# ------------------------------------------------------------------ experiment

@dataclass
class Experiment:
    """A reference experiment directory, loaded.  ``sigma`` is the model the
    reference program that reads this directory would start from: the true
    model (rec2rec, true_model) or the initial model (inv_fwi, init_model)."""
    data_dir: str
    env: ModelEnv
    minval: float
    maxval: float
    iseabed: int
    sigma: np.ndarray
    signal_x: np.ndarray
    signal_y: np.ndarray
    signal_z: np.ndarray
    signal_true: np.ndarray
    dt: float
    cmax: float

    @property
    def grid(self) -> Grid:
        e = self.env
        return Grid(e.ix, e.iy, e.iz, e.ncpml, e.dx, e.dy, e.dz, self.dt, e.omega0, self.cmax)

    @property
    def sources(self) -> np.ndarray:
        return self.env.sources0()

    @property
    def receivers(self) -> np.ndarray:
        return self.env.receivers0()

    def read_eobs(self, isource: int) -> np.ndarray:
        return read_eobs(self.data_dir, isource, len(self.env.receivers), self.env.upper_f)

    def compat_fwi(self):
        """The reference's FWI iteration (fwi3d_cpu_inv_fwi_v1.4.c:27-109) over
        this directory's set-up: z-directed sources, waveformX for the source
        and Green's runs, true_waveformX for T."""
        from .fwi_compat import CompatFWI
        e = self.env
        return CompatFWI(self.grid, self.sources, self.receivers, e.it, self.signal_x,
                         self.signal_true, e.start_f, e.upper_f, src_comp=2,
                         minval=self.minval, maxval=self.maxval, iseabed=self.iseabed)


def load_experiment(data_dir: str, model: str = "initial") -> Experiment:
    """What init() (function.c:2552-2563, inversion: ``model="initial"``) or
    rec2rec's set-up (fwi3d_cpu_rec2rec_v1.4.c:22-29, ``model="true"``) read,
    with dt and cmax from lattice_time_3d at minval (lattice.c:35-39)."""
    if model not in ("initial", "true"):
        raise ValueError("model must be 'initial' or 'true'")
    p = lambda name: os.path.join(data_dir, name)
    env = read_model_env(p("model_env.dat"))
    minval, maxval = read_minmax(p("minmax.dat"))
    iseabed = read_iseabed(p("iseabed.dat"))
    sigma = read_conductivity(p("conductivity.dat"), (env.iz, env.iy, env.ix),
                              initial=(model == "initial"))
    waves = [read_waveform(p(f"waveform{a}.dat"), env.it) for a in "XYZ"]
    strue = read_waveform(p("true_waveformX.dat"), env.it)
    dt, cmax = lattice_time_3d(minval, env.omega0, env.dx, env.dy, env.dz)
    return Experiment(data_dir, env, minval, maxval, iseabed, sigma, *waves, strue, dt, cmax)


def write_experiment(data_dir: str, env: ModelEnv, sigma: np.ndarray, minval: float,
                     maxval: float, iseabed: int, signal_x: np.ndarray,
                     signal_true: np.ndarray, dt: float,
                     signal_y: Optional[np.ndarray] = None,
                     signal_z: Optional[np.ndarray] = None) -> None:
    """Write the input files of a reference experiment directory (what
    posiinf, sigmodel and waveform produce); missing Y/Z waveforms are zero."""
    os.makedirs(data_dir, exist_ok=True)
    p = lambda name: os.path.join(data_dir, name)
    write_model_env(p("model_env.dat"), env)
    write_minmax(p("minmax.dat"), minval, maxval)
    write_iseabed(p("iseabed.dat"), iseabed)
    write_conductivity(p("conductivity.dat"), sigma, env.dx, env.dy, env.dz)
    zero = np.zeros(len(signal_x))
    for a, s in zip("XYZ", (signal_x, zero if signal_y is None else signal_y,
                            zero if signal_z is None else signal_z)):
        write_waveform(p(f"waveform{a}.dat"), s, dt)
    write_waveform(p("true_waveformX.dat"), signal_true, dt)


def record_observed(exp: Experiment, sources: Optional[Iterable[int]] = None) -> List[dict]:
    """rec2rec on the GPU (fwi3d_cpu_rec2rec_v1.4.c:38-60): each source runs on
    ``exp.sigma`` with an Ez dipole driven by waveformX, and laplaceToFreq
    writes its spectra into the experiment directory.  Returns the spectra."""
    from .fwi3d import Model
    e = exp.env
    srcs = exp.sources
    todo = range(len(srcs)) if sources is None else list(sources)
    out = []
    model = Model(exp.grid)
    try:
        model.media_coeff_3d(exp.sigma)
        for s in todo:
            sh = model.shot(e.it, exp.signal_x, tuple(int(v) for v in srcs[s]), 2, exp.receivers)
            try:
                sh.run()
                sh.sync()
                lap = sh.laplace_to_freq(e.it, exp.signal_true)
            finally:
                sh.close()
            write_laplace_spectra(exp.data_dir, s, lap, e.it, exp.dt, e.upper_f)
            out.append(lap)
    finally:
        model.close()
    return out
