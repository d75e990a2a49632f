"""The reference's experiment-directory files (model_env.dat, minmax.dat,
iseabed.dat, conductivity.dat, waveform*.dat, spectrum_*.dat) read and
written by mittet_main_amd.refio.

Pinned by: the model_env.dat that the reference's own writer
(satan_i/fwi3d_cpu_posiinf_v1.4.c, compiled into oracle/_ref/posiinf) produced,
kept as tests/golden/model_env_posiinf.dat; the C library's printf and strtod
applied to the reference's own format strings (the writers of conductivity,
waveform and spectrum files include fftw3.h, which this image lacks, so those
programs are not built); and layouts built so that a swapped axis, component,
receiver or source index lands a value in the wrong place."""
import ctypes as C
import ctypes.util
import os
import subprocess

import numpy as np
import pytest

from mittet_main_amd import refio
from mittet_main_amd.fwi3d import lattice_time_3d

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "model_env_posiinf.dat")
POSIINF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "posiinf")

_libc = C.CDLL(ctypes.util.find_library("c"))
_libc.strtod.restype = C.c_double
_libc.strtod.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]


def c_printf(fmt: str, *args) -> str:
    buf = C.create_string_buffer(8192)
    cargs = [C.c_int(int(a)) if isinstance(a, (int, np.integer)) else C.c_double(float(a))
             for a in args]
    _libc.snprintf(buf, len(buf), fmt.encode(), *cargs)
    return buf.value.decode()


def c_strtod(tok: str) -> float:
    return _libc.strtod(tok.encode(), None)


def c_reread(fmt: str, *args) -> list:
    """What the reference reads back from a value it printed with fmt."""
    return [c_strtod(t) for t in c_printf(fmt, *args).split()]


def hashed(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal(n)


# ------------------------------------------------------------------ model_env

POSIINF_VALUES = dict(ix=31, iy=31, iz=31, dx=30.0, dy=30.0, dz=30.0, dt_ratio=100.0, it=1302,
                      mstep=4, trstep=1302 // 4, ncpml=10, start_f=1, upper_f=20)


def test_model_env_reads_the_reference_writers_file():
    env = refio.read_model_env(FIXTURE)
    for k, v in POSIINF_VALUES.items():
        assert getattr(env, k) == v, k
    assert env.omega0 == 6.283185            # 2 pi through "%e"
    assert env.fmax_w == 3.0
    assert env.glim == 10.4                  # 10.4f through "%e"
    assert env.bandpass == (4.0, 8.0, 12.0, 0.5, 5.0)
    i = np.arange(7)
    np.testing.assert_array_equal(env.shots, np.stack([10 + 2 * i, 16 + 0 * i, 16 + 0 * i], 1))
    np.testing.assert_array_equal(env.receivers, np.stack([11 + 2 * i, 16 + 0 * i, 16 + 0 * i], 1))
    np.testing.assert_array_equal(env.sources0(), env.shots - 1)
    np.testing.assert_array_equal(env.receivers0(), env.receivers - 1)


def test_model_env_writer_reproduces_the_reference_file(tmp_path):
    env = refio.read_model_env(FIXTURE)
    with open(FIXTURE) as f:
        want = f.read()
    assert refio.format_model_env(env) == want
    # written from the unquantized values the reference program holds
    env.omega0 = 2.0 * np.pi
    env.glim = float(np.float32(10.4))
    p = tmp_path / "model_env.dat"
    refio.write_model_env(str(p), env)
    assert p.read_text() == want


@pytest.mark.skipif(not os.path.exists(POSIINF), reason="oracle/_ref/posiinf not built here")
def test_model_env_fixture_is_what_the_reference_writes(tmp_path):
    (tmp_path / "data").mkdir()
    subprocess.run([POSIINF], cwd=tmp_path, check=True, stdout=subprocess.DEVNULL)
    with open(FIXTURE) as f:
        assert (tmp_path / "data" / "model_env.dat").read_text() == f.read()


def test_model_env_roundtrip_with_other_values(tmp_path):
    env = refio.ModelEnv(40, 36, 33, 12.5, 25.0, 7.75, 0.5, 901, 3, 300, 12.566370614359172,
                         2.5, 9.0, 6, (1.0, 2.0, 3.0, 0.25, 4.0), 2, 11,
                         np.array([[7, 8, 9], [30, 31, 32]]), np.array([[5, 6, 7]]))
    p = str(tmp_path / "m.dat")
    refio.write_model_env(p, env)
    back = refio.read_model_env(p)
    assert (back.ix, back.iy, back.iz, back.ncpml, back.it) == (40, 36, 33, 6, 901)
    assert (back.dx, back.dy, back.dz) == (12.5, 25.0, 7.75)
    assert back.omega0 == c_reread("%e", env.omega0)[0]
    np.testing.assert_array_equal(back.shots, env.shots)
    np.testing.assert_array_equal(back.receivers, env.receivers)
    assert refio.format_model_env(back) == refio.format_model_env(env)


# ---------------------------------------------------------------- formats

FORMATS = [
    (refio.FMT_F, 1), (refio.FMT_E, 1), (refio.FMT_MINMAX, 1), (refio.FMT_SIGMA, 4),
    (refio.FMT_WAVE, 2), (refio.FMT_SPEC, 5), (refio.FMT_SPEC_E, 5),
]


@pytest.mark.parametrize("fmt,nargs", FORMATS)
def test_writer_formats_match_c_printf(fmt, nargs):
    rng = np.random.default_rng(3)
    vals = list(rng.standard_normal(200) * 10.0 ** rng.integers(-12, 9, 200))
    vals += [0.0, -0.0, 1.0, -1.0, 0.5, 2.5, 1e-300, 123456.5, 9.9999995e-5, 0.1234565]
    for i in range(0, len(vals) - nargs + 1, nargs):
        args = tuple(float(v) for v in vals[i:i + nargs])
        assert fmt % args == c_printf(fmt, *args), (fmt, args)


@pytest.mark.parametrize("fmt", [refio.FMT_INT, refio.FMT_POS, refio.FMT_SEABED])
def test_integer_formats_match_c_printf(fmt):
    n = fmt.count("%")
    for v in (0, 1, 7, 16, 99, 1302, 12345, -3):
        assert fmt % ((v,) * n) == c_printf(fmt, *((v,) * n))


def test_reader_parses_like_strtod():
    rng = np.random.default_rng(4)
    toks = []
    for v in rng.standard_normal(300) * 10.0 ** rng.integers(-20, 20, 300):
        toks += [refio.FMT_SPEC.split()[0] % v, "%.17g" % v, "%f" % v]
    for t in toks:
        assert float(t) == c_strtod(t), t


# ------------------------------------------------------------- conductivity

def test_conductivity_layout_quantization_and_initial_model(tmp_path):
    iz, iy, ix = 3, 4, 5
    sig = np.exp(np.clip(hashed(ix * iy * iz, 5), -3, 1.3)).reshape(iz, iy, ix)
    sig[2, 1, 3] = 4.0            # replaced by init_model (>= 4.0f)
    sig[0, 3, 0] = 3.99999        # kept
    sig[1, 0, 4] = 10.0           # replaced
    p = str(tmp_path / "conductivity.dat")
    refio.write_conductivity(p, sig, 12.5, 25.0, 40.0)
    lines = open(p).read().splitlines()
    assert len(lines) == ix * iy * iz
    # line of cell (i, j, k) = (3, 1, 2): coordinates and value in the C format
    q = (2 * iy + 1) * ix + 3
    assert lines[q] + "\n" == c_printf(refio.FMT_SIGMA, 3 * 12.5, 1 * 25.0, 2 * 40.0, 4.0)
    true = refio.read_conductivity(p, (iz, iy, ix))
    want = np.array([c_reread("%15.5e", v)[0] for v in sig.ravel()]).reshape(sig.shape)
    np.testing.assert_array_equal(true, want)
    init = refio.read_conductivity(p, (iz, iy, ix), initial=True)
    assert init[2, 1, 3] == init[1, 0, 4] == np.float32(0.45)
    assert init[0, 3, 0] == true[0, 3, 0]
    mask = np.ones(sig.shape, bool)
    mask[2, 1, 3] = mask[1, 0, 4] = False
    np.testing.assert_array_equal(init[mask], true[mask])
    with pytest.raises(ValueError):
        refio.read_conductivity(p, (iz + 1, iy, ix))


# ------------------------------------------------------------ waveforms

def test_waveform_roundtrip_and_short_file(tmp_path):
    sig = hashed(50, 6) * 1e-3
    dt = 0.0123456789
    p = str(tmp_path / "waveformX.dat")
    refio.write_waveform(p, sig, dt)
    lines = open(p).read().splitlines()
    assert lines[7] + "\n" == c_printf(refio.FMT_WAVE, 7 * dt, sig[7])
    got = refio.read_waveform(p, 50)
    np.testing.assert_array_equal(got, [c_reread("%e", v)[0] for v in sig])
    np.testing.assert_array_equal(refio.read_waveform(p, 20), got[:20])
    with pytest.raises(ValueError):
        refio.read_waveform(p, 51)


# -------------------------------------------------------------- spectra

def test_eobs_reads_each_source_receiver_component(tmp_path):
    nsrc, nrec, upper = 2, 3, 6
    d = str(tmp_path)
    truth = {}
    for s in range(nsrc):
        for r in range(nrec):
            for c, kind in enumerate(("ex", "ey", "ez")):
                # value encodes (source, receiver, component, bin)
                b = np.arange(upper)
                v = (1000 * s + 100 * r + 10 * c + b) + 1j * (b + 0.5)
                truth[s, r, c] = v
                refio.write_spectrum(refio.spectrum_path(d, kind, s, r),
                                     refio.bin_frequencies(80, 0.01, upper), v, kind)
    assert os.path.basename(refio.spectrum_path(d, "ey", 1, 2)) == "spectrum_ey_001_002.dat"
    for s in range(nsrc):
        e = refio.read_eobs(d, s, nrec, upper)
        assert e.shape == (3, nrec, upper)
        for r in range(nrec):
            for c in range(3):
                np.testing.assert_array_equal(e[c, r], truth[s, r, c])
    # columns 4 and 5 are |v| and arg v, as cabs/carg print them
    line = open(refio.spectrum_path(d, "ex", 1, 2)).read().splitlines()[3]
    v = truth[1, 2, 0][3]
    f3 = refio.bin_frequencies(80, 0.01, upper)[3]
    assert line + "\n" == c_printf(refio.FMT_SPEC, f3, v.real, v.imag, abs(v), np.angle(v))
    with pytest.raises(ValueError):
        refio.read_eobs(d, 0, nrec, upper + 1)
    with pytest.raises(FileNotFoundError):
        refio.read_eobs(d, 2, nrec, upper)


def test_laplace_spectra_file_set(tmp_path):
    it, nrec, upper, dt = 12, 2, 5, 0.02
    rng = np.random.default_rng(8)
    cz = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    lap = {"E": cz(3, nrec, it), "J": cz(it), "G": cz(3, nrec, it), "OBS": cz(3, nrec, it)}
    d = str(tmp_path)
    refio.write_laplace_spectra(d, 4, lap, it, dt, upper)
    for r in range(nrec):
        for kind, want in (("e", lap["E"][0, r]), ("j", lap["J"]), ("g", lap["G"][0, r]),
                           ("c", lap["OBS"][0, r])):
            txt = open(refio.spectrum_path(d, kind, 4, r)).read().splitlines()
            assert len(txt) == it
            got = refio.read_spectrum(refio.spectrum_path(d, kind, 4, r), it)
            np.testing.assert_allclose(got, want, rtol=1e-6)
        e = refio.read_eobs(d, 4, nrec, upper)
        for c in range(3):
            assert len(open(refio.spectrum_path(d, ("ex", "ey", "ez")[c], 4, r)).read()
                       .splitlines()) == upper
            np.testing.assert_allclose(e[c, r], lap["OBS"][c, r, :upper], rtol=1e-6)
    # frequency column: n om/2/pi = n/(it dt)
    first = [float(l.split()[0]) for l in open(refio.spectrum_path(d, "e", 4, 0))]
    np.testing.assert_allclose(first, np.arange(it) / (it * dt), rtol=1e-6)
    # spectrum_e keeps its own spacing (three blanks before the last column)
    v = lap["E"][0, 0, 2]
    line = open(refio.spectrum_path(d, "e", 4, 0)).read().splitlines()[2]
    f2 = refio.bin_frequencies(it, dt, it)[2]
    assert line + "\n" == c_printf(refio.FMT_SPEC_E, f2, v.real, v.imag, abs(v), np.angle(v))


# ------------------------------------------------------------ experiment

def small_env():
    return refio.ModelEnv(22, 20, 18, 50.0, 45.0, 40.0, 100.0, 90, 4, 22, 2 * np.pi, 3.0, 10.4, 4,
                          (4.0, 8.0, 12.0, 0.5, 5.0), 1, 4,
                          np.array([[12, 11, 7], [9, 13, 8]]), np.array([[8, 8, 6], [15, 13, 6]]))


def test_experiment_roundtrip(tmp_path):
    env = small_env()
    shape = (env.iz, env.iy, env.ix)
    sig = np.exp(hashed(np.prod(shape), 9)).reshape(shape)
    sig[5, 5, 5] = 4.5
    dtw = 0.001
    sx, st = hashed(env.it, 10), hashed(env.it, 11)
    d = str(tmp_path / "data")
    refio.write_experiment(d, env, sig, 0.05, 5.0, 9, sx, st, dtw)
    for name in ("model_env.dat", "minmax.dat", "iseabed.dat", "conductivity.dat",
                 "waveformX.dat", "waveformY.dat", "waveformZ.dat", "true_waveformX.dat"):
        assert os.path.exists(os.path.join(d, name)), name
    ex = refio.load_experiment(d, model="true")
    assert (ex.minval, ex.maxval, ex.iseabed) == (0.05, 5.0, 9)
    assert ex.sigma[5, 5, 5] == 4.5
    np.testing.assert_array_equal(ex.signal_x, [c_reread("%e", v)[0] for v in sx])
    np.testing.assert_array_equal(ex.signal_true, [c_reread("%e", v)[0] for v in st])
    assert not np.any(ex.signal_y) and not np.any(ex.signal_z)
    dt, cmax = lattice_time_3d(0.05, ex.env.omega0, 50.0, 45.0, 40.0)
    g = ex.grid
    assert (g.dt, g.cmax, g.omega0) == (dt, cmax, 6.283185)
    assert (g.ix, g.iy, g.iz, g.ncpml) == (22, 20, 18, 4)
    np.testing.assert_array_equal(ex.sources, env.shots - 1)
    np.testing.assert_array_equal(ex.receivers, env.receivers - 1)
    init = refio.load_experiment(d)
    assert init.sigma[5, 5, 5] == np.float32(0.45)
    with pytest.raises(ValueError):
        refio.load_experiment(d, model="both")
    os.remove(os.path.join(d, "iseabed.dat"))
    with pytest.raises(FileNotFoundError):
        refio.load_experiment(d)


def test_reader_rejects_malformed_model_env(tmp_path):
    p = tmp_path / "model_env.dat"
    text = open(FIXTURE).read()
    p.write_text(text.replace("1302", "13o2"))
    with pytest.raises(ValueError, match="it"):
        refio.read_model_env(str(p))
    p.write_text("\n".join(text.splitlines()[:-2]))
    with pytest.raises(ValueError, match="receiver"):
        refio.read_model_env(str(p))
