"""Diffusive-time back-transform and G*J convolution on the GPU.

Parity: fwi3d_dft_rows against numpy's FFT (FFTW_FORWARD = numpy fft,
FFTW_BACKWARD = nd x numpy ifft); fwi3d_laplace_fft against the oracle's
laplace_to_freq at every bin followed by numpy's FFT (laplace_fft,
satan_i/fwi3d_cpu_function_v1.4.c:1430-1564); fwi3d_convolve_gj against the
product of numpy transforms (function.c:1566-1635).  All floating point, at
1e-12 (transforms of given data) and 1e-9 relative L2 (the whole laplace_fft
chain, whose spectra the device evaluates with its own exp/sincos).

Physics regression (the reference's efld_fic / efld_dip comparisons,
function.c:1726-1818): an x-dipole in a sigma = 1 S/m whole space, receiver
1500 m along x.  The fictitious-domain trace follows the retarded dipole field
(25 % at 30 cells, lattice dispersion), and the diffusive impulse response G_t
from laplace_fft follows dE_x/dt of the diffusive step response: peak within
3 samples, amplitude above the late-time floor (the transform's constant
aliasing term) within 10 %, shape within 8 %."""
import numpy as np
import pytest

import oracle
from mittet_main_amd import Grid, Model
from mittet_main_amd.fwi3d import (convolve_gj, dft_rows, firstderiv_gauss, lattice_time_3d,
                                   to_diffusive_time)

pytestmark = pytest.mark.gpu

OM0 = 2 * np.pi


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


@pytest.mark.parametrize("nd", [1, 7, 256, 1000])
def test_dft_rows_match_numpy_fft(nd):
    rng = np.random.default_rng(nd)
    x = rng.standard_normal((5, nd)) + 1j * rng.standard_normal((5, nd))
    assert rel(dft_rows(x, -1), np.fft.fft(x, axis=-1)) <= 1e-12
    assert rel(dft_rows(x, +1), nd * np.fft.ifft(x, axis=-1)) <= 1e-12
    assert rel(dft_rows(x, -1, 0.25), 0.25 * np.fft.fft(x, axis=-1)) <= 1e-12
    dt = 0.0123
    assert rel(to_diffusive_time(x, dt), np.fft.fft(x, axis=-1) * 2 / nd / dt) <= 1e-12
    assert rel(to_diffusive_time(x, dt, +1), np.fft.ifft(x, axis=-1) * 2 / dt) <= 1e-12


def test_convolve_gj_is_the_product_of_backward_transforms():
    rng = np.random.default_rng(3)
    nd = 300
    g = rng.standard_normal((4, nd)) + 1j * rng.standard_normal((4, nd))
    j = rng.standard_normal(nd)
    want = (nd * np.fft.ifft(g, axis=-1)) * (nd * np.fft.ifft(j))
    assert rel(convolve_gj(g, j), want) <= 1e-12
    with pytest.raises(ValueError):
        convolve_gj(g, j[:-1])


def test_laplace_fft_matches_oracle():
    ix, iy, iz, nc, it = 26, 24, 22, 5, 160
    rng = np.random.default_rng(8)
    sig = np.exp(rng.uniform(np.log(0.05), np.log(2.0), (iz, iy, ix)))
    dt, cmax = lattice_time_3d(float(sig.min()), OM0, 50.0, 45.0, 40.0)
    m = Model(Grid(ix, iy, iz, nc, 50.0, 45.0, 40.0, dt, OM0, cmax))
    m.media_coeff_3d(sig)
    signal = firstderiv_gauss(it, dt, 4.0)
    rec = np.array([[8, 9, 10], [15, 12, 7], [13, 13, 11]])
    sh = m.shot(it, signal, (13, 12, 11), 0, rec)
    sh.run()
    tr = sh.traces()
    got = sh.laplace_fft()
    for c in range(3):
        for r in range(len(rec)):
            want = oracle.laplace_fft(tr[c, r], signal, dt, OM0)
            for key in ("E_t", "G_t"):
                assert rel(got[key][c, r], want[key]) <= 1e-9, (key, c, r)
    assert rel(got["J_t"], want["J_t"]) <= 1e-9
    sh.close()
    m.close()


def test_whole_space_dipole_in_both_domains():
    n, nc, h, off, it = 100, 10, 50.0, 30, 400
    fmax = 3.0
    dt, cmax = lattice_time_3d(1.0, OM0, h, h, h)
    m = Model(Grid(n, n, n, nc, h, h, h, dt, OM0, cmax))
    m.media_coeff_3d(np.ones((n, n, n)))
    signal = firstderiv_gauss(it, dt, fmax)
    c0 = n // 2
    print(f"whole-space dipole: {n}^3, offset {off} cells, it {it}", flush=True)
    sh = m.shot(it, signal, (c0, c0, c0), 0, np.array([[c0 + off, c0, c0]]))
    sh.run()
    ex = sh.traces()[0, 0]
    r = (off * h, 0.0, 0.0)
    # fictitious domain: sample k holds E after k + 1 updates
    fic = oracle.efld_fic(r, (np.arange(it) + 1) * dt, cmax, fmax)
    assert rel(ex[:it - 1], fic[:it - 1]) < 0.25
    # diffusive domain: impulse response vs d/dt of the step response
    g = sh.laplace_fft()["G_t"][0, 0].real
    _, dex = oracle.efld_dip(r, np.arange(it) * dt)
    floor = g[int(0.9 * it):].mean()
    kg, kd = int(np.argmax(g)), int(np.argmax(dex))
    assert abs(kg - kd) <= 3
    assert 0.9 < (g[kg] - floor) / dex[kd] < 1.1
    w = slice(kd // 2, 3 * kd)
    assert rel(g[w] - floor, dex[w]) < 0.08
    sh.close()
    m.close()
