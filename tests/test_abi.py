"""C ABI checks that need no GPU: the library loads, exports every entry point
include/mittet_fdtd.h declares, and its host-only functions and argument
validation behave (no compute is launched here)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from mittet_main_amd import _lib
from mittet_main_amd.fwi3d import Grid, MU0, f_to_d_h, lattice_time_3d, shot_bytes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mittet_fdtd.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fwi3d_\w+)\s*\(", txt)))


def test_header_declarations_all_exported():
    names = declared()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_ctypes_binding_covers_header():
    assert sorted(_lib.exported_symbols()) == declared()


def test_library_loads_without_device():
    L = _lib.load()
    assert L.fwi3d_version().decode().startswith("mittet_fdtd")


def test_lattice_equals_oracle_bitwise():
    for sig_min, d in [(0.01, (50.0, 50.0, 50.0)), (0.45, (30.0, 30.0, 30.0)), (0.2, (25.0, 60.0, 35.0))]:
        for factor in (oracle.DT_FACTOR_REF, 0.99):
            a = lattice_time_3d(sig_min, 2 * np.pi, *d, factor=factor)
            b = oracle.lattice_time_3d(sig_min, 2 * np.pi, *d, factor=factor)
            assert a == b


def test_f_to_d_h_factor():
    om = 2 * np.pi * np.array([0.1, 1.0, 4.0])
    x = np.random.default_rng(0).standard_normal((5, 3)) + 1j
    y = f_to_d_h(x, om, 2 * np.pi)
    np.testing.assert_allclose(y, x * np.sqrt(-2 * 2 * np.pi / (1j * om)), rtol=1e-15)


def test_f_to_d_h_bin_zero_is_nan():
    """sqrt(-2 w0/(i w)) diverges at w = 0 (the reference divides by zero at
    k = 0, ver0040/cpu_f_to_d_h.f90:146-148): that bin is NaN, the others finite."""
    om = 2 * np.pi * np.array([0.0, 0.5])
    y = f_to_d_h(np.ones((2, 2), dtype=complex), om, 2 * np.pi)
    assert np.all(np.isnan(y[:, 0])) and np.all(np.isfinite(y[:, 1]))


def test_time_reversed_residual():
    from mittet_main_amd.fwi3d import time_reversed_residual
    obs = np.arange(6.0)
    calc = np.array([0.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    np.testing.assert_array_equal(time_reversed_residual(obs, calc), [4.0, 3.0, 2.0, 1.0, 0.0, 0.0])
    with pytest.raises(ValueError):
        time_reversed_residual(obs, calc[:-1])


def grid(**kw):
    g = dict(ix=40, iy=40, iz=40, ncpml=10, dx=50.0, dy=50.0, dz=50.0, dt=7.7e-4,
             omega0=2 * np.pi, cmax=31622.0)
    g.update(kw)
    return Grid(**g)


@pytest.mark.parametrize("bad", [dict(ncpml=1), dict(ix=20), dict(dt=0.0), dict(dx=-1.0),
                                 dict(cmax=0.0)])
def test_model_create_rejects_bad_grids(bad):
    p = grid(**bad).to_c()
    h = C.c_void_p()
    rc = _lib.load().fwi3d_model_create(C.byref(p), None, C.byref(h))
    assert rc == -1
    assert _lib.load().fwi3d_last_error()
    assert not h.value


@pytest.mark.parametrize("fuse", ["0", "1"])
def test_shot_bytes_accounts_ring_and_accumulators(fuse, monkeypatch):
    monkeypatch.setenv("MITTET_FDTD_FUSE", fuse)
    g = grid()
    n = 40 ** 3
    plain = shot_bytes(g, it=100)
    with_dft = shot_bytes(g, it=100, nf_vol=4, ring=16)
    assert plain > 8 * 6 * n
    # 15 extra ring slots of E (3 comps) + 4 complex accumulators x 3 comps + weights;
    # with MITTET_FDTD_FUSE=1 a ring of more than one slot on the LDS family
    # (forced on every grid in these tests) runs the fused E+H step, which
    # holds a second H (3 comps)
    h2 = 8 * 3 * n if fuse == "1" else 0
    assert with_dft - plain == 8 * 3 * n * 15 + 16 * 3 * 4 * n + 16 * 4 * 100 + h2
    # H sums add an H ring of the same 16 slots and 4 x 3 complex H accumulators
    with_h = shot_bytes(g, it=100, nf_vol=4, ring=16, volume_dft_h=True)
    assert with_h - with_dft == 8 * 3 * n * 16 + 16 * 3 * 4 * n
    assert shot_bytes(g, it=100, volume_dft_h=True) == plain


def test_null_arguments_are_errors():
    L = _lib.load()
    assert L.fwi3d_step_n(None, 1) == -1
    assert L.fwi3d_get_traces(None, None) == -1
    assert L.fwi3d_batch_step(None, 1) == -1
    assert L.fwi3d_model_destroy(None) == 0


def test_fmax_check_follows_the_reference_rule():
    """fmax = cmin/Glim/max(dx,dy,dz), cmin = sqrt(2 w0/mu0/maxval)
    (function.c:1392-1405, lattice.c:36): the 64^3 workload (sigma_max 3.3,
    Glim 10.4 of posiinf.c:40, 50 m) resolves 3.35 Hz, so a 3 Hz waveform
    passes and a 4 Hz one is refused."""
    L = _lib.load()
    fm = C.c_double(0.0)
    assert L.fwi3d_fmax_check(3.3, 2 * np.pi, MU0, 10.4, 50.0, 40.0, 50.0, 3.0, C.byref(fm)) == 0
    cmin = np.sqrt(2 * 2 * np.pi / MU0 / 3.3)
    assert fm.value == cmin / 10.4 / 50.0
    assert abs(fm.value - 3.35) < 0.01
    assert L.fwi3d_fmax_check(3.3, 2 * np.pi, MU0, 10.4, 50.0, 40.0, 50.0, 4.0, None) == -1
    assert b"fmax" in L.fwi3d_last_error()
    # sub-metre spacings stay doubles (the reference's int maxin3 would divide by 0)
    assert L.fwi3d_fmax_check(3.2, 2 * np.pi, MU0, 10.4, 4.5e-3, 4.5e-3, 4.5e-3, 2e3, C.byref(fm)) == 0
    assert np.isfinite(fm.value) and fm.value > 2e3
    assert L.fwi3d_fmax_check(-1.0, 2 * np.pi, MU0, 10.4, 50.0, 50.0, 50.0, 1.0, None) == -1


def test_stable_dt_is_the_lattice_limit():
    """Variant 0: the reference's courant*6/7 at the lowest conductivity
    (lattice_time_3d with factor 1); variant 1: 1/1.16667 of the Fortran dt
    rule, with per-cell mu entering through min(mu*sigma)."""
    L = _lib.load()
    sig = np.exp(np.random.default_rng(0).uniform(np.log(0.02), np.log(3.0), (12, 11, 10)))
    g = Grid(10, 11, 12, 2, 50.0, 40.0, 30.0, 1e-3, 2 * np.pi, 1.0)
    gp = g.to_c()
    d0 = L.fwi3d_stable_dt(C.byref(gp), 0, sig.ctypes.data_as(C.c_void_p), None)
    np.testing.assert_allclose(
        d0, lattice_time_3d(float(sig.min()), 2 * np.pi, 50.0, 40.0, 30.0, factor=1.0)[0],
        rtol=1e-15)
    mu = np.full(sig.shape, MU0)
    c = np.unravel_index(np.argmin(sig), sig.shape)
    mu[c] = 0.25 * MU0                     # 2x the wave speed at the slowest cell
    d1 = L.fwi3d_stable_dt(C.byref(gp), 1, sig.ctypes.data_as(C.c_void_p),
                           mu.ctypes.data_as(C.c_void_p))
    cmax = np.sqrt(2 * 2 * np.pi / (0.25 * MU0 * sig.min()))
    inv = np.sqrt(1 / 50.0 ** 2 + 1 / 40.0 ** 2 + 1 / 30.0 ** 2)
    np.testing.assert_allclose(d1, 1.0 / cmax / inv / 1.16667, rtol=1e-15)
    assert L.fwi3d_stable_dt(C.byref(gp), 2, sig.ctypes.data_as(C.c_void_p), None) == -1.0
