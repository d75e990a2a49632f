"""fwi3d_receiver_dft_noalloc (the receiver spectra staged in a fixed device
array instead of per-call buffers) against fwi3d_receiver_dft, through the C
ABI: two calls in a row, the second on a smaller request (fewer receivers and
frequencies), each bit for bit equal to the allocating call; and a request
beyond the fixed array's size refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from mittet_main_amd import _lib

from .test_gpu_parity import setup_pair

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(C.c_void_p)


def spectra(fn, shot, om, hfac):
    out = np.zeros(2 * 6 * shot.nrec * om.size)
    rc = fn(shot._h, om.size, _dp(om), hfac, _dp(out))
    _lib.check(rc)
    return out


def test_noalloc_receiver_dft_equals_the_allocating_call():
    L = _lib.load()
    for nrec, freqs, seed in ((12, (0.25, 0.5, 1.0, 2.0, 3.0, 4.5), 81), (3, (0.7, 1.9), 82)):
        print(f"noalloc receiver DFT: {nrec} receivers, {len(freqs)} frequencies", flush=True)
        m, shot, *_ = setup_pair(seed=seed, it=48, nrec=nrec, ring=8)
        shot.run()
        om = 2 * np.pi * np.asarray(freqs)
        for hfac in (1, 0):
            want = spectra(L.fwi3d_receiver_dft, shot, om, hfac)
            got = spectra(L.fwi3d_receiver_dft_noalloc, shot, om, hfac)
            np.testing.assert_array_equal(got, want)
            assert np.abs(got).max() > 0
        shot.close()
        m.close()


def test_noalloc_receiver_dft_refuses_requests_beyond_its_array():
    L = _lib.load()
    m, shot, *_ = setup_pair(seed=83, it=20, nrec=64, ring=8)
    shot.run()
    nf = 1400        # 12 * 64 * 1400 + 1408 doubles > 2^19
    om = 2 * np.pi * np.linspace(0.1, 5.0, nf)
    out = np.zeros(2 * 6 * 64 * nf)
    rc = L.fwi3d_receiver_dft_noalloc(shot._h, nf, _dp(om), 1, _dp(out))
    assert rc == -1
    assert b"scratch" in L.fwi3d_last_error()
    # the same request through the allocating call still works, and a small
    # one through the allocation-free call after the refusal too
    _lib.check(L.fwi3d_receiver_dft(shot._h, nf, _dp(om), 1, _dp(out)))
    small = om[:4].copy()
    a = spectra(L.fwi3d_receiver_dft, shot, small, 1)
    b = spectra(L.fwi3d_receiver_dft_noalloc, shot, small, 1)
    np.testing.assert_array_equal(b, a)
    shot.close()
    m.close()
