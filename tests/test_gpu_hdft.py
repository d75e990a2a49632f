"""Volume running DFT of H (fwi3d_enable_volume_dft_h): bit for bit
sum_n H[n] w_n with H after step n, against the oracle stepped one step at a
time and the sums formed in numpy in the same order (the weights are the ones
the E volume DFT is checked with); alone, in a batch, and after set_zero_eh."""
import numpy as np
import pytest

import oracle
from mittet_main_amd import Batch

from .test_gpu_parity import OM0, setup_pair

pytestmark = pytest.mark.gpu


def oracle_h_sums(S, signal, src, it, w):
    nf = w.shape[1]
    re = np.zeros((3, nf, S.n))
    im = np.zeros((3, nf, S.n))
    src1 = tuple(c + 1 for c in src)
    for n in range(it - 1):
        S.forward(0, signal, src1, 1, it=it, step0=n)
        for d, nm in enumerate(("HX", "HY", "HZ")):
            h = S.arr(nm).reshape(-1)
            for f in range(nf):
                re[d, f] += h * w[n, f].real
                im[d, f] += h * w[n, f].imag
    return re, im


@pytest.mark.parametrize("family", ["lds", "auto"])
def test_h_volume_dft_matches_oracle_steps(family, monkeypatch):
    if family == "auto":
        monkeypatch.delenv("MITTET_FDTD_LDS", raising=False)
    it = 40
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(seed=31, it=it, ring=8)
    print(f"H volume DFT: {family} family, it {it}, nf {len(om)}", flush=True)
    shot.enable_volume_dft_h()
    shot.run()
    w = oracle.dft_weights(it, dt, OM0, om)
    re, im = oracle_h_sums(S, signal, src, it, w)
    got = shot.volume_dft_h().reshape(3, len(om), -1)
    np.testing.assert_array_equal(got.real, re)
    np.testing.assert_array_equal(got.imag, im)
    assert np.abs(got).max() > 0
    # the E sums beside it are unchanged: equal to a shot without the H sums
    _, plain, *_ = setup_pair(seed=31, it=it, ring=8)
    plain.run()
    np.testing.assert_array_equal(shot.volume_dft(), plain.volume_dft())
    np.testing.assert_array_equal(shot.traces(), plain.traces())
    # reset, run again: the same sums
    shot.set_zero_eh()
    assert not np.any(shot.volume_dft_h())
    shot.run()
    np.testing.assert_array_equal(shot.volume_dft_h().reshape(3, len(om), -1), got)


def test_h_volume_dft_in_a_batch_equals_alone():
    specs = [dict(ix=30, iy=26, iz=22, seed=32, it=50), dict(ix=26, iy=24, iz=20, seed=33, it=40)]
    alone = []
    for q, sp in enumerate(specs):
        m, shot, *_ = setup_pair(ring=8, **sp)
        if q == 0:
            shot.enable_volume_dft_h()
        shot.run()
        alone.append((shot.volume_dft_h() if q == 0 else None, shot.volume_dft(), shot.traces()))
    b = Batch(4, chunk=8)
    keep = []
    for q, sp in enumerate(specs):
        m, shot, *_ = setup_pair(ring=8, **sp)
        if q == 0:
            shot.enable_volume_dft_h()
        keep.append((m, shot))
        b.add(shot)
    for _ in range(8):
        b.step(8)
    b.sync()
    for q, (m, shot) in enumerate(keep):
        b.remove(shot)
        h, vol, tr = alone[q]
        if h is not None:
            np.testing.assert_array_equal(shot.volume_dft_h(), h)
        np.testing.assert_array_equal(shot.volume_dft(), vol)
        np.testing.assert_array_equal(shot.traces(), tr)
    b.close()


def test_h_volume_dft_guards():
    from mittet_main_amd._lib import Fwi3dError
    m, shot, *_ = setup_pair(seed=34, it=20, ring=8)
    with pytest.raises(Fwi3dError):
        shot.volume_dft_h()               # not enabled
    shot.step_n(3)
    with pytest.raises(Fwi3dError):
        shot.enable_volume_dft_h()        # after the first step
    m2, noft, *_ = setup_pair(seed=34, it=20, omegas=())
    with pytest.raises(Fwi3dError):
        noft.enable_volume_dft_h()        # no volume DFT frequencies
