"""Failure detection on the device path: a model whose lowest conductivity
makes the grid's dt unstable is refused before anything is uploaded, and a
shot whose fields turn NaN reports it through its divergence flag (set by
the receiver gathers and the end-of-chunk probe, cleared by set_zero_eh).
Nothing here faults: NaN arithmetic is ordinary floating point."""
import numpy as np
import pytest

from mittet_main_amd import Grid, Model
from mittet_main_amd._lib import Fwi3dError
from mittet_main_amd.fwi3d import firstderiv_gauss, lattice_time_3d

pytestmark = pytest.mark.gpu

OM0 = 2 * np.pi


def grid_for(sig_min, shape=(22, 24, 26), nc=5):
    iz, iy, ix = shape
    dt, cmax = lattice_time_3d(sig_min, OM0, 50.0, 45.0, 40.0, factor=0.99)
    return Grid(ix, iy, iz, nc, 50.0, 45.0, 40.0, dt, OM0, cmax)


def test_unstable_model_is_refused_and_the_model_kept():
    g = grid_for(0.1)
    rng = np.random.default_rng(0)
    sig = np.exp(rng.uniform(np.log(0.1), np.log(2.0), g.shape))
    m = Model(g)
    m.media_coeff_3d(sig)
    it = 40
    signal = firstderiv_gauss(it, g.dt, 6.0) * 1e3
    rec = np.array([[13, 12, 11], [8, 9, 10]], dtype=np.int32)
    print("stability refusal: 26 x 24 x 22, dt for sigma_min 0.1, model with 0.05", flush=True)

    def traces():
        sh = m.shot(it, signal, (13, 12, 11), 0, rec)
        sh.run()
        out = sh.traces()
        assert not sh.diverged
        sh.close()
        return out

    before = traces()
    bad = sig.copy()
    bad[3, 4, 5] = 0.05       # c grows by sqrt(2): dt = 0.99 x the old limit is unstable
    with pytest.raises(Fwi3dError) as e:
        m.media_coeff_3d(bad)
    assert "stability" in str(e.value)
    np.testing.assert_array_equal(traces(), before)
    m.close()


def test_divergence_flag_set_by_nan_and_cleared():
    g = grid_for(0.1)
    sig = np.full(g.shape, 0.5)
    m = Model(g)
    m.media_coeff_3d(sig)
    it = 30
    signal = firstderiv_gauss(it, g.dt, 6.0) * 1e3
    src = (13, 12, 11)
    rec = np.array([[16, 12, 11], [13, 15, 11]], dtype=np.int32)
    print("divergence flag: NaN planted next to the source, receivers 3 cells away",
          flush=True)
    for ring, om in ((0, None), (8, 2 * np.pi * np.array([0.5, 1.0]))):
        sh = m.shot(it, signal, src, 0, rec, om, ring=ring)
        sh.step_n(5)
        assert not sh.diverged
        e = sh.field(0)
        e[11, 12, 14] = np.nan
        sh.set_field(0, e)
        sh.step_n(10)
        assert sh.diverged
        assert np.isnan(sh.traces()).any()
        sh.set_zero_eh()
        assert not sh.diverged
        sh.run()
        assert not sh.diverged
        assert np.all(np.isfinite(sh.traces()))
        sh.close()
    m.close()
