"""A reference experiment directory (data/*.dat) replayed on the GPU.

rec2rec (satan_i/fwi3d_cpu_rec2rec_v1.4.c:38-60) runs on the directory's true
model and writes the observed spectra; an inversion iteration
(fwi3d_cpu_inv_fwi_v1.4.c:27-109) then starts from the directory's initial
model (init_model: sigma >= 4.0f -> 0.45f) and reads them back with read_Eobs.
Both are compared with the oracle run on the same files: the set-up the
directory implies (dt from minmax.dat, omega_0 as printed, 1-based positions,
Ez sources driven by waveformX) must be the one the reference would use."""
import numpy as np
import pytest

import oracle
from mittet_main_amd import refio
from mittet_main_amd.fwi3d import firstderiv_gauss, lattice_time_3d

from .test_gpu_compat import check_iteration, oracle_iteration, rel

pytestmark = pytest.mark.gpu

MINVAL, MAXVAL, ISEABED = 0.05, 5.0, 9


def make_experiment(root) -> str:
    env = refio.ModelEnv(22, 20, 18, 50.0, 45.0, 40.0, 100.0, 90, 4, 22, 2 * np.pi, 3.0, 10.4, 4,
                         (4.0, 8.0, 12.0, 0.5, 5.0), 1, 4,
                         np.array([[12, 11, 7], [9, 13, 8]]),        # 1-based
                         np.array([[8, 8, 6], [15, 13, 6], [11, 6, 6]]))
    rng = np.random.default_rng(21)
    sig = np.exp(rng.uniform(np.log(0.2), np.log(2.0), (env.iz, env.iy, env.ix)))
    sig[8:11, 7:12, 8:13] = 4.5          # a conductor the initial model replaces
    dt, _ = lattice_time_3d(MINVAL, 2 * np.pi, env.dx, env.dy, env.dz)
    d = str(root / "data")
    refio.write_experiment(d, env, sig, MINVAL, MAXVAL, ISEABED,
                           firstderiv_gauss(env.it, dt, 6.0), firstderiv_gauss(env.it, dt, 2.0), dt)
    return d


def oracle_traces(ex, src1):
    g, e = ex.grid, ex.env
    S = oracle.OracleSim(g.ix, g.iy, g.iz, g.ncpml, g.dx, g.dy, g.dz, g.dt, g.omega0, g.cmax)
    S.set_sigma(ex.sigma)
    S.init_cpml()
    S.media_coeff_3d()
    tr, _ = S.forward(2, ex.signal_x, tuple(int(c) for c in src1), e.it - 1, rec1=e.receivers,
                      it=e.it)
    return tr


def test_rec2rec_replay_writes_the_oracles_spectra(tmp_path):
    d = make_experiment(tmp_path)
    ex = refio.load_experiment(d, model="true")
    print(f"rec2rec replay: grid {ex.grid.shape}, it {ex.env.it}, dt {ex.dt:.6e}", flush=True)
    laps = refio.record_observed(ex)
    e = ex.env
    for s, src1 in enumerate(e.shots):
        tr = oracle_traces(ex, src1)
        for c in range(3):
            for r in range(len(e.receivers)):
                want = oracle.laplace_to_freq(tr[c, r], ex.signal_x, ex.signal_true, ex.dt,
                                              ex.grid.omega0, e.it)
                for key in ("E", "G", "OBS"):
                    ref = want[key]
                    np.testing.assert_allclose(laps[s][key][c, r], ref, rtol=1e-10,
                                               atol=1e-10 * np.abs(ref).max() + 1e-300,
                                               err_msg=f"{key} source {s} comp {c} rec {r}")
        # the files hold the same spectra at the printed precision
        obs = laps[s]["OBS"][:, :, :e.upper_f]
        back = ex.read_eobs(s)
        assert np.abs(back - obs).max() <= 1e-6 * np.abs(obs).max()
        np.testing.assert_allclose(
            refio.read_spectrum(refio.spectrum_path(d, "j", s, 1), e.it), laps[s]["J"],
            rtol=1e-6, atol=1e-6 * np.abs(laps[s]["J"]).max())


def test_inversion_iteration_from_the_directory(tmp_path):
    d = make_experiment(tmp_path)
    refio.record_observed(refio.load_experiment(d, model="true"))
    ex = refio.load_experiment(d)                       # initial model
    assert ex.sigma[9, 9, 10] == np.float32(0.45)
    e = ex.env
    eobs = [ex.read_eobs(s) for s in range(len(e.shots))]
    F = ex.compat_fwi()
    print(f"inversion iteration: {len(e.shots)} sources, {len(e.receivers)} receivers",
          flush=True)
    res = F.iterate(ex.sigma, eobs)
    want = oracle_iteration(ex.grid, ex.sigma, ex.sources, ex.receivers, e.it, ex.signal_x,
                            ex.signal_true, e.start_f, e.upper_f, ex.iseabed, ex.minval,
                            ex.maxval, eobs, F.omegas)
    check_iteration(res, want, ex.sigma)
    F.close()


def test_volume_and_trace_transforms_agree_on_the_true_model(tmp_path):
    """laplaceToFreq_3 (volume running DFT at the receivers, the inversion's
    forward) and laplaceToFreq (DFT of the receiver traces, rec2rec) give the
    same E_ff at bins [startF, upperF) on the model that made the data."""
    d = make_experiment(tmp_path)
    ex = refio.load_experiment(d, model="true")
    laps = refio.record_observed(ex)
    e = ex.env
    F = ex.compat_fwi()
    F.model.media_coeff_3d(ex.sigma)
    for s, src in enumerate(ex.sources):
        eff = F.forward_receivers(F.model, src)[:, :, e.start_f:]
        obs = laps[s]["OBS"][:, :, e.start_f:e.upper_f]
        assert rel(eff, obs) <= 1e-9, s
    F.close()
