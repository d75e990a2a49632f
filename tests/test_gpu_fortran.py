"""The Fortran lineage's step (model variant 1) on the GPU against the oracle's
literal restatement of its loops (ver0040/cpu_main.f90:104-148): Hz source
points before the E update, E/H over the Fortran ranges, CPML over the
Fortran shells with its psi_Hyx stencil, per-cell permeability.  Bit for bit:
fields after the run and E/H traces at every step, also with a volume DFT
(ring slots) and as members of a multi-shot batch."""
import numpy as np
import pytest

import oracle
from mittet_main_amd import Batch
from mittet_main_amd.fwi3d import (F_MU0, SRC_HZ, Model, fortran_dt, fortran_grid,
                                   fortran_waveform)

pytestmark = pytest.mark.gpu

OM0 = 2 * np.pi


def setup(ix, iy, iz, nc=5, h=4.5e-3, seed=0, nt=40):
    rng = np.random.default_rng(seed)
    sig = rng.uniform(0.5, 5.0, (iz, iy, ix))
    mu = np.full((iz, iy, ix), F_MU0)
    mu[iz // 2:, :, : ix // 2] = 0.999991 * F_MU0          # seawater-like region
    mu[iz // 3, iy // 3, ix // 3] = 4 * F_MU0
    cmax = np.sqrt(2 * OM0 / F_MU0 / sig.min())
    dt = fortran_dt(h, cmax)
    g = fortran_grid(ix, iy, iz, nc, h, h, h, dt, OM0, cmax)
    sigw = fortran_waveform(nt, dt, 40.0)
    src = [(ix // 2, iy // 2, iz // 2), (ix // 2 + 3, iy // 2 - 1, iz // 2 + 1)]
    rec = np.array([[ix // 2 + 4, iy // 2, iz // 2], [nc + 1, iy // 2, iz // 2],
                    [ix - nc - 2, iy - nc - 2, iz - nc - 2], [0, 0, 0], [ix - 1, iy - 1, iz - 1]])
    return g, sig, mu, sigw, src, rec, dt, cmax


def oracle_run(g, sig, mu, sigw, src, rec, nt):
    kw = dict(c2=g.c2, mu0=g.mu0, rcoef=g.rcoef, alpha_max=g.alpha_max, opt_to_max=g.opt_to_max)
    S = oracle.OracleSim(g.ix, g.iy, g.iz, g.ncpml, g.dx, g.dy, g.dz, g.dt, g.omega0, g.cmax, **kw)
    S.set_sigma(sig)
    S.f_init_cpml()
    S.f_media_coeff(mu)
    S.set_zero_eh()
    flat = lambda c: (c[2] * g.iy + c[1]) * g.ix + c[0]
    tr = S.f_forward([flat(c) for c in src], sigw, nt, [flat(c) for c in rec], nt + 1)
    return S, tr


def gpu_shot(model, g, sigw, src, rec, nt, omegas=None, ring=0):
    sh = model.shot(nt + 1, np.append(sigw, 0.0), src[0], SRC_HZ, rec, omegas, ring)
    for c in src[1:]:
        sh.add_source(*c)
    return sh


@pytest.mark.parametrize("dims", [(30, 26, 22), (27, 23, 21)])
def test_fortran_step_bitwise(dims):
    nt = 40
    g, sig, mu, sigw, src, rec, *_ = setup(*dims, nt=nt)
    S, tr_o = oracle_run(g, sig, mu, sigw, src, rec, nt)
    m = Model(g, variant=1)
    m.set_mu(mu)
    m.media_coeff_3d(sig)
    print(f"fortran variant: {dims}, {nt} steps", flush=True)
    sh = gpu_shot(m, g, sigw, src, rec, nt)
    for p in range(3):
        np.testing.assert_array_equal(m.profile(p, "ESIG"), S.prof(p, "ESIG"))
        np.testing.assert_array_equal(m.profile(p, "CH"), S.prof(p, "CH"))
    sh.run()
    tr = sh.traces()
    np.testing.assert_array_equal(tr[:, :, :nt], tr_o[:, :, :nt])
    for q, nm in enumerate(oracle.FIELD_NAMES):
        np.testing.assert_array_equal(sh.field(q), S.arr(nm), err_msg=nm)
    assert np.any(tr[5]) and np.all(np.isfinite(tr))
    # a volume DFT (E in ring slots) leaves the fields and traces unchanged
    sh2 = gpu_shot(m, g, sigw, src, rec, nt, omegas=[1.0, 3.0], ring=8)
    sh2.run()
    np.testing.assert_array_equal(sh2.traces(), tr)
    for q in range(6):
        np.testing.assert_array_equal(sh2.field(q), sh.field(q))
    m.close()


def test_fortran_shots_in_a_batch_equal_alone():
    nt = 30
    g, sig, mu, sigw, src, rec, *_ = setup(28, 24, 22, nt=nt, seed=3)
    m = Model(g, variant=1)
    m.set_mu(mu)
    m.media_coeff_3d(sig)
    alone = []
    for s0 in (src, src[::-1]):
        sh = gpu_shot(m, g, sigw, s0, rec, nt)
        sh.run()
        alone.append(sh.traces())
        sh.close()
    b = Batch(4, chunk=8)
    shots = [gpu_shot(m, g, sigw, s0, rec, nt) for s0 in (src, src[::-1])]
    for sh in shots:
        b.add(sh)
    b.step(nt)
    b.sync()
    for sh, a in zip(shots, alone):
        np.testing.assert_array_equal(sh.traces(), a)
    b.close()
    m.close()


def test_variant_guards():
    g, sig, mu, sigw, src, rec, *_ = setup(24, 22, 20, nt=10)
    from mittet_main_amd._lib import Fwi3dError
    m0 = Model(g)
    with pytest.raises(Fwi3dError):
        m0.set_mu(mu)
    with pytest.raises(Fwi3dError):
        m0.shot(11, np.append(sigw[:10], 0.0), src[0], SRC_HZ, rec)
    m0.close()
    m1 = Model(g, variant=1)
    m1.media_coeff_3d(sig)
    with pytest.raises(Fwi3dError):
        m1.shot(11, np.append(sigw[:10], 0.0), src[0], 0, rec)
    sh = m1.shot(11, np.append(sigw[:10], 0.0), src[0], SRC_HZ, rec)
    with pytest.raises(Fwi3dError):
        sh.e_field_cpml42()
    for c in [(1, 1, 1)] * 3:
        sh.add_source(*c)
    with pytest.raises(Fwi3dError):
        sh.add_source(2, 2, 2)
    m1.close()
