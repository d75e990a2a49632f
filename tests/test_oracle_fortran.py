"""The oracle's restatement of the Fortran lineage (ver0040), pinned by
mathematics on grids of any shape.  The reference's own Fortran, compiled with
amdflang at its compile-time grid, pins it bit for bit in
tests/test_oracle_fortran_ref.py; the checks here cover odd grids and the
properties each formula must have.

* CPML profiles: closed forms at the ends of the ramp (the edge cell carries
  sigma_max, cell ncpml-1 carries 0: cpu_cpml4.f90:25-26), the mirror, the
  interior zeros, b = exp(-(sigma/kappa + alpha) dt).
* Bulk stencils: away from the CPML the Fortran E and H updates are the C
  ones, so with the same fields, coefficients and constants the two oracles
  agree bit for bit there (the C oracle is pinned against the reference's own
  C build and by closed forms).
* What differs from C, each checked where it shows: the Fortran ranges update
  the component's own axis end to end; the Hz source sits before the E update
  and scales by 1/mu; mu enters the H coefficients per cell; psi_Hyx takes c2
  where the difference has c1 (cpu_cpml4.f90:400, 413), so a constant Ez feeds
  it (c1 - c2) Ez/dx while every other CPML difference of a constant is 0."""
import numpy as np
import pytest

import oracle

OM0 = 2 * np.pi
F_MU0 = 1.2566370614e-6
F_C2 = -0.04167


def fsim(ix=26, iy=22, iz=20, nc=5, h=4.5e-3, sig=None, seed=0, mu=None):
    rng = np.random.default_rng(seed)
    sig = sig if sig is not None else rng.uniform(0.5, 5.0, (iz, iy, ix))
    cmax = np.sqrt(2 * OM0 / F_MU0 / sig.min())
    dt = 0.999 * h / cmax / np.sqrt(3.0) / 1.16667
    kw = dict(c2=F_C2, mu0=F_MU0, rcoef=0.01, alpha_max=0.2, opt_to_max=(1.0, 1.0, 1.0))
    F = oracle.OracleSim(ix, iy, iz, nc, h, h, h, dt, OM0, cmax, **kw)
    F.set_sigma(sig)
    F.f_init_cpml()
    mu = np.full(sig.size, F_MU0) if mu is None else mu
    F.f_media_coeff(mu)
    return F, sig, dt, cmax, kw, mu


def test_profiles_closed_forms():
    F, sig, dt, cmax, kw, _ = fsim()
    nc, h = 5, 4.5e-3
    smax = (3.0 + 0.0 + 1.0) * cmax * np.log(1.0 / 0.01) / (2.0 * nc * h)
    for a, n in enumerate((26, 22, 20)):
        es, ms = F.prof(a, "ESIG"), F.prof(a, "MSIG")
        assert es[0] == smax and es[nc - 1] == 0.0        # no x-only factor
        np.testing.assert_allclose(ms[0], smax * ((nc - 1.5) / (nc - 1)) ** 3, rtol=1e-15)
        np.testing.assert_array_equal(es, es[::-1])
        assert not np.any(es[nc:n - nc]) and not np.any(ms[nc:n - nc])
        ae = F.prof(a, "AE")
        np.testing.assert_allclose(ae[:nc], 0.2 * np.arange(1, nc + 1) / (nc - 1), rtol=1e-15)
        be = F.prof(a, "BE")
        np.testing.assert_allclose(be[:nc], np.exp(-(es[:nc] + ae[:nc]) * dt), rtol=1e-15)
        assert not np.any(be[nc:n - nc])
        np.testing.assert_array_equal(F.prof(a, "KED"), np.full(n, h))


def c_twin(kw, sig, dt, cmax, nc=5, h=4.5e-3):
    iz, iy, ix = sig.shape
    Cs = oracle.OracleSim(ix, iy, iz, nc, h, h, h, dt, OM0, cmax, **kw)
    Cs.set_sigma(sig)
    Cs.init_cpml()
    Cs.media_coeff_3d()
    return Cs


def load(S, fields):
    for q, nm in enumerate(oracle.FIELD_NAMES):
        S.arr(nm)[...] = fields[q]


def test_bulk_updates_equal_the_c_oracle():
    """Away from the CPML the two lineages do the same arithmetic except that
    the Fortran divides complex differences by real spacings as a complex
    quotient, (x c)/(c c) (the reference's amdflang build), where C divides
    x / c: equal to the last place, not bit for bit."""
    F, sig, dt, cmax, kw, _ = fsim(seed=1)
    Cs = c_twin(kw, sig, dt, cmax)
    rng = np.random.default_rng(2)
    fields = [rng.standard_normal(sig.shape) for _ in range(6)]
    load(F, fields)
    load(Cs, fields)
    F.f_e_field_cpml4()
    Cs.e_field_cpml42(1)
    nc = 5
    inner = (slice(nc + 1, -nc - 1),) * 3
    for nm in ("EX", "EY", "EZ"):
        c = Cs.arr(nm)[inner]
        np.testing.assert_allclose(F.arr(nm)[inner], c, rtol=0, atol=1e-14 * np.abs(c).max(),
                                   err_msg=nm)
    # the Fortran ranges reach the component's own axis ends: Ex at i = 0, ix-1
    assert np.any(F.arr("EX")[2:-1, 2:-1, 0] != fields[0][2:-1, 2:-1, 0])
    np.testing.assert_array_equal(Cs.arr("EX")[:, :, 0], fields[0][:, :, 0])
    F.f_h_field_cpml4()
    Cs.h_field_cpml42(1)
    for nm in ("HX", "HY", "HZ"):
        c = Cs.arr(nm)[inner]
        np.testing.assert_allclose(F.arr(nm)[inner], c, rtol=0, atol=1e-14 * np.abs(c).max(),
                                   err_msg=nm)


def test_permeability_enters_h_per_cell():
    """From zero H one update gives -db curl E, db = dt/mu/(1 + ...): tripling
    mu at one cell divides that cell's change by 3 and leaves its neighbours."""
    sig = np.full((20, 22, 26), 2.0)
    mu = np.full(sig.size, F_MU0)
    mu.reshape(sig.shape)[10, 11, 12] = 3 * F_MU0
    rng = np.random.default_rng(3)
    E = [rng.standard_normal(sig.shape) for _ in range(3)]
    out = []
    for m in (None, mu):
        S, *_ = fsim(sig=sig, mu=m)
        S.set_zero_eh()
        for nm, v in zip(("EX", "EY", "EZ"), E):
            S.arr(nm)[...] = v
        S.f_h_field_cpml4()
        out.append(S.arr("HZ").copy())
    a, b = out[0][10, 11, 12], out[1][10, 11, 12]
    assert a != 0 and abs(b - a / 3) <= 1e-15 * abs(a)
    np.testing.assert_array_equal(out[0][10, 11, 13], out[1][10, 11, 13])


def test_source_points_scale_by_one_over_mu():
    sig = np.full((20, 22, 26), 2.0)
    mu = np.full(sig.size, F_MU0)
    cells = [(8 * 22 + 9) * 26 + 10, (8 * 22 + 9) * 26 + 14]
    mu[cells[1]] = 2 * F_MU0
    F, _, dt, *_ = fsim(sig=sig, mu=mu)
    F.set_zero_eh()
    F.f_source_hz(cells, 0.75)
    hz = F.arr("HZ").reshape(-1)
    h = 4.5e-3
    # s dt / mu / h^3, each division as the Fortran's complex-by-real quotient
    np.testing.assert_allclose(hz[cells[0]], -(0.75 * dt / F_MU0 / h / h / h), rtol=4e-16)
    np.testing.assert_allclose(hz[cells[1]], -(0.75 * dt / (2 * F_MU0) / h / h / h), rtol=4e-16)
    assert hz[cells[1]] == 0.5 * hz[cells[0]]
    assert np.count_nonzero(hz) == 2


def test_psi_hyx_takes_c2_for_c1():
    F, sig, dt, *_ = fsim(seed=4)
    F.set_zero_eh()
    e0 = 0.37
    F.arr("EZ")[...] = e0
    F.arr("EY")[...] = e0
    F.f_cpml_h4()
    chx = F.prof(0, "CH")
    i = 2                                   # inside the low x shell [1, ncpml)
    want = chx[i] * (1.125 * e0 - F_C2 * e0 + F_C2 * e0 - F_C2 * e0) / 4.5e-3
    assert F.arr("PSI_HYX")[5, 6, i] == want and want != 0
    # the correct differences of a constant field vanish
    assert not np.any(F.arr("PSI_HZX")) and not np.any(F.arr("PSI_HXY"))
    assert not np.any(F.arr("PSI_HXZ"))


def test_forward_is_linear_in_the_source():
    F, sig, dt, *_ = fsim(seed=5)
    cells = [(10 * 22 + 11) * 26 + 12, (10 * 22 + 11) * 26 + 15]
    rec = [(6 * 22 + 7) * 26 + 8, (12 * 22 + 14) * 26 + 16]
    nt = 30
    s = np.sin(np.arange(nt) * 0.3)
    F.set_zero_eh()
    t1 = F.f_forward(cells, s, nt, rec, nt)
    F.set_zero_eh()
    t2 = F.f_forward(cells, 2.0 * s, nt, rec, nt)
    assert np.any(t1) and np.all(np.isfinite(t1))
    np.testing.assert_allclose(t2, 2.0 * t1, rtol=1e-12, atol=1e-12 * np.abs(t1).max())
