"""The oracle's Fortran lineage against the reference's own Fortran.

tests/golden/ref_fortran_step.npz holds what ver0040/cpu_cpml4.f90,
cpu_eh_field_cpml42.f90 and cpu_waveform.f90, compiled where they lie with
amdflang (oracle/Makefile, driver oracle/ref_fortran_harness.f90), computed on
the module's compile-time grid 121 x 51 x 51 with ncpml 10
(cpu_parameter_wafe.f90:26-29): the module's dt and cmax, init_cpml's
profiles, media_coeff's coefficient volumes with per-cell mu, and the fields
and psi after three steps of the reference main loop's order
(cpu_main.f90:104-140: read_source_3d with two Hz points, e_field_cpml4,
CPML_E4, h_field_cpml4, CPML_H4).  The oracle (orf_*) must reproduce every
one of them bit for bit, at sample cells and through sums over every plane of
each volume.  The reference works on complex fields with real coefficients,
so the oracle runs the real and the imaginary parts as two real runs, the
source only in the real one (the waveform is real).  The same fixture also
pins the product's Fortran constants (fortran_dt, fortran_waveform).

A planted mistake in the oracle's inputs (sigma_e and sigma_m profiles
swapped, or the psi_Hyx stencil written with c1 as the other psi are) must
make the comparison fail, on the CPU."""
import os

import numpy as np
import pytest

import oracle
from .golden import make_fortran_fixtures as fx

FIX = fx.FIXTURE


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(FIX):
        pytest.fail("tests/golden/ref_fortran_step.npz is missing")
    return dict(np.load(FIX))


def sim(ref, sig, mu):
    dx = float(ref["dx"])
    S = oracle.OracleSim(fx.NX, fx.NY, fx.NZ, fx.NC, dx, dx, dx, float(ref["dt"]),
                         float(ref["omega0"]), float(ref["cmax"]), mu0=float(ref["mu0"]),
                         c1=float(ref["c1"]), c2=float(ref["c2"]), rcoef=0.01, alpha_max=0.2,
                         opt_to_max=(1.0, 1.0, 1.0))
    S.set_sigma(sig)
    S.f_init_cpml()
    S.f_media_coeff(mu)
    return S


def src_cells(ref):
    s = ref["src"]
    return [((s[3 * q + 2] - 1) * fx.NY + (s[3 * q + 1] - 1)) * fx.NX + (s[3 * q] - 1)
            for q in range(2)]


def run_parts(ref, sig, mu, fields, mutate=None):
    """Oracle volumes after the fixture's steps: {name_re/_im: (z, y, x)}."""
    nsteps = int(ref["dims"][4])
    signal = oracle.fortran_waveform(nsteps, float(ref["dt"]), 2.0e3)
    out = {}
    for part, tag in ((0, "re"), (1, "im")):
        S = sim(ref, sig, mu)
        if mutate:
            mutate(S)
        for q, nm in enumerate(oracle.FIELD_NAMES):
            S.arr(nm)[...] = fields[q][part]
        S.f_forward(src_cells(ref), signal if part == 0 else None, nsteps, [], nsteps + 1)
        for nm in fx.FIELDS + fx.PSIS:
            out[f"{nm}_{tag}"] = S.arr(nm).copy()
    return out


def assert_volume(ref, name, v):
    cells = ref["cells"]
    np.testing.assert_array_equal(v.reshape(-1)[cells], ref[f"at_{name}"], err_msg=name)
    np.testing.assert_array_equal(fx.axis_sums(v), ref[f"sum_{name}"], err_msg=name + " sums")


def test_module_constants(ref):
    from mittet_main_amd.fwi3d import F_MU0, F_C2, fortran_dt
    pi = 3.14159265358979
    assert float(ref["omega0"]) == 2.0 * pi * 1.0
    assert float(ref["mu0"]) == F_MU0 and float(ref["c2"]) == F_C2 and float(ref["c1"]) == 1.125
    # cmax of sea water (the smaller conductivity), mu = 0.999991 mu0
    cmax = np.sqrt(2.0 * float(ref["omega0"]) / (0.999991 * F_MU0) / 3.2)
    assert float(ref["cmax"]) == cmax
    assert float(ref["dt"]) == fortran_dt(float(ref["dx"]), cmax)


def test_waveform_equals_the_reference_source(ref):
    """Three steps from the same fields with and without the source differ
    only at the two Hz points and what the steps spread from them; the Hz
    decrement of step 1 at a point far from the other is the reference's
    s(1) dt/mu/dx^3, and the product's fortran_waveform equals the oracle's."""
    from mittet_main_amd.fwi3d import fortran_waveform
    dt = float(ref["dt"])
    want = oracle.fortran_waveform(5, dt, 2.0e3)
    got = fortran_waveform(5, dt, 2.0e3)
    np.testing.assert_array_equal(got, want)
    # a slope check of the closed form: s'(t) at t0 is -2 beta sqrt(beta/pi)
    beta = 3.14159265358979 * 4.0e6
    t0 = 3.14159265358979 / 2.0e3
    n0 = int(round(t0 / dt))
    s = oracle.fortran_waveform(n0 + 2, dt, 2.0e3)
    slope = (s[n0] - s[n0 - 2]) / (2 * dt)
    np.testing.assert_allclose(slope, -2 * beta * np.sqrt(beta / np.pi), rtol=1e-4)


def test_profiles_and_coefficients_bitwise(ref):
    sig, mu, _ = fx.inputs()
    S = sim(ref, sig, mu)
    for a in range(3):
        for nm in fx.PROFILES:
            np.testing.assert_array_equal(S.prof(a, nm), ref[f"prof_{a}_{nm}"], err_msg=f"{a} {nm}")
    for nm in fx.COEFFS:
        assert_volume(ref, nm, S.arr(nm))


def test_three_steps_bitwise(ref):
    sig, mu, fields = fx.inputs()
    got = run_parts(ref, sig, mu, fields)
    for nm, v in got.items():
        assert_volume(ref, nm, v)
    # the CPML terms are exercised: psi is nonzero in every shell
    for nm in fx.PSIS:
        assert np.count_nonzero(ref[f"at_{nm}_re"]) > 50, nm


def test_planted_mistakes_are_caught(ref):
    """Each plausible slip in the oracle's inputs or order fails the fixture."""
    sig, mu, fields = fx.inputs()
    cells = ref["cells"]

    def swap_e_m(S):
        for a in range(3):
            e, m = S.prof(a, "ESIG").copy(), S.prof(a, "MSIG").copy()
            S.prof(a, "ESIG")[...] = m
            S.prof(a, "MSIG")[...] = e
        S.f_media_coeff(mu)

    got = run_parts(ref, sig, mu, fields, mutate=swap_e_m)
    assert not np.array_equal(got["EX_re"].reshape(-1)[cells], ref["at_EX_re"])
    # mu0 everywhere instead of the per-cell mu in db
    S = sim(ref, sig, np.full(sig.shape, float(ref["mu0"])))
    assert not np.array_equal(S.arr("DB_X").reshape(-1)[cells], ref["at_DB_X"])
    # the Hz source after the E update instead of before it
    S = sim(ref, sig, mu)
    for q, nm in enumerate(oracle.FIELD_NAMES):
        S.arr(nm)[...] = fields[q][0]
    signal = oracle.fortran_waveform(3, float(ref["dt"]), 2.0e3)
    for n in range(3):
        S.f_e_field_cpml4()
        S.f_cpml_e4()
        S.f_source_hz(src_cells(ref), signal[n])
        S.f_h_field_cpml4()
        S.f_cpml_h4()
    assert not np.array_equal(fx.axis_sums(S.arr("EY")), ref["sum_EY_re"])
