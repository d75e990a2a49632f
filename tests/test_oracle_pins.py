"""One-step closed forms that pin the CPML branch of the oracle's
media_coeff_3d (satan_i/fwi3d_cpu_function_v1.4.c:403-423) and its source
injection read_source_3d (:514-529), whose reference file cannot be built here
(it includes fftw3.h).  The 1-D profiles they use are pinned against the
reference's own cpmlabc.c build (tests/test_oracle_ref.py).

Each check isolates one coefficient by choosing the fields so that every
other term of the update vanishes, and states the expected value as the
physics of that term rather than in the reference's arrangement:

* ca: H = 0, psi = 0, random E.  The curl and the psi correction vanish, so
  one E step is the trapezoidal step of eps' dE/dt = -sigma_e E, with
  sigma_e the E profile of the component's OWN axis at the cell:
  E1 = E0 (2 eps'/dt - sigma_e) / (2 eps'/dt + sigma_e).
* cb: E = 0, psi = 0, H linear (D- of a ramp of slope s is (c1 + 3 c2) s):
  eps' (E1 - 0)/dt = -sigma_e (E1 + 0)/2 + curl H, so
  E1 = curl / (eps'/dt + sigma_e/2).
* da, db: the same for H with the magnetic profile over eps' (the
  reference's H coefficients carry msig dt/(2 eps'), :411-419).
* source: from zero fields one step leaves only the injected current,
  E[src] = -J dt / eps', J = (s[n] + s[n+1])/2 / (dx dy dz), eps' = sigma/(2 w0),
  at two conductivities of the source cell other than 1 and with
  s[n] != s[n+1].
For every check the test also shows that the plausible slips (another
axis's profile, the magnetic for the electric profile, x sigma for / sigma,
s[n] for the half-step average) miss the oracle's numbers.
"""
import numpy as np
import pytest

import oracle
from mittet_main_amd.fwi3d import firstderiv_gauss

OM0 = 2 * np.pi
DX, DY, DZ = 50.0, 40.0, 30.0
E_NAMES, H_NAMES = ("EX", "EY", "EZ"), ("HX", "HY", "HZ")


def make_sim(ix=30, iy=26, iz=22, nc=6, seed=0):
    sig = np.exp(np.random.default_rng(seed).uniform(np.log(0.05), np.log(3.0), (iz, iy, ix)))
    dt, cmax = oracle.lattice_time_3d(float(sig.min()), OM0, DX, DY, DZ, factor=0.99)
    S = oracle.OracleSim(ix, iy, iz, nc, DX, DY, DZ, dt, OM0, cmax)
    S.set_sigma(sig)
    S.init_cpml()
    S.media_coeff_3d()
    S.set_zero_eh()
    return S, sig, dt


def own_axis_profile(S, d, name):
    """(z, y, x) volume of axis d's 1-D profile at each cell's index on that axis."""
    iz, iy, ix = S.shape
    shape = [1, 1, 1]
    shape[2 - d] = (ix, iy, iz)[d]
    return np.broadcast_to(S.prof(d, name).reshape(shape), S.shape)


def e_range(S, d):
    """E_d update range: own axis [1, n-1), the others [2, n-1) (function.c:1193-1230)."""
    n = (S.p.ix, S.p.iy, S.p.iz)
    sl = [slice(2, n[a] - 1) for a in range(3)]
    sl[d] = slice(1, n[d] - 1)
    return (sl[2], sl[1], sl[0])


def h_range(S, d):
    """H_d update range: own axis [1, n-1), the others [1, n-2) (function.c:1247-1284)."""
    n = (S.p.ix, S.p.iy, S.p.iz)
    sl = [slice(1, n[a] - 2) for a in range(3)]
    sl[d] = slice(1, n[d] - 1)
    return (sl[2], sl[1], sl[0])


def shell_mask(S):
    nc = S.p.ncpml
    iz, iy, ix = S.shape
    K, J, I = np.meshgrid(np.arange(iz), np.arange(iy), np.arange(ix), indexing="ij")
    return ((I < nc) | (I >= ix - nc) | (J < nc) | (J >= iy - nc) | (K < nc) | (K >= iz - nc))


def trapezoid_gain(two_eps_dt, s):
    return (two_eps_dt - s) / (two_eps_dt + s)


def test_ca_is_the_own_axis_trapezoidal_decay():
    S, sig, dt = make_sim()
    rng = np.random.default_rng(1)
    e0 = [rng.standard_normal(S.shape) for _ in range(3)]
    for nm, v in zip(E_NAMES, e0):
        S.arr(nm)[...] = v
    S.e_field_cpml42(1)
    S.epml_abcs4()
    two_eps_dt = 2 * (sig / 2.0 / OM0) / dt
    shell = shell_mask(S)
    for d, nm in enumerate(E_NAMES):
        got = S.arr(nm)
        want = e0[d] * trapezoid_gain(two_eps_dt, own_axis_profile(S, d, "ESIG"))
        r = e_range(S, d)
        # the gain is formed to a few ulp of 1, so the error scales with |E0|
        assert np.all(np.abs(got[r] - want[r]) <= 2e-15 * np.abs(e0[d][r])), nm
        # outside the update range E is untouched
        outside = np.ones(S.shape, bool)
        outside[r] = False
        np.testing.assert_array_equal(got[outside], e0[d][outside], err_msg=nm)
        damped = shell[r] & (own_axis_profile(S, d, "ESIG")[r] > 0)
        assert damped.sum() > 100
        # slips: another axis's profile, or the magnetic profile
        for slip in (own_axis_profile(S, (d + 1) % 3, "ESIG"), own_axis_profile(S, d, "MSIG")):
            bad = e0[d] * trapezoid_gain(two_eps_dt, slip)
            assert not np.allclose(got[r][damped], bad[r][damped], rtol=1e-6, atol=0), nm


def test_cb_is_the_trapezoidal_drive_of_the_curl():
    S, sig, dt = make_sim(seed=2)
    iz, iy, ix = S.shape
    K, J, I = np.meshgrid(np.arange(iz), np.arange(iy), np.arange(ix), indexing="ij")
    slopes = {"HX": (0.3, 0.7, 1.1), "HY": (1.3, -0.4, 0.9), "HZ": (-0.8, 0.5, 2.1)}
    for nm, (a, b, c) in slopes.items():
        S.arr(nm)[...] = a * I + b * J + c * K
    S.e_field_cpml42(1)
    g = oracle.C1 + 3 * oracle.C2
    dl = (DX, DY, DZ)
    d_ = {nm: [s[q] * g / dl[q] for q in range(3)] for nm, s in slopes.items()}
    curl = (d_["HZ"][1] - d_["HY"][2], d_["HX"][2] - d_["HZ"][0], d_["HY"][0] - d_["HX"][1])
    eps = sig / 2.0 / OM0
    for d, nm in enumerate(E_NAMES):
        se = own_axis_profile(S, d, "ESIG")
        want = curl[d] / (eps / dt + se / 2.0)
        r = e_range(S, d)
        np.testing.assert_allclose(S.arr(nm)[r], want[r], rtol=1e-12, atol=0, err_msg=nm)
        bad = curl[d] / (eps / dt + own_axis_profile(S, (d + 2) % 3, "ESIG") / 2.0)
        assert not np.allclose(S.arr(nm)[r], bad[r], rtol=1e-6, atol=0), nm


def test_da_is_the_own_axis_magnetic_decay():
    S, sig, dt = make_sim(seed=3)
    rng = np.random.default_rng(4)
    h0 = [rng.standard_normal(S.shape) for _ in range(3)]
    for nm, v in zip(H_NAMES, h0):
        S.arr(nm)[...] = v
    S.h_field_cpml42(1)
    S.hpml_abcs4()
    two_eps_dt = 2 * (sig / 2.0 / OM0) / dt
    for d, nm in enumerate(H_NAMES):
        got = S.arr(nm)
        want = h0[d] * trapezoid_gain(two_eps_dt, own_axis_profile(S, d, "MSIG"))
        r = h_range(S, d)
        assert np.all(np.abs(got[r] - want[r]) <= 2e-15 * np.abs(h0[d][r])), nm
        outside = np.ones(S.shape, bool)
        outside[r] = False
        np.testing.assert_array_equal(got[outside], h0[d][outside], err_msg=nm)
        for slip in (own_axis_profile(S, (d + 1) % 3, "MSIG"), own_axis_profile(S, d, "ESIG")):
            bad = h0[d] * trapezoid_gain(two_eps_dt, slip)
            assert not np.allclose(got[r], bad[r], rtol=1e-6, atol=0), nm


def test_db_is_the_magnetic_drive_of_the_curl():
    S, sig, dt = make_sim(seed=5)
    iz, iy, ix = S.shape
    K, J, I = np.meshgrid(np.arange(iz), np.arange(iy), np.arange(ix), indexing="ij")
    slopes = {"EX": (0.2, -0.9, 1.7), "EY": (1.4, 0.6, -0.3), "EZ": (-1.1, 0.8, 0.25)}
    for nm, (a, b, c) in slopes.items():
        S.arr(nm)[...] = (a * I + b * J + c * K) * 1e-3
    S.h_field_cpml42(1)
    g = oracle.C1 + 3 * oracle.C2
    dl = (DX, DY, DZ)
    d_ = {nm: [s[q] * 1e-3 * g / dl[q] for q in range(3)] for nm, s in slopes.items()}
    curl = (d_["EZ"][1] - d_["EY"][2], d_["EX"][2] - d_["EZ"][0], d_["EY"][0] - d_["EX"][1])
    eps = sig / 2.0 / OM0
    mu0 = oracle.MU0_C
    for d, nm in enumerate(H_NAMES):
        sm = own_axis_profile(S, d, "MSIG")
        # mu0 (H1 - 0)/dt = -(mu0 sm/eps') (H1 + 0)/2 - curl E
        want = -curl[d] / (mu0 / dt + mu0 * sm / (2.0 * eps))
        r = h_range(S, d)
        np.testing.assert_allclose(S.arr(nm)[r], want[r], rtol=1e-12, atol=0, err_msg=nm)
        bad = -curl[d] / (mu0 / dt + sm / 2.0)    # sigma_m over mu0 instead of over eps'
        assert not np.allclose(S.arr(nm)[r], bad[r], rtol=1e-6, atol=0), nm


@pytest.mark.parametrize("sig_src", [0.07, 2.9])
@pytest.mark.parametrize("comp", [0, 2])
def test_source_is_the_half_step_current_over_eps(sig_src, comp):
    S, sig, dt = make_sim(seed=6)
    src = (14, 11, 9)                       # 0-based x, y, z
    sig = sig.copy()
    sig[src[2], src[1], src[0]] = sig_src
    S.set_sigma(sig)
    S.media_coeff_3d()
    it, n = 40, 17
    signal = firstderiv_gauss(it, dt, 3.0) * 1e4
    assert signal[n] != signal[n + 1]
    S.forward(comp, signal, tuple(q + 1 for q in src), 1, it=it, step0=n)
    e = S.arr(E_NAMES[comp])
    eps = sig_src / (2.0 * OM0)
    J = 0.5 * (signal[n] + signal[n + 1]) / (DX * DY * DZ)
    np.testing.assert_allclose(e[src[2], src[1], src[0]], -J * dt / eps, rtol=1e-15)
    # the current is the only thing a zero state produces in E
    assert np.count_nonzero(e) == 1
    for other in set(range(3)) - {comp}:
        assert not np.any(S.arr(E_NAMES[other]))
    v = e[src[2], src[1], src[0]]
    for bad in (-J * dt * sig_src * 2.0 * OM0,                      # x sigma
                -signal[n] / (DX * DY * DZ) * dt / eps,            # s[n], no average
                -signal[n + 1] / (DX * DY * DZ) * dt / eps):
        assert not np.isclose(v, bad, rtol=1e-6, atol=0)


def test_product_waveform_equals_the_oracle():
    """The waveform the Python host hands to every shot (firstderiv_gauss,
    satan_i/fwi3d_cpu_waveform_v1.4.c:132-143) against the oracle's C
    restatement, which tests/test_oracle_physics.py pins by its closed form:
    equal to the last place over long traces at several fmax and dt."""
    for nt, dt, fm in ((8000, 7.746e-4, 3.0), (2000, 1.3e-3, 1.0), (500, 1e-4, 40.0)):
        a = firstderiv_gauss(nt, dt, fm)
        b = oracle.firstderiv_gauss(nt, dt, fm)
        scale = np.abs(b).max()
        assert scale > 0
        np.testing.assert_allclose(a, b, rtol=0, atol=4e-16 * scale)
        # both take sample n at t = n dt: a one-sample shift would fail
        assert np.abs(a[1:] - b[:-1]).max() > 1e-6 * scale
