"""Pins the oracle's restatement of the reference's FWI gradient chain
(satan_i/fwi3d_cpu_function_v1.4.c:2202-2384, 2565-2603, 2757-2832) by what
the mathematics fixes: scalings that must be exact (powers of two), special
cases that reduce to one complex product, the cosine bound of a single
correlation, the step-length closed form, and the update's masks and clamps.
Each case is built so that a plausible slip (a missing conj, a swapped
direction/component index, a wrong reduction mask) fails it."""
import numpy as np

import oracle

RNG = np.random.default_rng(7)


def crand(*shape):
    return RNG.normal(size=shape) + 1j * RNG.normal(size=shape)


def spectra(n=60, nrec=2, upper=4, start=1):
    eff = crand(3, upper, n)
    ebb = crand(3, nrec, 3, upper, n)
    eff[:, :start] = 0
    ebb[:, :, :, :start] = 0
    de = crand(3, nrec, upper)
    return eff, ebb, de


def test_laplace_volume_divides_by_j_and_scales_by_t():
    n, upper, start = 40, 5, 2
    acc = crand(3, upper - start, n)
    jw, tw = crand(upper - start), crand(upper - start)
    g = oracle.laplace_volume(acc, jw, None, upper, start)
    eff = oracle.laplace_volume(acc, jw, tw, upper, start)
    assert not np.any(g[:, :start]) and not np.any(eff[:, :start])
    np.testing.assert_allclose(g[:, start:] * jw[None, :, None], acc, rtol=1e-14)
    np.testing.assert_allclose(eff[:, start:], g[:, start:] * tw[None, :, None], rtol=1e-14)


def test_residual_sums_moduli_per_component():
    eobs = np.zeros((3, 2, 3), dtype=complex)
    eff = np.zeros_like(eobs)
    eobs[0, 1, 2] = 3 + 4j          # |dE| = 5
    eobs[2, 0, 1] = -6j             # |dE| = 6
    eff[1, 1, 0] = 1.0              # dE = -1
    de, err = oracle.residual_e(eobs, eff)
    assert err == 12.0
    assert de[1, 1, 0] == -1.0 and de[0, 1, 2] == 3 + 4j


def test_sensitivity_is_positively_homogeneous():
    eff, ebb, _ = spectra()
    s = oracle.sensitivity(eff, ebb)
    assert np.all(s > 0)
    np.testing.assert_array_equal(oracle.sensitivity(eff, 2 * ebb), 2 * s)
    np.testing.assert_array_equal(oracle.sensitivity(0.5 * eff, ebb), 0.5 * s)


def test_sensitivity_single_term_is_modulus_of_product():
    n, upper = 30, 3
    eff = np.zeros((3, upper, n), dtype=complex)
    ebb = np.zeros((3, 2, 3, upper, n), dtype=complex)
    eff[1, 2] = crand(n)
    ebb[1, 1, 0, 2] = crand(n)      # y component, receiver 1, direction x, bin 2
    np.testing.assert_allclose(oracle.sensitivity(eff, ebb), np.abs(eff[1, 2] * ebb[1, 1, 0, 2]),
                               rtol=1e-15)


def test_conv_uses_the_direction_of_the_residual_and_its_conjugate():
    n, upper = 25, 3
    ebb = np.zeros((3, 2, 3, upper, n), dtype=complex)
    de = np.zeros((3, 2, upper), dtype=complex)
    ebb[:, 0, 1, 2] = crand(3, n)   # receiver 0, direction y, bin 2, all components
    de[1, 0, 2] = 0.5 - 2j          # y residual at receiver 0, bin 2
    de[0, 0, 2] = 7.0               # an x residual must not reach a y-direction G
    eb = oracle.conv_gdele(ebb, de)
    np.testing.assert_allclose(eb[:, 2], ebb[:, 0, 1, 2] * np.conj(de[1, 0, 2]), rtol=1e-15)
    assert not np.any(eb[:, :2])
    # conjugate-linear in the residual: i dE -> -i E_back
    np.testing.assert_array_equal(oracle.conv_gdele(ebb, 1j * de), -1j * eb)


def test_gamma_single_correlation_is_a_cosine_bounded_by_the_residual():
    shape, nc = (6, 5, 7), 1
    iz, iy, ix = shape
    n = ix * iy * iz
    upper = 2
    eff = np.zeros((3, upper, n), dtype=complex)
    ebb = np.zeros((3, 1, 3, upper, n), dtype=complex)
    de = np.zeros((3, 1, upper), dtype=complex)
    eff[2, 1] = crand(n)
    ebb[2, 0, 2, 1] = crand(n)
    de[2, 0, 1] = 1.5 * np.exp(0.3j)
    sense = oracle.sensitivity(eff, ebb)
    eb = oracle.conv_gdele(ebb, de)
    g = oracle.calc_gamma(shape, nc, eff, eb, sense).reshape(shape)
    u = (eff[2, 1] * ebb[2, 0, 2, 1]).reshape(shape)
    expect = 1.5 * np.cos(np.angle(u) - 0.3)
    inner = (slice(nc, iz - nc), slice(nc, iy - nc), slice(nc, ix - nc))
    np.testing.assert_allclose(g[inner], expect[inner], rtol=1e-12, atol=1e-14)
    mask = np.ones(shape, bool)
    mask[inner] = False
    assert not np.any(g[mask])                      # CPML cells untouched
    # invariant under a positive rescaling of the Green's functions
    g2 = oracle.calc_gamma(shape, nc, eff, oracle.conv_gdele(2 * ebb, de),
                           oracle.sensitivity(eff, 2 * ebb)).reshape(shape)
    np.testing.assert_array_equal(g2, g)


def test_phi_takes_maxima_above_the_seabed_only():
    shape = (5, 4, 3)
    sig = RNG.uniform(0.1, 1.0, shape)
    grad = RNG.uniform(-1, 1, shape)
    iseabed = 2
    sig[4, 0, 0], grad[3, 1, 1] = 50.0, 80.0        # below the seabed: ignored by the maxima
    sig[1, 2, 2], grad[0, 3, 1] = 3.0, 4.0          # the maxima
    sig2, scaler = oracle.calc_phi(shape, iseabed, sig, grad)
    assert scaler == np.float32(0.01) * 3.0 / 4.0
    np.testing.assert_array_equal(sig2, (sig - scaler * grad).ravel())
    # at the gradient maximum the step is 0.01 (as a float literal) x max sigma
    assert abs(sig2.reshape(shape)[0, 3, 1] - (sig[0, 3, 1] - float(np.float32(0.01)) * 3.0)) < 1e-15


def test_alpha_closed_form_for_a_proportional_change():
    de = crand(3, 3, 4)
    effl = crand(3, 3, 4)
    c = 0.7 - 1.1j
    k1, k2 = oracle.calc_alpha(de, effl + c * de, effl)
    s = np.sum(de * de)
    np.testing.assert_allclose(k1, c * s, rtol=1e-12)
    np.testing.assert_allclose(k2, c * c * s, rtol=1e-12)
    alpha, _ = oracle.update_para2((3, 3, 3), 1, 0, k1, k2, 1.0, 0.0, 1e9, np.ones(27), np.zeros(27))
    assert abs(alpha - 3.0 / abs(c)) < 1e-12 * alpha


def test_update_masks_clamps_and_zeroes_grad():
    shape, nc, iseabed = (6, 5, 5), 1, 2
    sig = np.full(shape, 1.0)
    grad = np.full(shape, 10.0)
    grad[2, 2, 2] = -10.0
    alpha, new = oracle.update_para2(shape, nc, iseabed, 3 + 0j, 1 + 0j, 0.02, 0.5, 1.5, sig, grad)
    assert alpha == 9.0                               # 3 |k1| / |k2|
    new = new.reshape(shape)
    assert new[1, 2, 2] == 1.5                        # 1 + 9*0.02*10 clamped to maxval
    assert new[2, 2, 2] == 0.5                        # 1 - 1.8 clamped to minval
    assert new[3, 2, 2] == 1.0                        # below the seabed: no step
    assert new[0, 2, 2] == 1.0 and new[2, 0, 2] == 1.0   # CPML cells untouched
    sig[3, 2, 2] = 9.0
    _, new = oracle.update_para2(shape, nc, iseabed, 3 + 0j, 1 + 0j, 0.02, 0.5, 1.5, sig, grad)
    assert new.reshape(shape)[3, 2, 2] == 1.5         # clamped even below the seabed
