"""Oracle pinned against the reference's own code (CPU only).

The reference's satan_i/fwi3d_cpu_cpmlabc_v1.4.c and fwi3d_cpu_lattice_v1.4.c
build from their own sources; tests/golden/ref_cpml.npz holds their outputs
(written by tests/golden/make_ref_fixtures.py).  The oracle must reproduce them
bit for bit.  Where the reference tree is mounted the same comparison also runs
live against oracle/_ref/libref_cpml.so on fresh random inputs.
"""
import os

import numpy as np
import pytest

import oracle
from tests.golden.make_ref_fixtures import DX, DY, DZ, IX, IY, IZ, NC, OMEGA0, hash_field, sigma

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_cpml.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


def test_lattice_matches_reference(gold):
    dt, cmax = oracle.lattice_time_3d(float(gold["minval"]), OMEGA0, DX, DY, DZ)
    assert dt == float(gold["dt"])
    assert cmax == float(gold["cmax"])


def test_profiles_match_reference(gold):
    S = oracle.OracleSim(IX, IY, IZ, NC, DX, DY, DZ, float(gold["dt"]), OMEGA0, float(gold["cmax"]))
    S.init_cpml()
    for a in range(3):
        for nm in oracle.PROFILE_IDS:
            np.testing.assert_array_equal(S.prof(a, nm), gold[f"prof_{a}_{nm}"], err_msg=f"{a} {nm}")


def test_cpml_psi_updates_match_reference(gold):
    S = oracle.OracleSim(IX, IY, IZ, NC, DX, DY, DZ, float(gold["dt"]), OMEGA0, float(gold["cmax"]))
    S.init_cpml()
    N = IX * IY * IZ
    for q, nm in enumerate(["CB_X", "CB_Y", "CB_Z", "DB_X", "DB_Y", "DB_Z"]):
        S.arr(nm)[...] = hash_field(10 + q, N, lo=0.5, scale=1e-3).reshape(IZ, IY, IX)
    for q, nm in enumerate(oracle.FIELD_NAMES + oracle.PSI_NAMES):
        S.arr(nm)[...] = hash_field(20 + q, N).reshape(IZ, IY, IX)
    for _ in range(2):
        S.epml_abcs4()
        S.hpml_abcs4()
    for nm in oracle.FIELD_NAMES + oracle.PSI_NAMES:
        np.testing.assert_array_equal(S.arr(nm), gold[f"after_{nm}"], err_msg=nm)


def test_fixture_inputs_are_reproducible(gold):
    sig = sigma()
    assert sig.min() == float(gold["minval"]) and sig.max() == float(gold["maxval"])


@pytest.mark.skipif(not oracle.RefCpml.available(), reason="reference tree not mounted")
@pytest.mark.parametrize("dims", [(30, 26, 24, 8, 50.0, 50.0, 50.0), (21, 33, 18, 6, 25.0, 60.0, 35.0)])
def test_live_reference_cpml(dims):
    ix, iy, iz, nc, dx, dy, dz = dims
    rng = np.random.default_rng(ix * 7 + iy)
    sig = np.exp(rng.uniform(np.log(0.01), np.log(3.3), (iz, iy, ix)))
    R = oracle.RefCpml(ix, iy, iz, nc, dx, dy, dz, OMEGA0, sig.min(), sig.max())
    R.put("SIG", sig)
    dtr, cmr = R.lattice()
    dt, cm = oracle.lattice_time_3d(sig.min(), OMEGA0, dx, dy, dz)
    assert (dt, cm) == (dtr, cmr)
    R.init_cpml()
    S = oracle.OracleSim(ix, iy, iz, nc, dx, dy, dz, dt, OMEGA0, cm)
    S.set_sigma(sig)
    S.init_cpml()
    S.media_coeff_3d()
    for a in range(3):
        for nm in oracle.PROFILE_IDS:
            np.testing.assert_array_equal(S.prof(a, nm), R.profile(a, nm))
    for nm in ["CA_X", "CA_Y", "CA_Z", "CB_X", "CB_Y", "CB_Z", "DA_X", "DA_Y", "DA_Z", "DB_X",
               "DB_Y", "DB_Z"]:
        R.put(nm, S.arr(nm))
    for nm in oracle.FIELD_NAMES + oracle.PSI_NAMES:
        v = rng.standard_normal((iz, iy, ix))
        R.put(nm, v)
        S.arr(nm)[...] = v
    for _ in range(3):
        R.epml()
        S.epml_abcs4()
        R.hpml()
        S.hpml_abcs4()
    for nm in oracle.FIELD_NAMES + oracle.PSI_NAMES:
        np.testing.assert_array_equal(S.arr(nm), R.get(nm), err_msg=nm)
