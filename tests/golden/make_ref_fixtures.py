"""Writes tests/golden/ref_cpml.npz from the reference's own code.

Runs ONLY where the reference tree is mounted (it builds oracle/_ref from
satan_i/fwi3d_cpu_cpmlabc_v1.4.c and fwi3d_cpu_lattice_v1.4.c via
oracle/Makefile) and records, for fixed inputs:
  * lattice_time_3d's dt and cmax (lattice.c:6-54),
  * init_cpml's twelve 1-D profiles per axis (cpmlabc.c:7-175),
  * fields and psi after two rounds of epml_abcs4 / hpml_abcs4
    (cpmlabc.c:191-203, 402-679) started from integer-hash inputs.
Grid 14x13x12 with ncpml 4 and dx, dy, dz = 50, 40, 30 m, so that an axis or
spacing mix-up changes the numbers.

Usage:  python tests/golden/make_ref_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

IX, IY, IZ, NC = 14, 13, 12, 4
DX, DY, DZ = 50.0, 40.0, 30.0
OMEGA0 = 6.283185307179586


def hash_field(a: int, n: int, lo: float = -0.5, scale: float = 1.0) -> np.ndarray:
    """Deterministic inputs from exact integer arithmetic (no libm)."""
    idx = np.arange(n, dtype=np.uint64)
    h = (idx * np.uint64(2654435761) + np.uint64(97 * a + 13)) % np.uint64(2 ** 32)
    return (h.astype(np.float64) / 2.0 ** 32 + lo) * scale


def sigma() -> np.ndarray:
    s = 0.01 * 330.0 ** (hash_field(0, IX * IY * IZ, lo=0.0))  # values in [0.01, 3.3)
    return s.reshape(IZ, IY, IX)


def main():
    oracle.build()
    if not oracle.RefCpml.available():
        raise SystemExit("reference tree not mounted: oracle/_ref was not built")
    sig = sigma()
    R = oracle.RefCpml(IX, IY, IZ, NC, DX, DY, DZ, OMEGA0, float(sig.min()), float(sig.max()))
    R.put("SIG", sig)
    dt, cmax = R.lattice()
    R.init_cpml()
    out = {"dims": np.array([IX, IY, IZ, NC]), "steps": np.array([DX, DY, DZ]),
           "omega0": np.array(OMEGA0), "dt": np.array(dt), "cmax": np.array(cmax),
           "minval": np.array(sig.min()), "maxval": np.array(sig.max())}
    for a in range(3):
        for nm in oracle.PROFILE_IDS:
            out[f"prof_{a}_{nm}"] = R.profile(a, nm)
    N = IX * IY * IZ
    coeffs = ["CB_X", "CB_Y", "CB_Z", "DB_X", "DB_Y", "DB_Z"]
    for q, nm in enumerate(coeffs):
        R.put(nm, hash_field(10 + q, N, lo=0.5, scale=1e-3).reshape(IZ, IY, IX))
    for q, nm in enumerate(oracle.FIELD_NAMES + oracle.PSI_NAMES):
        R.put(nm, hash_field(20 + q, N).reshape(IZ, IY, IX))
    for _ in range(2):
        R.epml()
        R.hpml()
    for nm in oracle.FIELD_NAMES + oracle.PSI_NAMES:
        out[f"after_{nm}"] = R.get(nm)
    path = os.path.join(HERE, "ref_cpml.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
