"""Writes tests/golden/ref_fortran_step.npz from the reference's own Fortran.

Runs ONLY where the reference tree is mounted: oracle/Makefile compiles
ver0040/cpu_parameter_wafe.f90, cpu_cpml4.f90, cpu_eh_field_cpml42.f90 and
cpu_waveform.f90 where they lie, with amdflang, into oracle/_ref/fortran_step
together with the driver oracle/ref_fortran_harness.f90.  The driver runs the
reference main loop's order (ver0040/cpu_main.f90:96-99, 104-140): init_cpml,
media_coeff, then per step read_source_3d (two Hz points), e_field_cpml4,
CPML_E4, h_field_cpml4, CPML_H4, on the module's compile-time grid
121 x 51 x 51 with ncpml 10 (cpu_parameter_wafe.f90:26-29).

Inputs come from exact integer hashes (no libm), so the test regenerates them
with `inputs()`: sigma in [0.5, 5.5) S/m, mu = MU0 * [0.95, 1.05), complex
fields with independent real and imaginary parts.  Recorded:
  * dt, cmax, omega0, MU0, c1, c2, dx as the module computes them;
  * the eight 1-D profiles per axis (init_cpml);
  * after NSTEPS steps, at sample cells (lines through every CPML shell plus
    hashed cells): the twelve coefficient volumes, the six fields and the
    twelve psi volumes (real and imaginary parts);
  * for every one of those volumes its sums over planes of constant x, y and z
    (numpy float64 sums of the whole volume), so every cell takes part.

Usage:  python tests/golden/make_fortran_fixtures.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NX, NY, NZ, NC = 121, 51, 51, 10
NSTEPS = 3
SRC = (61, 26, 20, 69, 26, 20)          # 1-based x, y, z of the two Hz points
F_MU0 = 1.2566370614e-6
N = NX * NY * NZ
PROFILES = ("ESIG", "MSIG", "BE", "BH", "CE", "CH", "KED", "KHD")
COEFFS = ("CA_X", "CA_Y", "CA_Z", "CB_X", "CB_Y", "CB_Z",
          "DA_X", "DA_Y", "DA_Z", "DB_X", "DB_Y", "DB_Z")
FIELDS = ("EX", "EY", "EZ", "HX", "HY", "HZ")
PSIS = ("PSI_EZX", "PSI_EYX", "PSI_EXY", "PSI_EZY", "PSI_EYZ", "PSI_EXZ",
        "PSI_HYX", "PSI_HZX", "PSI_HZY", "PSI_HXY", "PSI_HXZ", "PSI_HYZ")
FIXTURE = os.path.join(HERE, "ref_fortran_step.npz")


def hash_field(a: int, n: int, lo: float = -0.5, scale: float = 1.0) -> np.ndarray:
    """Deterministic inputs from exact integer arithmetic (no libm)."""
    idx = np.arange(n, dtype=np.uint64)
    h = (idx * np.uint64(2654435761) + np.uint64(97 * a + 13)) % np.uint64(2 ** 32)
    return (h.astype(np.float64) / 2.0 ** 32 + lo) * scale


def inputs():
    """sigma, mu and the complex fields (z, y, x arrays, x fastest)."""
    shp = (NZ, NY, NX)
    sig = (0.5 + 5.0 * hash_field(1, N, lo=0.0)).reshape(shp)
    mu = (F_MU0 * (0.95 + 0.1 * hash_field(2, N, lo=0.0))).reshape(shp)
    fields = []
    for q in range(6):
        scale = 1e-3 if q < 3 else 1.0
        re = hash_field(10 + q, N, scale=scale).reshape(shp)
        im = hash_field(20 + q, N, scale=scale).reshape(shp)
        fields.append((re, im))
    return sig, mu, fields


def sample_cells() -> np.ndarray:
    """Flat 0-based cells: x-, y- and z-lines crossing every CPML shell, plus hashed cells."""
    cells = []
    for j, k in ((1, 1), (5, 25), (25, 25), (49, 12)):
        cells += [(k * NY + j) * NX + i for i in range(NX)]
    for i, k in ((3, 30), (60, 25), (117, 8)):
        cells += [(k * NY + j) * NX + i for j in range(NY)]
    for i, j in ((2, 2), (60, 25), (110, 45)):
        cells += [(k * NY + j) * NX + i for k in range(NZ)]
    h = (np.arange(400, dtype=np.uint64) * np.uint64(2246822519) + np.uint64(7)) % np.uint64(N)
    cells += [int(c) for c in h]
    return np.unique(np.asarray(cells, dtype=np.int64))


def axis_sums(v: np.ndarray) -> np.ndarray:
    """Sums over planes of constant x, y and z of a (z, y, x) volume."""
    return np.concatenate([v.sum(axis=(0, 1)), v.sum(axis=(0, 2)), v.sum(axis=(1, 2))])


def run_reference(exe: str):
    sig, mu, fields = inputs()
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            np.array([NSTEPS], dtype=np.int32).tofile(f)
            np.array(SRC, dtype=np.int32).tofile(f)
            sig.tofile(f)
            mu.tofile(f)
            for re, im in fields:
                re.tofile(f)
                im.tofile(f)
        subprocess.run([exe, fin, fout], cwd=tmp, check=True)
        raw = np.fromfile(fout, dtype=np.float64)
    pos = 0

    def take(n):
        nonlocal pos
        v = raw[pos:pos + n]
        pos += n
        return v

    out = {}
    hdr = take(7)
    for q, nm in enumerate(("dt", "cmax", "omega0", "mu0", "c1", "c2", "dx")):
        out[nm] = np.array(hdr[q])
    for a, n in enumerate((NX, NY, NZ)):
        for nm in PROFILES:
            out[f"prof_{a}_{nm}"] = take(n).copy()
    vols = {}
    for nm in COEFFS:
        vols[nm] = take(N).reshape(NZ, NY, NX)
    for nm in FIELDS:
        vols[nm + "_re"] = take(N).reshape(NZ, NY, NX)
        vols[nm + "_im"] = take(N).reshape(NZ, NY, NX)
    for nm in PSIS:
        vols[nm + "_re"] = take(N).reshape(NZ, NY, NX)
        vols[nm + "_im"] = take(N).reshape(NZ, NY, NX)
    assert pos == raw.size, (pos, raw.size)
    return out, vols


def main():
    import oracle
    oracle.build()
    exe = os.path.join(ROOT, "oracle", "_ref", "fortran_step")
    if not os.path.exists(exe):
        raise SystemExit("reference tree or amdflang missing: oracle/_ref/fortran_step was not built")
    out, vols = run_reference(exe)
    cells = sample_cells()
    out["dims"] = np.array([NX, NY, NZ, NC, NSTEPS])
    out["src"] = np.array(SRC)
    out["cells"] = cells
    for nm, v in vols.items():
        out[f"at_{nm}"] = v.reshape(-1)[cells]
        out[f"sum_{nm}"] = axis_sums(v)
    np.savez_compressed(FIXTURE, **out)
    print("wrote", FIXTURE, os.path.getsize(FIXTURE), "bytes")


if __name__ == "__main__":
    main()
