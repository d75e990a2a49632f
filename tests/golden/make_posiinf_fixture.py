"""Writes tests/golden/model_env_posiinf.dat with the reference's own writer.

Runs ONLY where the reference tree is mounted: oracle/Makefile compiles
satan_i/fwi3d_cpu_posiinf_v1.4.c (a self-contained program) into
oracle/_ref/posiinf, which writes ./data/model_env.dat (and posi4gmt.dat) in its
working directory from the values set in its source: a 31^3 grid of 30 m cells,
1302 steps, omega_0 = 2 pi, fmax_w = 3, Glim = 10.4f, ncpml 10, startF 1,
upperF 20 and seven transmitters and receivers on a line.  The file it writes
is kept as the fixture.

Usage:  python tests/golden/make_posiinf_fixture.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

FIXTURE = os.path.join(HERE, "model_env_posiinf.dat")


def run_posiinf() -> str:
    """Text of the model_env.dat the reference's program writes."""
    exe = os.path.join(ROOT, "oracle", "_ref", "posiinf")
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    with tempfile.TemporaryDirectory() as td:
        os.makedirs(os.path.join(td, "data"))
        subprocess.run([exe], cwd=td, check=True, stdout=subprocess.DEVNULL)
        with open(os.path.join(td, "data", "model_env.dat")) as f:
            return f.read()


def main():
    oracle.build()
    text = run_posiinf()
    with open(FIXTURE, "w") as f:
        f.write(text)
    print(f"wrote {FIXTURE} ({len(text)} bytes)")


if __name__ == "__main__":
    main()
