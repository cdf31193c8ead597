#!/usr/bin/env python3
"""Writes tests/golden/fm_torus_reference.npz: the seven input fields the
reference's own example-data generator (create_example_data.py,
create_fishbone_moncrief_torus, plot off) produces for a few small grids.
The CPU tests compare grt_amd.fm_torus with these arrays.

Runs in the build container only (it imports the reference from
/root/reference/src, which no GPU run sees); the committed fixture is data.
"""
import contextlib
import io
import os
import sys
import tempfile

import numpy as np

REF_SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "fm_torus_reference.npz")

# name -> generator keyword arguments
CASES = {
    # the first reference configuration's torus on a coarse grid
    "a09": dict(nr=12, ntheta=8, nphi=4, r_min=1.5, r_max=100.0, mass=1.0, spin=0.9,
                r_in=6.0, r_pressure_max=12.0),
    # mass != 1, retrograde-free high spin, no density floor
    "m2_a094_nofloor": dict(nr=10, ntheta=6, nphi=3, r_min=3.0, r_max=80.0, mass=2.0,
                            spin=1.88, r_in=12.0, r_pressure_max=24.0, rho_floor=0.0),
    # gamma 5/3, an equipotential that does not close before r_max
    "a05_open_g53": dict(nr=9, ntheta=7, nphi=2, r_min=2.0, r_max=40.0, mass=1.0, spin=0.5,
                         r_in=6.0, r_pressure_max=30.0, gamma=5.0 / 3.0),
    # default r_min (1.2 horizon radii), Schwarzschild
    "a0_default_rmin": dict(nr=8, ntheta=5, nphi=2, r_max=50.0, mass=1.0, spin=0.0,
                            r_in=6.0, r_pressure_max=10.0),
}
FIELDS = ("r", "theta", "phi", "rho", "Tgas", "vel", "mag")


def main():
    sys.path.insert(0, REF_SRC)
    import create_example_data as ref
    arrays = {}
    for name, kw in CASES.items():
        with tempfile.TemporaryDirectory() as d, contextlib.redirect_stdout(io.StringIO()):
            ref.create_fishbone_moncrief_torus(output_dir=d, plot=False, **kw)
            for f in FIELDS:
                arrays[f"{name}/{f}"] = np.load(os.path.join(d, f + ".npy"), allow_pickle=False)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(CASES)} cases)")


if __name__ == "__main__":
    main()
