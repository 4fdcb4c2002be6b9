"""Writes the grid-file fixtures under ``tests/golden/grid_32x16x24``.

A 32 x 16 x 24 grid in the reference's ``x{1,2,3}_grid.in`` format
(``utils/loadGrid.F90:38-209``): uniform x on [0, 2*pi), uniform y on
[0, pi), and z on [0, 1] with mid-plane tanh stretching (beta = 1.5),
``z = 1/2 + atanh((2s-1) tanh(beta)) / (2 beta)``.  Run from the repository
root: ``python tests/golden/make_grid_files.py``.
"""
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from anelastic_solve_amd.grid import tanh_axis, uniform_axis, write_grid_file  # noqa: E402

OUT = os.path.join(HERE, "grid_32x16x24")
SPEC = [("X1", uniform_axis(32, 2 * math.pi), 2 * math.pi, "uniform"),
        ("X2", uniform_axis(16, math.pi), math.pi, "uniform"),
        ("X3", tanh_axis(24, 1.0, 1.5), 1.0, "tanh-midplane-beta1.5")]


def main():
    os.makedirs(OUT, exist_ok=True)
    for d, (label, ax, L, kind) in enumerate(SPEC, start=1):
        write_grid_file(os.path.join(OUT, f"x{d}_grid.in"), ax, L, 0.0, label, kind)


if __name__ == "__main__":
    main()
