"""The oracle's celestial checkerboard (imaging.py:27-50) against what the
sphere's geometry fixes: a direction at the centre of polar cell i and
azimuthal cell j is light exactly when i + j is odd, for any cell count; the
poles and the coordinate axes sit on cell boundaries whose indices follow
from acos and atan2's exact values there; |z| > 1 clips to the pole."""

import numpy as np
import pytest

from oracle import bh_oracle as orc

LIGHT = (104, 128, 178)  # imaging.py:22
DARK = (24, 30, 52)      # imaging.py:21


def cell_centres(cells):
    """Unit directions at the centres of the cells x 2 cells grid, with the
    expected light/dark flag of each."""
    i, j = np.meshgrid(np.arange(cells), np.arange(2 * cells), indexing="ij")
    theta = (i + 0.5) * np.pi / cells
    phi = -np.pi + (j + 0.5) * np.pi / cells
    d = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)
    return d.reshape(-1, 3), ((i + j) % 2 == 1).reshape(-1)


@pytest.mark.parametrize("cells", [1, 2, 5, 12, 31])
def test_cell_centres(cells):
    d, light = cell_centres(cells)
    rgb = orc.celestial_checkerboard(d, cells)
    assert rgb.dtype == np.uint8 and rgb.shape == d.shape
    np.testing.assert_array_equal(rgb, np.where(light[:, None], LIGHT, DARK))
    # every cell is hit once: both colours cover half the sphere's cells
    assert light.sum() == cells * cells


def test_axes_and_poles():
    # +z: theta 0, phi atan2(0, 0) = 0 -> cells (0, 12); -z: theta pi -> (12, 12);
    # +x: theta pi/2 -> (6, 12); -x: phi pi -> (6, 24); -x below the x axis
    # (y = -0.0): phi -pi -> (6, 0); +y: phi pi/2 -> (6, 18); -y: (6, 6)
    d = np.array([[0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [-1, -0.0, 0], [0, 1, 0],
                  [0, -1, 0], [0, 0, 7.5], [0, 0, -3.0]], dtype=np.float64)
    rgb = orc.celestial_checkerboard(d, 12)
    np.testing.assert_array_equal(rgb, [DARK] * 9)
    # with 5 cells: +x -> (floor 2.5 = 2, 5) odd; +y -> (2, floor 7.5 = 7) odd
    rgb5 = orc.celestial_checkerboard(d[[2, 5, 0]], 5)
    np.testing.assert_array_equal(rgb5, [LIGHT, LIGHT, LIGHT])


def test_neighbouring_cells_alternate():
    """Stepping across one cell boundary in either angle flips the colour."""
    cells = 12
    d, light = cell_centres(cells)
    rgb = orc.celestial_checkerboard(d, cells).reshape(cells, 2 * cells, 3)
    lit = (rgb == LIGHT).all(-1)
    assert (lit[1:, :] != lit[:-1, :]).all()
    assert (lit[:, 1:] != lit[:, :-1]).all()
