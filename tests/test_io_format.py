"""Text format of the thermodynamic-force table (DumpProfile.cpp:26-36: values separated
by one space, default C++ stream formatting of doubles = six significant digits)."""
import numpy as np

from mrmd_amd.io import dump_thermo_force


class _Table:
    def __init__(self, grid, force):
        self._grid, self._force = np.asarray(grid), np.asarray(force)

    def grid(self):
        return self._grid

    def get_force(self):
        return self._force


def test_dump_format(tmp_path):
    t = _Table([1.015, 1.045, 1.075], [[1.0, 2.0], [2.0, 4.0], [1.0 / 3.0, 1e-7]])
    p = tmp_path / "all.txt"
    dump_thermo_force(str(p), t)
    assert p.read_text() == "1.015 1.045 1.075\n1 2 0.333333\n2 4 1e-07\n"
    p1 = tmp_path / "one.txt"
    dump_thermo_force(str(p1), t, 1)
    assert p1.read_text() == "1.015 1.045 1.075\n2 4 1e-07\n"
