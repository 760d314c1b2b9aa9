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


def test_checkpoint_frame_round_trip(tmp_path):
    """checkpoint container: H5MD member paths, every dtype and value back exactly; files
    of another format are refused"""
    import pytest

    from mrmd_amd import checkpoint
    g = np.random.default_rng(3)
    frame = {"format": np.array(checkpoint.FORMAT), "h5md/author/name": np.array("someone"),
             "particles/atoms/position/value": g.random((17, 3)),
             "particles/atoms/type/value": g.integers(0, 3, 17, dtype=np.int64),
             "particles/atoms/box/boundary": np.array(["periodic"] * 3),
             "langevin/seed": np.array(2**63 + 12345, np.uint64),
             "integrator/step": np.array(1234567, np.int64)}
    p = tmp_path / "c.npz"
    checkpoint.write_frame(str(p), frame)
    back = checkpoint.read_frame(str(p))
    assert set(back) == set(frame)
    for k, v in frame.items():
        assert back[k].dtype == v.dtype and np.array_equal(back[k], v), k
    assert int(back["langevin/seed"]) == 2**63 + 12345
    other = tmp_path / "o.npz"
    np.savez(str(other), x=np.zeros(2))
    with pytest.raises(ValueError):
        checkpoint.read_frame(str(other))
