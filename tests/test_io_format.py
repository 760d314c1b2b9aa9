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


def _gro_fixture(tmp_path):
    import gzip
    import os
    import shutil
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lennardJonesNVT_final.gro.gz")
    dst = tmp_path / "lennardJonesNVT_final.gro"
    with gzip.open(src, "rb") as f, open(dst, "wb") as g:
        shutil.copyfileobj(f, g)
    return dst


def test_read_gro_reference_input(tmp_path):
    """io::restoreGRO's reading (RestoreGRO.cpp:24-147) of the reference's NVE input
    (examples/02_LennardJones_NVE/lennardJonesNVT_final.gro: 4,096 argon atoms with
    velocities in a 20^3 box): the library's sscanf-style scan equals the oracle's
    fixed-column read bit for bit, and the first and last atoms are the file's"""
    import pyoracle as O

    from mrmd_amd.io import read_gro
    path = _gro_fixture(tmp_path)
    pos, vel, edges = read_gro(str(path))
    opos, ovel, oedges = O.read_gro(str(path))
    assert pos.shape == (4096, 3) and np.array_equal(edges, [20.0, 20.0, 20.0])
    assert np.array_equal(pos, opos) and np.array_equal(vel, ovel) and np.array_equal(edges, oedges)
    assert pos[0].tolist() == [14.587, 5.263, 17.966] and vel[0].tolist() == [-1.0355, 0.7038, -0.6883]
    assert pos[-1].tolist() == [4.579, 8.706, 3.082] and vel[-1].tolist() == [1.3601, -1.2157, -0.2029]


def test_read_gro_without_velocities_and_invalid_lines(tmp_path):
    """a line of 7 fields gives velocity 0; a line of another field count is refused, as
    restoreGRO exits on it"""
    import pytest

    from mrmd_amd.io import read_gro
    p = tmp_path / "a.gro"
    p.write_text("title\n2\n    1Argon   Ar    1   1.000   2.000   3.000  0.5000 -0.2500  0.1250\n"
                 "    2Argon   Ar    2   4.000   5.000   6.000\n   10.0 11.0 12.0\n")
    pos, vel, edges = read_gro(str(p))
    assert pos.tolist() == [[1, 2, 3], [4, 5, 6]] and vel.tolist() == [[0.5, -0.25, 0.125], [0, 0, 0]]
    assert edges.tolist() == [10, 11, 12]
    bad = tmp_path / "b.gro"
    bad.write_text("title\n1\n    1Argon   Ar    1   1.000   2.000\n   1 1 1\n")
    with pytest.raises(ValueError):
        read_gro(str(bad))
