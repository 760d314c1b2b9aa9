"""Replica generator of the ensemble engine (host logic, no GPU): sizes, densities and box
shapes of the 256-replica ensemble."""
import numpy as np
import pytest

try:
    from mrmd_amd.ensemble import adress_box, cubic_box, ensemble_config
except ImportError:  # the library must be built for the package to import
    pytest.skip("libmrmd_hip.so not built", allow_module_level=True)


def test_ensemble_config_shapes():
    reps, pos, vel = ensemble_config(n_rep=256, seed=2024)
    assert len(reps) == 256
    n = np.array([r["num_atoms"] for r in reps])
    assert n.min() >= 500 and n.max() <= 50000
    assert 5e5 < n.sum() < 6e6
    assert sum(r["adress"] for r in reps) == 128
    for r, p, v in zip(reps, pos, vel):
        box = np.asarray(r["box"])
        assert (box >= 5.6).all()
        assert p.shape == (r["num_atoms"], 3) and (p >= 0).all() and (p < box).all()
        assert np.abs(v.mean(axis=0)).max() < 1e-12
        rho = r["num_atoms"] / box.prod()
        assert 0.29 <= rho <= 0.91
        if r["adress"]:
            assert box[0] >= box[1] == box[2]
            # the hybrid region ends well inside the box along x
            assert 0.5 * r["atomistic_diameter"] + r["hybrid_diameter"] < 0.5 * box[0] - 2.8 or box[0] < 30


def test_box_helpers():
    assert adress_box(160000, 0.5) == pytest.approx([200 * (160000 / 0.5 / 320000) ** (1 / 3) * 1.0,
                                                     40 * (160000 / 0.5 / 320000) ** (1 / 3),
                                                     40 * (160000 / 0.5 / 320000) ** (1 / 3)])
    assert cubic_box(4000, 0.8442)[0] == pytest.approx(16.795961913825074)
    assert min(adress_box(500, 0.9)) >= 5.6
