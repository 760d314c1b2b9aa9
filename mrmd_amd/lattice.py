"""Initial configurations shared by the benchmarks and the parity tests.

The reference has no lattice or velocity generator (it restores GRO/TXT files or draws
uniform positions with a Kokkos RNG, util/simulationSetup.cpp:22-59); these are the
standard constructions, seeded through NumPy's Philox bit generator so that identical
arrays go to the device and to any CPU comparison.
"""
import numpy as np


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def fcc(cells, density):
    """FCC lattice of cells^3 unit cells x 4 atoms at number density `density`;
    returns positions (N x 3) and the cubic box edge L = (N / density)^(1/3)"""
    n = 4 * cells ** 3
    L = (n / density) ** (1.0 / 3.0)
    a = L / cells
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    idx = np.stack(np.meshgrid(np.arange(cells), np.arange(cells), np.arange(cells), indexing="ij"),
                   axis=-1).reshape(-1, 3)
    pos = (idx[:, None, :] + basis[None, :, :] + 0.25) * a
    return pos.reshape(-1, 3), L


def maxwell_boltzmann(n, temperature, seed, mass=1.0):
    """velocities ~ N(0, T/m) per component with the centre-of-mass momentum removed"""
    v = rng(seed).standard_normal((n, 3)) * np.sqrt(temperature / mass)
    v -= v.mean(axis=0)
    return v


def uniform(n, lo, hi, seed):
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    return lo + rng(seed).random((n, 3)) * (hi - lo)
