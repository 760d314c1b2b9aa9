"""The reference's own test programs, restated against the drop-in C++ headers
(mrmd_amd/facade/include: same include paths, class names and call sequences), run on
the GPU.  Built by __graft_entry__.build() into mrmd_amd/facade/bin/."""
import gzip
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mrmd_amd", "facade", "bin")


def _run(name, *args, cwd=None):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), f"{exe} missing: run __graft_entry__.build()"
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120, cwd=cwd)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    return r.stdout


def test_espp_comparison_dropin(tmp_path):
    txt = tmp_path / "positions.txt"
    with gzip.open(os.path.join(ROOT, "tests", "golden", "espp_positions.txt.gz"), "rb") as f, open(txt, "wb") as g:
        shutil.copyfileobj(f, g)
    out = _run("espp_comparison", str(txt))
    assert "ghosts 22104 pairs 1310403" in out


def test_unit_actions_dropin(tmp_path):
    _run("unit_actions", cwd=tmp_path)  # the thermodynamic-force files go to tmp_path


def test_nve_example_on_reference_input(tmp_path):
    """examples/02_LennardJones_NVE/02_LennardJones_NVE.cpp:87-216 restated against the
    drop-in headers (tests/cpp/nve_gro.cpp) on the reference's own input, the 4,096-atom
    argon phase point lennardJonesNVT_final.gro: restoreGRO, velocity Verlet at dt 0.002,
    the unshifted LennardJones(rc 2.5, sigma, epsilon, r_cap 0.7), skin 0.1, 1,000 steps.
    Against the oracle's restatement of the same loop from the same file: the rebuild count,
    the early energy series (before chaos separates the trajectories) and the NVE drift."""
    import numpy as np

    import pyoracle as O
    with gzip.open(os.path.join(ROOT, "tests", "golden", "lennardJonesNVT_final.gro.gz"), "rb") as f, \
            open(tmp_path / "in.gro", "wb") as g:
        shutil.copyfileobj(f, g)
    nsteps, every = 1000, 10
    out = _run("nve_gro", str(tmp_path / "in.gro"), str(nsteps), str(every))
    rec = [ln.split() for ln in out.splitlines() if ln.startswith("step ")]
    e0 = np.array([float(r[3]) for r in rec])
    ek = np.array([float(r[5]) for r in rec])
    reb_g = int(out.split("rebuilds ")[-1].split()[0])
    pos, vel, edges = O.read_gro(str(tmp_path / "in.gro"))
    n = pos.shape[0]
    a = O.Atoms.from_arrays(pos, vel=vel)
    box = O.make_box([0, 0, 0], edges, 2.6)
    prm = O.md_params(dt=0.002, skin=0.1, rc=2.5, cap=0.7, max_neigh=60, nsteps=nsteps, energy_every=every,
                      shifted=0, langevin=0, permute=0)
    ec = np.zeros((nsteps // every, 3))
    O.lib().or_set_threads(4)
    reb_c = O.lib().or_run_md(a.h, O.C.byref(box), O.C.byref(prm), O.ptr(ec))
    assert len(e0) == nsteps // every
    assert reb_g == reb_c and reb_c > 10
    np.testing.assert_allclose(e0[:20], ec[:20, 0] / n, rtol=1e-9)
    np.testing.assert_allclose(ek[:20], ec[:20, 1] / n, rtol=1e-9)
    t = np.arange(len(e0)) * every * 0.002
    tot_g, tot_c = e0 + ek, (ec[:, 0] + ec[:, 1]) / n
    slope_g, slope_c = np.polyfit(t, tot_g, 1)[0], np.polyfit(t, tot_c, 1)[0]
    assert abs(tot_g - tot_g[0]).max() < 5e-3
    assert abs(slope_g - slope_c) < 2e-5 + 0.5 * abs(slope_c)
