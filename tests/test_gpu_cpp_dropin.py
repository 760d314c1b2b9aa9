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
