"""The gfx950 library loads and exports every entry point include/mrmd_hip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mrmd_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(mrmd_\w+)\s*\(", text))


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) > 60
    for core in ["mrmd_lj_apply", "mrmd_verlet_build", "mrmd_halo_create", "mrmd_vv_pre",
                 "mrmd_langevin_pre", "mrmd_ljig_run", "mrmd_tf_update", "mrmd_lj_adjoint"]:
        assert core in names


def test_library_exports_every_declared_symbol():
    lib = os.path.join(ROOT, "mrmd_amd", "libmrmd_hip.so")
    assert os.path.exists(lib), "build the library first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mrmd_\w+)", out))
    missing = _declared() - exported
    assert not missing, missing
    ctypes.CDLL(lib)


def test_python_binding_binds_all():
    import mrmd_amd.capi as capi
    assert set(capi.DECLS) == _declared()


def test_library_targets_gfx950_only():
    lib = os.path.join(ROOT, "mrmd_amd", "libmrmd_hip.so")
    blob = open(lib, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    for other in (b"--gfx942", b"--gfx90a", b"--gfx1100"):
        assert other not in blob
