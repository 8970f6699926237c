"""The Fortran side of the boundary (horses3d_amd/fortran): the ISO_C_BINDING module compiles
with the image's Fortran compiler and drives the C ABI from Fortran (descriptor validation and
error paths, no GPU needed); the HORSES3D adapter compiles against the modules of a HORSES3D
build when one is present (the oracle's reference build)."""
import os
import shutil
import subprocess

import pytest

from conftest import REPO

FC = shutil.which("amdflang") or shutil.which("flang")
FDIR = os.path.join(REPO, "horses3d_amd", "fortran")
LIB = os.path.join(REPO, "horses3d_amd", "libhx3d.so")
REF_MOD = os.path.join(REPO, "oracle", "_ref", "mod")
REF_INC = "/root/reference/Solver/src/libs/foundation"

pytestmark = pytest.mark.skipif(FC is None, reason="no Fortran compiler in this image")


def _run(cmd, cwd, env=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_binding_drives_the_c_abi(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("engine library not built")
    _run([FC, "-c", "-module-dir", str(tmp_path), os.path.join(FDIR, "hx3d_binding.f90"),
          "-o", str(tmp_path / "hx3d_binding.o")], tmp_path)
    exe = str(tmp_path / "abi_check")
    _run([FC, "-I", str(tmp_path), os.path.join(REPO, "tests", "fortran", "abi_check.f90"),
          str(tmp_path / "hx3d_binding.o"), LIB, "-Wl,-rpath," + os.path.dirname(LIB), "-o", exe], tmp_path)
    from horses3d_amd import specmesh
    from test_control import BASE, PERIODIC
    specmesh.write_specmesh(str(tmp_path / "box.mesh"), 2, 0.0, None)
    (tmp_path / "case.control").write_text(BASE + PERIODIC)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    out = _run([exe], tmp_path, env=env)
    assert "abi_check ok" in out and "control ok" in out
    assert "K <= 0" in out


def test_adapter_compiles_against_horses3d_modules(tmp_path):
    if not (os.path.isdir(REF_MOD) and os.path.isdir(REF_INC)):
        pytest.skip("no HORSES3D build modules here")
    _run([FC, "-c", "-module-dir", str(tmp_path), os.path.join(FDIR, "hx3d_binding.f90"),
          "-o", str(tmp_path / "b.o")], tmp_path)
    _run([FC, "-cpp", "-DNAVIERSTOKES", "-I", REF_INC, "-I", REF_MOD, "-I", str(tmp_path),
          "-module-dir", str(tmp_path), "-c", os.path.join(FDIR, "hx3d_horses3d.f90"),
          "-o", str(tmp_path / "a.o")], tmp_path)
