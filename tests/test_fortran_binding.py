"""The Fortran ISO_C_BINDING module (bmad-ecosystem_amd/fortran) compiles with
the image's flang, and its interoperable derived types have the sizes the C
library reports -- the layout contract a Bmad program binds to."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
LIBDIR = os.path.join(ROOT, "bmad-ecosystem_amd", "bmx")

PROG = """program t
use bmx_c_binding
implicit none
type(bmx_ele_rec) :: e
type(bmx_particles) :: p
type(bmx_stats) :: s
type(bmx_grid) :: g
type(bmx_config) :: c
type(bmx_schedule) :: sc
type(bmx_rad_map) :: rm
type(bmx_radiation) :: ra
print '(i0,1x,i0)', c_sizeof(e), bmx_sizeof(0)
print '(i0,1x,i0)', c_sizeof(p), bmx_sizeof(1)
print '(i0,1x,i0)', c_sizeof(s), bmx_sizeof(2)
print '(i0,1x,i0)', c_sizeof(g), bmx_sizeof(3)
print '(i0,1x,i0)', c_sizeof(c), bmx_sizeof(4)
print '(i0,1x,i0)', c_sizeof(sc), bmx_sizeof(5)
print '(i0,1x,i0)', c_sizeof(rm), bmx_sizeof(6)
print '(i0,1x,i0)', c_sizeof(ra), bmx_sizeof(7)
print '(i0)', bmx_abi_version()
end program
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler in this image")
def test_fortran_binding_layout(tmp_path):
    src = os.path.join(ROOT, "bmad-ecosystem_amd", "fortran", "bmx_c_binding.f90")
    (tmp_path / "t.f90").write_text(PROG)
    subprocess.check_call([FLANG, "-c", src, "-o", "bmx_c_binding.o"], cwd=tmp_path)
    subprocess.check_call([FLANG, "t.f90", "bmx_c_binding.o", f"-L{LIBDIR}", "-lbmx", f"-Wl,-rpath,{LIBDIR}", "-o", "t"],
                          cwd=tmp_path)
    out = subprocess.check_output([str(tmp_path / "t")], cwd=tmp_path, text=True).split("\n")
    for line in out[:8]:
        a, b = line.split()
        assert a == b, out
    assert out[8].strip() == "4"
