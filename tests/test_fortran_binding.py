"""The Fortran side of the boundary: include/ffhip_mod.f, the ISO_C_BINDING module a
FINFLO maintainer adds, compiles under the reference's fixed-form rules (72 columns,
-fdefault-real-8 as in the reference's gfortran build, Makefile:73), its BIND(C) derived
types have the sizes of the C structs of include/ffhip.h, and a program that takes the
address of every bound entry point links against libffhip.so and calls one of them (no GPU
is touched)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/bin/amdflang"
MOD = os.path.join(ROOT, "include", "ffhip_mod.f")
LIBDIR = os.path.join(ROOT, "finflo_amd")

pytestmark = pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")


def bound_functions():
    src = open(MOD).read()
    return re.findall(r"FUNCTION\s+(FF_[A-Z_0-9]+)\s*\(", src)


def test_module_is_fixed_form_72_columns():
    for n, line in enumerate(open(MOD), start=1):
        assert len(line.rstrip("\n")) <= 72, f"ffhip_mod.f:{n} is {len(line)} columns"
        assert "\t" not in line


def test_bound_functions_are_declared_in_the_header():
    hdr = open(os.path.join(ROOT, "include", "ffhip.h")).read()
    for name in bound_functions():
        assert re.search(r"\b" + name.lower() + r"\s*\(", hdr), name


def test_module_compiles_links_and_matches_the_c_layouts(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libffhip.so")):
        pytest.skip("libffhip.so not built")
    names = bound_functions()
    assert len(names) >= 29
    refs = "\n".join(f"      P = C_FUNLOC({n})\n      IF (.NOT. C_ASSOCIATED(P)) STOP 2"
                     for n in names)
    prog = f"""      PROGRAM SIZES
      USE FFHIP
      IMPLICIT NONE
      TYPE(FF_PARAMS) :: PRM
      TYPE(FF_WALL_PARAMS) :: WP
      TYPE(FF_WALL_PATCH) :: WPA
      TYPE(FF_BLOCK) :: BLK
      TYPE(FF_BC_PATCH) :: BCP
      TYPE(FF_CON_PATCH) :: CNP
      TYPE(C_FUNPTR) :: P
      CHARACTER(KIND=C_CHAR), POINTER :: MSG(:)
{refs}
      CALL C_F_POINTER(FF_LAST_ERROR(C_NULL_PTR), MSG, [12])
      WRITE(*,'(6I8)') C_SIZEOF(PRM), C_SIZEOF(WP), C_SIZEOF(WPA),
     +  C_SIZEOF(BLK), C_SIZEOF(BCP), C_SIZEOF(CNP)
      WRITE(*,'(12A)') MSG
      END PROGRAM
"""
    (tmp_path / "sizes.f").write_text(prog)
    shutil.copy(MOD, tmp_path / "ffhip_mod.f")
    flags = ["-O2", "-fdefault-real-8", "-w"]
    subprocess.check_call([FC, "-c", *flags, "ffhip_mod.f"], cwd=tmp_path)
    subprocess.check_call([FC, *flags, "sizes.f", "ffhip_mod.o", "-o", "sizes",
                           f"-L{LIBDIR}", "-lffhip", f"-Wl,-rpath,{LIBDIR}"], cwd=tmp_path)
    out = subprocess.check_output([str(tmp_path / "sizes")], cwd=tmp_path, text=True).split("\n")
    fsz = [int(v) for v in out[0].split()]
    assert out[1].strip() == "null context"
    csrc = ('#include <stdio.h>\n#include "ffhip.h"\nint main(void){printf("%zu %zu %zu %zu %zu %zu\\n",'
            'sizeof(ff_params), sizeof(ff_wall_params), sizeof(ff_wall_patch), sizeof(ff_block),'
            'sizeof(ff_bc_patch), sizeof(ff_con_patch));'
            'return 0;}\n')
    (tmp_path / "sizes.c").write_text(csrc)
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "sizes.c", "-o", "csizes"],
                          cwd=tmp_path)
    csz = [int(v) for v in subprocess.check_output([str(tmp_path / "csizes")], text=True).split()]
    assert fsz == csz, (fsz, csz)
