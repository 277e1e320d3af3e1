"""The C ABI without a GPU: the library loads, exports every entry point the header
declares, and the ctypes structs have the header's layout (sizes and offsets measured
by compiling the header with gcc)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from sigmasph_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sigmasph.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sph_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert set(declared_functions()) == set(N.EXPORTS)


def test_library_exports_every_symbol():
    L = N.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert b"gfx950" in L.sph_version()


def test_null_arguments_rejected():
    L = N.lib()
    assert L.sph_ctx_create(None, None) == N.SPH_EINVAL
    assert L.sph_ctx_destroy(None) == N.SPH_OK
    assert L.sph_sort(None, None, None) == N.SPH_EINVAL
    assert L.sph_fluid_step(None, None, None, None) == N.SPH_EINVAL
    assert L.sph_solid_step(None, None, None, None, None, None) == N.SPH_EINVAL
    assert L.sph_solid_sort(None, None, None, None) == N.SPH_EINVAL


STRUCTS = {
    "sph_dims": (N.SphDims, ["n", "precision", "flags", "h", "cell", "origin", "period", "box_lo",
                             "box_n", "hash_bits", "dx0"]),
    "sph_sorted": (N.SphSorted, ["pid", "obj", "x", "x0", "u", "v", "shift", "V", "V_lo", "m", "p"]),
    "sph_state": (N.SphState, ["buf", "cur", "acc", "lam", "L6", "lam_grad", "div", "de_visc", "V_new",
                               "rho", "rho_new", "scal"]),
    "sph_fluid_params": (N.SphFluidParams, [f[0] for f in N.SphFluidParams._fields_]),
    "sph_material": (N.SphMaterial, [f[0] for f in N.SphMaterial._fields_]),
    "sph_solid_params": (N.SphSolidParams, [f[0] for f in N.SphSolidParams._fields_]),
    "sph_solid": (N.SphSolid, [f[0] for f in N.SphSolid._fields_]),
}


def test_struct_layout_matches_header():
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){"]
    for cname, (_, fields) in STRUCTS.items():
        src.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in fields:
            src.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    seen = {}
    for line in out.splitlines():
        parts = line.split()
        seen[(parts[0], parts[1])] = int(parts[2])
    for cname, (cls, fields) in STRUCTS.items():
        assert ctypes.sizeof(cls) == seen[(cname, "size")], cname
        for f in fields:
            assert getattr(cls, f).offset == seen[(cname, f)], (cname, f)
