"""The reference-side adapter styles (lammps_amd/lammps_styles: pair_style
lj/cut/mdx and eam/mdx, run_style verlet/mdx) compile against
the reference's own headers (src/pair.h, src/pair_lj_cut.h,
src/MANYBODY/pair_eam.h, src/verlet.h, the MPI stubs in src/STUBS): a syntax-only check
that the C ABI drops under the reference's Pair plugin surface.  Nothing of
the reference is built or run.  Skipped where the reference tree is absent."""
import os
import subprocess

import pytest

from conftest import ROOT

REF = "/root/reference/src"


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
@pytest.mark.parametrize("style", ["pair_lj_cut_mdx", "pair_eam_mdx", "verlet_mdx"])
def test_adapter_style_compiles_against_the_reference_headers(style):
    src = os.path.join(ROOT, "lammps_amd", "lammps_styles", style + ".cpp")
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-Werror", "-I" + REF, "-I" + os.path.join(REF, "STUBS"),
           "-I" + os.path.join(REF, "MANYBODY"), "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "lammps_amd", "lammps_styles"), src]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
