"""CPU checks of the script runner's command parsing (no device calls: the
scripts stop before `run`).  Messages are the reference's
(src/dump_custom.cpp, src/dump.cpp, src/compute_stress_atom.cpp,
src/input.cpp)."""
import os
import subprocess

import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "lammps_amd", "lmp_mdx")

SETUP = """units lj
atom_style atomic
lattice fcc 0.8442
region box block 0 4 0 4 0 4
create_box 1 box
create_atoms 1 box
mass 1 1.0
compute pe all pe/atom
compute s all stress/atom NULL
"""


def run_lines(tmp_path, extra):
    path = os.path.join(str(tmp_path), "in.script")
    with open(path, "w") as fh:
        fh.write(SETUP + extra)
    r = subprocess.run([EXE, "-in", path], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout


@pytest.mark.parametrize("line,message", [
    ("dump 1 all custom 10 d.txt id foo", "Invalid attribute foo in dump custom command"),
    ("dump 1 all custom 10 d.txt", "No dump custom arguments specified"),
    ("dump 1 all custom 10 d.txt id c_nope", "Could not find dump custom compute ID"),
    ("dump 1 all custom 10 d.txt id c_s", "Dump custom compute does not calculate per-atom vector"),
    ("dump 1 all custom 10 d.txt id c_pe[1]", "Dump custom compute does not calculate per-atom array"),
    ("dump 1 all custom 10 d.txt id c_s[7]", "Dump custom compute vector is accessed out-of-range"),
    ("dump 1 all custom 10 d.txt id q", "Dumping an atom property that isn't allocated"),
    ("dump 1 all custom 10 d.txt id x\ndump 1 all custom 10 e.txt id", "Reuse of dump ID"),
    ("dump 1 all custom 10 d.txt id x\ndump_modify 1 format \"%d\"", "Dump_modify format string is too short"),
    ("dump 1 all custom 10 d.txt id x\ndump_modify 1 format \"%d %d\"", "Invalid dump_modify format token %d"),
    ("dump 1 all custom 10 d.txt id x\ndump_modify 1 format \"%s %g\"", "Invalid dump_modify format token %s"),
    ("dump 1 all custom 10 d.txt id x\ndump_modify 1 sort 3", "Dump sort column is invalid"),
    ("dump 1 all custom 10 d.txt id x\ndump_modify 2 sort id", "Could not find dump_modify ID"),
    ("undump 3", "Could not find undump ID"),
    ("compute pe all pe/atom", "Reuse of compute ID"),
    ("compute t all stress/atom thermo_temp", "Could not find compute stress/atom temperature ID"),
    ("dump 1 all custom 10 \"d.txt id x", "Unbalanced quotes in input line"),
])
def test_dump_and_compute_errors(tmp_path, line, message):
    rc, out = run_lines(tmp_path, line + "\n")
    assert rc == 1
    assert out.splitlines()[-1] == "ERROR: " + message


def test_quoted_words_comments_and_undump(tmp_path):
    # a quoted format keeps its spaces and '#'; text after an unquoted '#' is a comment
    rc, out = run_lines(tmp_path, "dump 1 all custom 10 d.txt id x  # a comment \"\n"
                                  "dump_modify 1 format \"%d #%.3f\" sort -2\n"
                                  "undump 1\n"
                                  "dump 1 all custom 5 d.txt id\n")
    assert rc == 0, out
