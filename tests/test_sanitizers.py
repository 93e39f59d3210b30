"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code.

  * the FP64 oracle (oracle/mdref.c, oracle/fitref.c) and the library's
    host set-up code (lammps_amd/csrc/host_setup.cpp: lattice, velocity
    create, funcfl/setfl readers, the synthetic funcfl writer), driven by
    tests/sanitize/host_checks.cpp;
  * the script runner's command parser (lammps_amd/host/lmp_mdx.cpp,
    plugin.cpp, output.cpp), run on the scripts of test_runner_host.py
    (they stop before any device call).

Both are built with g++/gcc -fsanitize=address,undefined
-fno-sanitize-recover=all, so any report ends the process with a non-zero
status.  CPU only: no GPU code is instrumented (hipcc is not used here).
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="needs gcc/g++")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, "%s\n%s\n%s" % (" ".join(cmd), r.stdout[-4000:], r.stderr[-4000:])
    return r


@pytest.fixture(scope="module")
def build_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("sanitize"))


def test_oracle_and_host_setup_under_asan_ubsan(build_dir):
    import sys
    sys.path.insert(0, HERE)
    from setfl_util import binary_cu_like
    objs = []
    for src in ("mdref.c", "fitref.c"):
        obj = os.path.join(build_dir, src + ".o")
        _run(["gcc", "-std=c99", "-ffp-contract=off"] + SAN + ["-c", os.path.join(ROOT, "oracle", src), "-o", obj])
        objs.append(obj)
    exe = os.path.join(build_dir, "host_checks")
    _run(["g++", "-std=c++17"] + SAN + ["-I" + os.path.join(ROOT, "include"),
                                        os.path.join(ROOT, "lammps_amd", "csrc", "host_setup.cpp"),
                                        os.path.join(HERE, "sanitize", "host_checks.cpp")] + objs +
         ["-o", exe, "-lm"])
    alloy = os.path.join(build_dir, "ab.eam.alloy")
    fs = os.path.join(build_dir, "ab.eam.fs")
    binary_cu_like(alloy, fs=False)
    binary_cu_like(fs, fs=True)
    r = _run([exe, build_dir, alloy, fs], env=ENV, timeout=600)
    assert "host_checks ok" in r.stdout


RUNNER_SETUP = """units lj
atom_style atomic
lattice fcc 0.8442
region box block 0 4 0 4 0 4
create_box 1 box
create_atoms 1 box
mass 1 1.0
compute pe all pe/atom
compute s all stress/atom NULL
"""


@pytest.mark.parametrize("extra,want_rc", [
    ("dump 1 all custom 10 d.txt id foo\n", 1),
    ("dump 1 all custom 10 d.txt id x\ndump_modify 1 format \"%d #%.3f\" sort -2\nundump 1\n", 0),
    ("dump 1 all custom 10 \"d.txt id x\n", 1),
    ("dump 1 all custom 10 d.txt id c_s[7]\n", 1),
    ("pair_style lj/cut 2.5\npair_coeff 1 1 1.0 1.0 2.5\nneighbor 0.3 bin\nneigh_modify every 20 delay 0 check no\n"
     "velocity all create 1.44 87287 loop geom\nfix 1 all nve\nthermo 50\ntimestep 0.005\n", 0),
])
def test_runner_parser_under_asan_ubsan(build_dir, tmp_path, extra, want_rc):
    exe = os.path.join(build_dir, "lmp_mdx_asan")
    if not os.path.exists(exe):
        host = os.path.join(ROOT, "lammps_amd", "host")
        lib = os.path.join(ROOT, "lammps_amd", "libmdx.so")
        if not os.path.exists(lib):
            pytest.skip("libmdx.so not built")
        _run(["g++", "-std=c++17"] + SAN + ["-I" + os.path.join(ROOT, "include"),
                                            os.path.join(host, "plugin.cpp"), os.path.join(host, "output.cpp"),
                                            os.path.join(host, "lmp_mdx.cpp"), "-o", exe,
                                            "-L" + os.path.dirname(lib), "-l:libmdx.so",
                                            "-Wl,-rpath," + os.path.dirname(lib)])
    path = os.path.join(str(tmp_path), "in.script")
    with open(path, "w") as fh:
        fh.write(RUNNER_SETUP + extra)
    r = subprocess.run([exe, "-in", path], capture_output=True, text=True, timeout=120, env=ENV, cwd=str(tmp_path))
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == want_rc, (r.stdout[-2000:], r.stderr[-2000:])
