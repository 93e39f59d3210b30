"""The LAMMPS-style runner (lmp_mdx: C++ Pair/Neighbor/Fix classes over the
C ABI) on the reference's benchmark commands.

FP64: thermo rows equal the reference logs' printed digits, with the same
neighbour totals and build counts.  Mixed: the reference's own regression
rule, mean-square error of every thermo column < 1e-6
(regress/benchmark.py:25,105-124)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "lammps_amd", "lmp_mdx")
with open(os.path.join(GOLDEN, "goldens.json")) as fh:
    GOLD = json.load(fh)


def run_script(name, extra=None):
    path = os.path.join(GOLDEN, name)
    if extra:
        text = open(path).read().replace("mdx precision double", extra)
        path = os.path.join("/tmp", "mdx_" + name)
        with open(path, "w") as fh:
            fh.write(text)
    r = subprocess.run([EXE, "-in", path], cwd=ROOT, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows, stats = [], {}
    lines = r.stdout.splitlines()
    for k, line in enumerate(lines):
        if line.startswith("Step Temp"):
            j = k + 1
            while j < len(lines) and not lines[j].startswith("Loop time"):
                rows.append(lines[j].split())
                j += 1
        for key, tag in (("total", "Total # of neighbors = "), ("builds", "Neighbor list builds = "),
                         ("dangerous", "Dangerous builds = ")):
            if line.startswith(tag):
                stats[key] = int(line[len(tag):])
    return rows, stats


@pytest.mark.parametrize("name,gold", [("lj_bench.in", "lj_bench_32k"), ("eam_bench.in", "eam_bench_32k")])
def test_fp64_runner_reproduces_reference_log(name, gold):
    rows, stats = run_script(name)
    g = GOLD[gold]
    assert rows == g["thermo_text"]
    assert stats["total"] == g["total_neighbors"]
    assert stats["builds"] == g["builds"]
    assert stats["dangerous"] == g["dangerous"]


@pytest.mark.parametrize("name,gold", [("lj_bench.in", "lj_bench_32k"), ("eam_bench.in", "eam_bench_32k")])
def test_mixed_runner_meets_reference_regression_rule(name, gold):
    rows, _ = run_script(name, extra="mdx precision mixed")
    g = GOLD[gold]
    got = np.array([[float(v) for v in r] for r in rows])
    ref = np.array([[float(v) for v in r] for r in g["thermo_text"]])
    assert got.shape == ref.shape
    assert np.array_equal(got[:, 0], ref[:, 0])
    for col in range(1, ref.shape[1]):
        mse = np.mean((got[:, col] - ref[:, col]) ** 2)
        scale = max(1.0, np.max(np.abs(ref[:, col])) ** 2)
        assert mse / scale < 1e-6, (col, mse)
