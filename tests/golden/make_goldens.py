"""Extract the golden thermo rows and neighbour statistics from the reference's
own benchmark/example logs into tests/golden/goldens.json.

Runs only where the reference tree is present (it reads the logs as text);
the committed goldens.json is what the tests use.
Sources: bench/log.1Feb14.lj.fixed.linux.1, bench/log.1Feb14.eam.fixed.linux.1,
examples/melt/log.melt.1Feb14.linux.1 (thermo rows and the 'Total # of
neighbors' / 'Neighbor list builds' / 'Nghost' lines).
"""
import json
import os
import re
import sys

REF = os.environ.get("LAMMPS_REFERENCE", "/root/reference")
LOGS = {
    "lj_bench_32k": "bench/log.1Feb14.lj.fixed.linux.1",
    "eam_bench_32k": "bench/log.1Feb14.eam.fixed.linux.1",
    "melt_4k": "examples/melt/log.melt.1Feb14.linux.1",
}


def parse(path):
    rows, stats = [], {}
    with open(path) as fh:
        lines = fh.read().splitlines()
    i = 0
    while i < len(lines):
        if lines[i].startswith("Step Temp"):
            i += 1
            while i < len(lines) and not lines[i].startswith("Loop time"):
                rows.append(lines[i].split())
                i += 1
        m = re.match(r"Total # of neighbors = (\d+)", lines[i]) if i < len(lines) else None
        if m:
            stats["total_neighbors"] = int(m.group(1))
        m = re.match(r"Neighbor list builds = (\d+)", lines[i]) if i < len(lines) else None
        if m:
            stats["builds"] = int(m.group(1))
        m = re.match(r"Dangerous builds = (\d+)", lines[i]) if i < len(lines) else None
        if m:
            stats["dangerous"] = int(m.group(1))
        m = re.match(r"Nghost:\s+(\d+)", lines[i]) if i < len(lines) else None
        if m:
            stats["nghost"] = int(m.group(1))
        m = re.match(r"Loop time of ([0-9.eE+-]+) on (\d+) procs for (\d+) steps with (\d+) atoms",
                     lines[i]) if i < len(lines) else None
        if m:
            stats["loop_time"] = float(m.group(1))
            stats["steps"] = int(m.group(3))
            stats["atoms"] = int(m.group(4))
        i += 1
    return {"thermo_text": rows, **stats, "source": os.path.relpath(path, REF)}


def main():
    out = {k: parse(os.path.join(REF, v)) for k, v in LOGS.items()}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "goldens.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    sys.exit(main())
