"""Throughput of the short-range MD step on one MI355X, in Matom-steps/s.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 50] [--no-cpu-baseline]

Headline workload (the `value`): lj/cut melt, fcc 80^3 = 2,048,000 atoms,
rho* = 0.8442, rc = 2.5, skin 0.3, neigh_modify every 20 delay 0 check yes,
Gaussian velocities at T* = 1.44 (seed 87287, loop geom), NVE dt = 0.005,
mixed precision (FP64 positions/distances/accumulation, FP32 pair
arithmetic).  The same line reports the FP64 device path on the same
workload, the EAM 256,000-atom case (synthetic funcfl Cu-like table, seed
4951; fcc 40^3, a = 3.615 A, 1600 K, skin 1.0, every 1 delay 5 check yes,
dt = 5 fs) in both precisions, force errors against the FP64 oracle, NVE
drift, the neighbour-build share of step time, the roofline of the dominant
kernel (LJ force) and the CPU baseline (the FP64 oracle on the host).

A step is one pass of the hot path: initial integrate (+ rebuild trigger),
neighbour rebuild when decided, pair forces, final integrate.  Inputs are
device resident when the timed region starts; the region is bracketed by
device synchronisation (mdx_run returns after its stream drains) and covers
exactly --steps steps.  Kernel durations come from HIP events recorded on the
library's stream inside the timed region.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Matom-steps/s, LJ 2.048M & EAM 256k atoms, mixed vs FP64 at stated force err"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def lj_inputs(ncell):
    from lammps_amd import mdx
    lo, hi, x = mdx.lattice("fcc", 0.8442, "lj", (0, 0, 0), (ncell,) * 3)
    t = np.ones(len(x), dtype=np.int32)
    mass = np.array([0.0, 1.0])
    v = mdx.velocity_create(x, t, mass, 1.44, 87287, dist="gaussian", loop="geom")
    return dict(lo=lo, hi=hi, x=x, t=t, mass=mass, v=v)


def eam_inputs(ncell, table):
    from lammps_amd import mdx
    lo, hi, x = mdx.lattice("fcc", 3.615, "metal", (0, 0, 0), (ncell,) * 3)
    t = np.ones(len(x), dtype=np.int32)
    mass = np.array([0.0, 63.55])
    v = mdx.velocity_create(x, t, mass, 1600.0, 376847, dist="gaussian", loop="geom", units="metal")
    return dict(lo=lo, hi=hi, x=x, t=t, mass=mass, v=v, table=table)


def make_ctx(kind, inp, precision):
    from lammps_amd import mdx
    ctx = mdx.Context(0)
    ctx.set_box(inp["lo"], inp["hi"], (1, 1, 1))
    ctx.set_atoms(inp["x"], inp["t"], inp["mass"], v=inp["v"])
    if kind == "lj":
        ctx.pair_lj_from_coeffs(1, {(1, 1): (1.0, 1.0, 2.5)})
        ctx.neigh_modify(0.3, every=20, delay=0, check=True)
    else:
        mdx.eam_context_tables(ctx, inp["table"])
        ctx.neigh_modify(1.0, every=1, delay=5, check=True)
    ctx.precision(precision)
    return ctx


def run_case(kind, inp, precision, steps, warmup):
    """Setup, warmup, then exactly `steps` timed steps; returns timing,
    kernel times and thermo sums at the start/end of the timed region."""
    from lammps_amd import mdx
    units = "lj" if kind == "lj" else "metal"
    dt = 0.005
    ftm2v = mdx.UNIT_CONSTANTS[units]["ftm2v"]
    ctx = make_ctx(kind, inp, precision)
    ctx.setup()
    if warmup:
        ctx.run(warmup, dt, ftm2v, 0)
    st0 = ctx.stats()
    ctx.profiling(True)
    t0 = time.perf_counter()
    rows = ctx.run(steps, dt, ftm2v, steps)
    t1 = time.perf_counter()
    st = ctx.stats()
    ctx.profiling(False)
    n = len(inp["t"])
    secs = t1 - t0
    res = dict(n=n, seconds=secs, ms_per_step=1e3 * secs / steps, value=n * steps / secs / 1e6,
               builds=st["ncalls"] - st0["ncalls"], total_neighbors=st["total_neighbors"],
               max_neighbors=st["max_neighbors"], kernel_ms=st["kernel_ms"], kernel_launches=st["kernel_launches"],
               end_row=rows[-1][1] if rows else None, ctx=ctx)
    names = ("build", "lj", "eam_rho", "eam_force", "nve_initial", "nve_final", "fused")
    log("  kernels (ms/step, launches): " + ", ".join(
        "%s %.4f (%d)" % (nm, st["kernel_ms"][k] / steps, st["kernel_launches"][k])
        for k, nm in enumerate(names) if st["kernel_launches"][k]))
    return res


def profiled_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "lj2m_eam256k_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(kernel + " ") or line.startswith("void " + kernel + " "):
                try:
                    return float(line.split()[-1]) * 1e6
                except ValueError:
                    return None
    return None


def energy_per_atom(raw, n, units):
    from lammps_amd import mdx
    mvv2e = mdx.UNIT_CONSTANTS[units]["mvv2e"]
    return (0.5 * mvv2e * raw[0] + raw[1]) / n


def drift_run(kind, inp, precision, steps):
    """(E_tot(t) - E_tot(0))/N over `steps` NVE steps, thermo every steps/10."""
    from lammps_amd import mdx
    units = "lj" if kind == "lj" else "metal"
    ctx = make_ctx(kind, inp, precision)
    rows = ctx.setup() + ctx.run(steps, 0.005, mdx.UNIT_CONSTANTS[units]["ftm2v"], max(1, steps // 10))
    n = len(inp["t"])
    e = [energy_per_atom(raw, n, units) for _, raw in rows]
    steps_list = [s for s, _ in rows]
    slope = float(np.polyfit(steps_list, e, 1)[0]) if len(e) > 2 else 0.0
    ctx.close()
    return dict(steps=steps, e0=e[0], e_end=e[-1], drift_per_atom=e[-1] - e[0], slope_per_step=slope)


def force_errors(f, fref):
    e = np.linalg.norm(f - fref, axis=1)
    s = np.sqrt(np.mean(np.sum(fref * fref, axis=1)))
    return dict(max_rel=float(e.max() / s), rms_rel=float(np.sqrt(np.mean(e * e)) / s))


def cpu_baseline_and_errors(kind, inp, ctxs, sample_steps):
    """The FP64 oracle (test infrastructure, used here only as the checker and
    the CPU baseline): (1) one force evaluation on the device's current
    configuration -> max/rms force error of each device precision;
    (2) a bounded NVE sample of the same workload on 1 host core."""
    from oracle import mdref as M
    from lammps_amd import mdx
    out = {}
    # (1) force error on the configuration the device reached
    ctx = ctxs["mixed"]
    x, _, _, _ = ctx.get_atoms()
    units = M.LJ if kind == "lj" else M.METAL
    r = M.Ref(units)
    r.box(inp["lo"], inp["hi"])
    r.add_atoms(x, inp["t"])
    if kind == "lj":
        r.mass(1, 1.0)
        r.pair_lj(2.5)
        r.pair_coeff(1, 1, 1.0, 1.0, 2.5)
        r.neighbor(0.3, every=20, delay=0, check=True)
    else:
        r.pair_eam(inp["table"])
        r.neighbor(1.0, every=1, delay=5, check=True)
    r.sortfreq(0)
    t0 = time.perf_counter()
    r.compute(1, 1)
    t_force = time.perf_counter() - t0
    _, _, fref, _, _ = r.state()
    for prec, c in ctxs.items():
        c.set_positions(x)
        c.neigh_build()  # the other context's own list belongs to its own trajectory
        out[prec] = force_errors(c.compute(0, 0)["f"], fref)
    # (2) bounded NVE sample, same commands as the device workload
    rs = M.Ref(units)
    if kind == "lj":
        rs.lattice_box(M.FCC, 0.8442, (0, 0, 0), (80, 80, 80))
        rs.mass(1, 1.0)
        rs.velocity_create(1.44, 87287, dist="gaussian", loop="geom")
        rs.pair_lj(2.5)
        rs.pair_coeff(1, 1, 1.0, 1.0, 2.5)
        rs.neighbor(0.3, every=20, delay=0, check=True)
    else:
        rs.lattice_box(M.FCC, 3.615, (0, 0, 0), (40, 40, 40))
        rs.pair_eam(inp["table"])
        rs.velocity_create(1600.0, 376847, dist="gaussian", loop="geom")
        rs.neighbor(1.0, every=1, delay=5, check=True)
    rs.timestep(0.005)
    rs.setup()
    t0 = time.perf_counter()
    rs.run(sample_steps)
    secs = time.perf_counter() - t0
    n = rs.natoms
    base = dict(value=n * sample_steps / secs / 1e6, unit="Matom-steps/s", cores=1, kind="port",
                sample="%s, %d atoms, %d NVE steps after setup (%.1f s); FP64 oracle oracle/mdref.c, gcc -O2, "
                       "1 thread" % (kind, n, sample_steps, secs))
    return out, base, t_force


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-eam", action="store_true")
    ap.add_argument("--drift-steps", type=int, default=1000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (spatial decomposition over "
              "several GPUs is described in DESIGN.md, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from lammps_amd import mdx
    K, W = args.steps, args.warmup

    log("LJ 2.048M inputs")
    lj = lj_inputs(80)
    res = {}
    for prec, p in (("mixed", mdx.MIXED), ("fp64", mdx.FP64)):
        log("LJ %s: %d warmup + %d timed steps" % (prec, W, K))
        res[prec] = run_case("lj", lj, p, K, W)
        log("LJ %s: %.1f Matom-steps/s (%.3f ms/step)" % (prec, res[prec]["value"], res[prec]["ms_per_step"]))
    n = res["mixed"]["n"]

    eam = None
    if not args.no_eam:
        tmpdir = tempfile.mkdtemp(prefix="mdx_bench_")
        table = os.path.join(tmpdir, "synthetic_cu.eam")
        params = mdx.write_synthetic_funcfl(table, seed=4951)
        ein = eam_inputs(40, table)
        eres = {}
        for prec, p in (("mixed", mdx.MIXED), ("fp64", mdx.FP64)):
            log("EAM %s: %d warmup + %d timed steps" % (prec, W, K))
            eres[prec] = run_case("eam", ein, p, K, W)
            log("EAM %s: %.1f Matom-steps/s" % (prec, eres[prec]["value"]))
        eam = dict(workload="eam fcc 40^3 = 256,000 atoms, a=3.615 A, synthetic funcfl (seed 4951), rc 4.95, "
                            "skin 1.0, every 1 delay 5 check yes, 1600 K, dt 5 fs",
                   params=params,
                   mixed=dict(value=eres["mixed"]["value"], ms_per_step=eres["mixed"]["ms_per_step"],
                              builds=eres["mixed"]["builds"]),
                   fp64=dict(value=eres["fp64"]["value"], ms_per_step=eres["fp64"]["ms_per_step"],
                             builds=eres["fp64"]["builds"]),
                   mixed_over_fp64=eres["mixed"]["value"] / eres["fp64"]["value"])

    # measured streaming-copy rate of this device (next to the 8 TB/s spec)
    probe = mdx.Context(0)
    copy_gbs = probe.copy_bandwidth(1 << 30, 20)
    probe.close()
    log("HBM copy probe: %.0f GB/s" % copy_gbs)

    # roofline of the dominant kernel: the LJ force kernel, which in the fused
    # path (the default) also runs fix nve's final integrate of step n and the
    # initial integrate of n+1.  Algorithmic bytes per atom and launch:
    # 4*(n_h+1) half-list indices + 32 B x_i (+ 72 B when fused: v read and
    # write, x(n+1) write; 24 B f_i write otherwise).
    km = res["mixed"]["kernel_ms"]
    kl = res["mixed"]["kernel_launches"]
    fused = kl[mdx.K_FUSED] > kl[mdx.K_LJ]
    kk = mdx.K_FUSED if fused else mdx.K_LJ
    lj_ms = km[kk] / max(1, kl[kk])
    n_h = res["mixed"]["total_neighbors"] / 2.0 / n
    bytes_per_launch = n * (4.0 * (n_h + 1.0) + 32.0 + (72.0 if fused else 24.0))
    achieved = bytes_per_launch / (lj_ms * 1e-3) / 1e9 if lj_ms > 0 else None
    total_ms = res["mixed"]["seconds"] * 1e3
    kname = "k_lj<true, false, 8, %s>" % ("true" if fused else "false")
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=None if achieved is None else achieved / HBM_PEAK_GBS,
                    traffic=profiled_traffic(kname), kernel=kname, avg_launch_ms=lj_ms,
                    bytes_per_launch=bytes_per_launch, n_half_neighbors_per_atom=n_h,
                    measured_copy_gbs=copy_gbs,
                    step_algorithmic_frac=(n * (4.0 * (n_h + 1) + 32 + 24 + 200 + 16 + (4.0 * n_h + 32) / 20.0)
                                           / (res["mixed"]["ms_per_step"] * 1e-3) / 1e9) / HBM_PEAK_GBS)
    shares = {name: km[k] / total_ms for name, k in
              (("neigh_build", mdx.K_BUILD), ("pair", mdx.K_LJ), ("pair_fused_nve", mdx.K_FUSED),
               ("nve_initial", mdx.K_NVE_INITIAL), ("nve_final", mdx.K_NVE_FINAL))}

    errors, cpu, t_force = None, None, None
    if not args.no_cpu_baseline:
        log("oracle force check + CPU baseline sample")
        errors, cpu, t_force = cpu_baseline_and_errors(
            "lj", lj, {"mixed": res["mixed"]["ctx"], "fp64": res["fp64"]["ctx"]}, sample_steps=5)

    drift = None
    if args.drift_steps:
        log("NVE drift over %d steps" % args.drift_steps)
        drift = {prec: drift_run("lj", lj, p, args.drift_steps) for prec, p in (("mixed", mdx.MIXED),
                                                                               ("fp64", mdx.FP64))}

    out = {
        "metric": METRIC,
        "value": res["mixed"]["value"],
        "unit": "Matom-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": res["mixed"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair math, f64 positions/accumulation",
        "data": "synthetic (seeded fcc lattice + velocity create, as the reference's own commands)",
        "config": {"workload": "lj/cut fcc 80^3 = 2,048,000 atoms, rho*=0.8442, rc=2.5, skin=0.3, "
                               "every 20 delay 0 check yes, T*=1.44 gaussian seed 87287, NVE dt=0.005",
                   "precision": "mixed", "atoms": n, "builds_in_timed_steps": res["mixed"]["builds"]},
        "fp64": {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["ms_per_step"]},
        "mixed_over_fp64": res["mixed"]["value"] / res["fp64"]["value"],
        "force_error_vs_fp64_oracle": errors,
        "oracle_force_eval_s": t_force,
        "step_time_share": shares,
        "nve_drift": drift,
        "eam_256k": eam,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    for r in res.values():
        r["ctx"].close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
