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
    launch = ctx.autotune()  # threads per atom and block of the pair kernels
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
               end_row=rows[-1][1] if rows else None, ctx=ctx, launch=launch)
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


def halo_share(G, L, reps=20):
    """One GPU's share of the 2.048 M-atom LJ run split 2x2x2 over eight GPUs:
    the octant's 256 k owned atoms in their own context with the ghost shell
    (rc + skin wide, about 70 k atoms) packed from the full system, which
    stands in for the seven neighbouring ranks, and unpacked into the octant.
    Device times from the library's HIP events: the pack of one exchange (the
    26 swaps in one launch), the unpack, and the octant's force kernel; the send lists (built once per
    reneighbouring, one pass over the atoms for the 26 regions) are timed on
    the host."""
    import torch
    from lammps_amd import mdx
    cut = 2.5 + 0.3
    G.neigh_build()
    xg = G.get_atoms()[0]
    regions = mdx.octant_shell_regions(L, cut)
    counts = G.halo_sendlists(regions)  # sizes the lists
    t0 = time.perf_counter()
    for _ in range(5):
        counts = G.halo_sendlists(regions)
    t_list = (time.perf_counter() - t0) * 1e3 / 5
    ng = sum(counts)
    buf = torch.empty(3 * ng, dtype=torch.float64, device="cuda")
    G.profiling(True)
    for _ in range(reps):
        G.halo_pack(-1, buf.data_ptr())  # the 26 swaps in one launch, in swap order
    st = G.stats()
    G.profiling(False)
    pack_ms = st["kernel_ms"][mdx.K_HALO] / reps
    h = L / 2.0
    owned = np.where(np.all((xg >= 0.0) & (xg < h), axis=1))[0]
    place = np.full((ng, 3), h / 2.0)
    D = mdx.Context(0)
    D.set_box((-cut - 0.5,) * 3, (h + cut + 0.5,) * 3, (0, 0, 0))
    D.set_atoms(np.vstack([xg[owned], place]), np.ones(len(owned) + ng, dtype=np.int32), np.array([0.0, 1.0]))
    D.pair_lj_from_coeffs(1, {(1, 1): (1.0, 1.0, 2.5)})
    D.precision(mdx.MIXED)
    D.neigh_modify(0.3)
    D.profiling(True)
    for _ in range(reps):
        D.halo_unpack(len(owned), ng, buf.data_ptr())
    D.neigh_build()
    D.set_launch(0, 4)
    for _ in range(reps):
        fo = D.compute(0, 0)["f"][: len(owned)]
    st = D.stats()
    D.profiling(False)
    fr = G.compute(0, 0)["f"][owned]
    D.close()
    err = float(np.max(np.linalg.norm(fo - fr, axis=1)) / np.sqrt(np.mean(np.sum(fr * fr, axis=1))))
    unpack_ms = st["kernel_ms"][mdx.K_HALO] / reps
    return dict(owned=int(len(owned)), ghosts=int(ng), ghost_bytes=24 * ng, send_lists_ms=t_list,
                pack_ms=pack_ms, unpack_ms=unpack_ms,
                pack_gbs=(24.0 * ng * 2 + 4.0 * ng) / (pack_ms * 1e-3) / 1e9 if pack_ms > 0 else None,
                force_kernel_ms=st["kernel_ms"][mdx.K_LJ] / max(1, st["kernel_launches"][mdx.K_LJ]),
                force_max_rel_vs_full_system=err)


def plugin_mode(ctx, reps=10):
    """The per-call pair-style path (pair_style lj/cut/mdx: positions in from
    the host, forces back, every step) on the same 2.048 M-atom state: the
    PCIe-inclusive rate, reported beside the resident `value`."""
    import ctypes
    from lammps_amd import mdx
    x = np.ascontiguousarray(ctx.get_atoms()[0])
    f = np.zeros_like(x)  # the caller's force array, reused every step as atom->f is
    L = ctx._L
    xp = x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    fp = f.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for _ in range(2):
        ctx._ck(L.mdx_set_positions(ctx.h, xp))
        ctx._ck(L.mdx_compute(ctx.h, 0, 0, fp, None, None, None, None))
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx._ck(L.mdx_set_positions(ctx.h, xp))
        ctx._ck(L.mdx_compute(ctx.h, 0, 0, fp, None, None, None, None))
    secs = (time.perf_counter() - t0) / reps
    n = len(x)
    return dict(ms_per_step=secs * 1e3, value=n / secs / 1e6, unit="Matom-steps/s",
                host_bytes_per_step=48 * n, note="set_positions (24 B/atom H2D) + compute + forces (24 B/atom D2H)")


def fit_lj_ka(reps=10, check_oracle=True):
    """Differentiable lj/cut force matching on the KA mixture: 64 snapshots of
    8,000 atoms (device MD preparation), forward forces in mixed precision,
    loss = mean |F - F_t|^2 / 3, backward through the autograd Function to
    dL/d(eps, sigma).  Reports evaluations/s and the gradient against the
    FP64 oracle gradient and central differences of the device FP64 loss."""
    import torch
    from lammps_amd import fit, workloads as WL
    from lammps_amd.batch import FP64, MIXED
    t0 = time.perf_counter()
    snaps, types, L = WL.ka_snapshots()
    prep = time.perf_counter() - t0
    S, N = snaps.shape[0], snaps.shape[1]
    natoms = [N] * S
    lo, hi, per = [[0, 0, 0]] * S, [[L, L, L]] * S, [[1, 1, 1]] * S
    t_all = np.tile(types, S)
    x = snaps.reshape(-1, 3)
    eps, sig, cut = WL.ka_pair_vectors()
    e_t, s_t = WL.ka_target_vectors()
    dev = torch.device("cuda", 0)
    out = {"workload": "KA 80:20 LJ, 8,000 atoms x %d snapshots (T*=0.8 device MD), cut 2.5 sigma_ab" % S,
           "prep_s": prep}
    fm64 = fit.LJForceMatch(natoms, lo, hi, per, 2, t_all, x, cut, precision=FP64)
    with torch.no_grad():
        Ft = fm64(torch.tensor(e_t), torch.tensor(s_t)).clone()

    def loss_and_grad(fm):
        pe = torch.tensor(eps, requires_grad=True)
        ps = torch.tensor(sig, requires_grad=True)
        F = fm(pe, ps)
        loss = ((F - Ft) ** 2).sum() / (3.0 * F.shape[0])
        loss.backward()
        return loss.item(), pe.grad.numpy().copy(), ps.grad.numpy().copy(), F

    res = {}
    for name, fm in (("mixed", fit.LJForceMatch(natoms, lo, hi, per, 2, t_all, x, cut, precision=MIXED)),
                     ("fp64", fm64)):
        fm.batch.autotune(True)
        loss_and_grad(fm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            loss, ge, gs, F = loss_and_grad(fm)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res[name] = dict(loss=loss, grad_eps=ge.tolist(), grad_sigma=gs.tolist(), ms_per_loss_and_grad=dt * 1e3,
                         configs_per_s=S / dt, atom_forces_per_s=S * N / dt)
    # central differences of the device FP64 loss (h = 1e-6 relative)
    fd = []
    for which, base in ((0, eps), (1, sig)):
        for k in range(3):
            vals = []
            for sgn in (1, -1):
                pe, ps = eps.copy(), sig.copy()
                (pe, ps)[which][k] = base[k] * (1 + sgn * 1e-6)
                with torch.no_grad():
                    F = fm64(torch.tensor(pe), torch.tensor(ps))
                    vals.append((((F - Ft) ** 2).sum() / (3.0 * F.shape[0])).item())
            fd.append((vals[0] - vals[1]) / (2e-6 * base[k]))
    g64 = np.concatenate([res["fp64"]["grad_eps"], res["fp64"]["grad_sigma"]])
    gmx = np.concatenate([res["mixed"]["grad_eps"], res["mixed"]["grad_sigma"]])
    fd = np.array(fd)
    out["central_fd"] = fd.tolist()
    out["fp64_grad_vs_fd_max_rel"] = float(np.max(np.abs(g64 - fd) / np.abs(fd)))
    out["mixed_grad_vs_fd_max_rel"] = float(np.max(np.abs(gmx - fd) / np.abs(fd)))
    if check_oracle:
        from oracle import fitref as FR
        ref = FR.Batch(natoms, lo, hi, per, 2, t_all, x)
        from lammps_amd.fit import pair_table
        Fth, _ = ref.lj_forces(pair_table(e_t, 2), pair_table(s_t, 2), pair_table(cut, 2))
        F0, _ = ref.lj_forces(pair_table(eps, 2), pair_table(sig, 2), pair_table(cut, 2))
        G = 2.0 * (F0 - Fth) / (3.0 * F0.shape[0])
        t0 = time.perf_counter()
        ge_r, gs_r = ref.lj_vjp(pair_table(eps, 2), pair_table(sig, 2), pair_table(cut, 2), G)
        out["oracle_vjp_s"] = time.perf_counter() - t0
        gr = np.array([ge_r[1, 1], ge_r[1, 2], ge_r[2, 2], gs_r[1, 1], gs_r[1, 2], gs_r[2, 2]])
        out["fp64_grad_vs_oracle_max_rel"] = float(np.max(np.abs(g64 - gr) / np.abs(gr)))
        out["mixed_grad_vs_oracle_max_rel"] = float(np.max(np.abs(gmx - gr) / np.abs(gr)))
    out.update(res)
    fm64.batch.close()
    return out


def eam_replicas(reps=10, check_oracle=True, fd_knots=6):
    """Batched EAM clusters: 512 non-periodic fcc Cu-like clusters of
    200-5000 atoms (seed 7), thermalised on the device, forces + energies +
    the force-matching adjoint w.r.t. the 3x64 knot values in one persistent
    launch per pass.  Reports replica-forces/s and the knot gradient against
    the FP64 oracle and central differences."""
    from lammps_amd import batch as B, workloads as WL
    t0 = time.perf_counter()
    d = WL.eam_clusters(nrep=512, thermal_steps=100)
    prep = time.perf_counter() - t0
    natoms = [len(q) for q in d["x"]]
    x = np.concatenate(d["x"])
    n = len(x)
    p = d["pot"]
    target = WL.perturbed_knots(p)
    per = [[0, 0, 0]] * len(natoms)
    out = {"workload": "512 fcc Cu-like clusters, %d atoms (200-5000 each, seed 7), 64-knot synthetic eam, "
                       "100 NVE steps thermalisation" % n, "prep_s": prep, "atoms": n}

    def make(prec):
        b = B.Batch(natoms, d["lo"], d["hi"], per, 1, np.ones(n, np.int32))
        b.set_positions(x)
        b.precision(prec)
        return b

    b64 = make(B.FP64)
    b64.eam_init(p["drho"], p["dr"], p["cut"], *target)
    Ft, _ = b64.eam_forces(energy=False)
    b64.eam_init(p["drho"], p["dr"], p["cut"], p["frho"], p["rhor"], p["z2r"])
    res = {}
    for name, prec in (("mixed", B.MIXED), ("fp64", B.FP64)):
        b = make(prec)
        b.eam_init(p["drho"], p["dr"], p["cut"], p["frho"], p["rhor"], p["z2r"])
        b.autotune(True)
        F, _ = b.eam_forces()
        b.eam_vjp(2.0 * (F - Ft) / (3.0 * n))
        t0 = time.perf_counter()
        for _ in range(reps):
            F, e = b.eam_forces()
        t_fwd = (time.perf_counter() - t0) / reps
        G = 2.0 * (F - Ft) / (3.0 * n)
        t0 = time.perf_counter()
        for _ in range(reps):
            g = b.eam_vjp(G)
        t_adj = (time.perf_counter() - t0) / reps
        res[name] = dict(ms_forward=t_fwd * 1e3, ms_adjoint=t_adj * 1e3, replica_forces_per_s=len(natoms) / t_fwd,
                         replica_force_and_grad_per_s=len(natoms) / (t_fwd + t_adj),
                         launch={B.KIND_NAMES[k]: b.get_launch(k) for k in (3, 4, 5, 6)},
                         loss=float(np.sum((F - Ft) ** 2) / (3.0 * n)))
        res[name]["grad"] = g
        b.close()
    g64 = res["fp64"].pop("grad")
    gmx = res["mixed"].pop("grad")
    # central differences of the device FP64 loss on a sample of knots
    F0, _ = b64.eam_forces(energy=False)
    knots = [p["frho"], p["rhor"], p["z2r"]]
    errs = []
    rng = np.random.default_rng(0)
    for which in range(3):
        nz = np.nonzero(np.abs(g64[which]) > 1e-3 * np.max(np.abs(g64[which])))[0]
        for k in rng.choice(nz, size=min(fd_knots, len(nz)), replace=False):
            h = 1e-6 * max(abs(knots[which][k]), 1e-3)
            vals = []
            for sgn in (1, -1):
                kk = [a.copy() for a in knots]
                kk[which][k] += sgn * h
                b64.eam_init(p["drho"], p["dr"], p["cut"], *kk)
                F, _ = b64.eam_forces(energy=False)
                vals.append(np.sum((F - Ft) ** 2) / (3.0 * n))
            fdv = (vals[0] - vals[1]) / (2 * h)
            scale = np.max(np.abs(g64[which]))
            errs.append(dict(fn=("frho", "rhor", "z2r")[which], knot=int(k), fd=float(fdv),
                             fp64_rel=float(abs(g64[which][k] - fdv) / scale),
                             mixed_rel=float(abs(gmx[which][k] - fdv) / scale)))
    b64.close()
    out["fd_check"] = errs
    out["fp64_grad_vs_fd_max_rel"] = max(e["fp64_rel"] for e in errs)
    out["mixed_grad_vs_fd_max_rel"] = max(e["mixed_rel"] for e in errs)
    if check_oracle:
        from oracle import fitref as FR
        ref = FR.Batch(natoms, d["lo"], d["hi"], per, 1, np.ones(n), x)
        ref.eam_set(p["drho"], p["dr"], p["cut"], *target)
        Ftr, _, _, _ = ref.eam_forces()
        ref.eam_set(p["drho"], p["dr"], p["cut"], p["frho"], p["rhor"], p["z2r"])
        F0r, _, _, _ = ref.eam_forces()
        t0 = time.perf_counter()
        gr = ref.eam_vjp(2.0 * (F0r - Ftr) / (3.0 * n))
        out["oracle_vjp_s"] = time.perf_counter() - t0
        out["fp64_grad_vs_oracle_max_rel"] = max(float(np.max(np.abs(a - r_)) / np.max(np.abs(r_)))
                                                  for a, r_ in zip(g64, gr))
        out["mixed_grad_vs_oracle_max_rel"] = max(float(np.max(np.abs(a - r_)) / np.max(np.abs(r_)))
                                                   for a, r_ in zip(gmx, gr))
    out.update(res)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-eam", action="store_true")
    ap.add_argument("--drift-steps", type=int, default=1000)
    ap.add_argument("--no-fit", action="store_true", help="skip the force-matching workloads (KA, EAM clusters)")
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
    # the kernel as rocprofv3 names it: k_lj<MIXED, EV, TPA, INTEG, Q, ONE> with
    # the autotuned threads per atom, the fixed-point positions of the mixed
    # path and the single-type specialisation (taken for TPA 4, 8 and 16)
    tpa = res["mixed"]["launch"]["tpa"]
    kname = "k_lj<true, false, %d, %s, true, %s>" % (tpa, "true" if fused else "false",
                                                      "true" if tpa in (4, 8, 16) else "false")
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
            "lj", lj, {"mixed": res["mixed"]["ctx"], "fp64": res["fp64"]["ctx"]}, sample_steps=40)

    drift = None
    if args.drift_steps:
        log("NVE drift over %d steps" % args.drift_steps)
        drift = {prec: drift_run("lj", lj, p, args.drift_steps) for prec, p in (("mixed", mdx.MIXED),
                                                                               ("fp64", mdx.FP64))}

    fit_ka, replicas = None, None
    if not args.no_fit:
        log("force matching: KA lj/cut 64 x 8,000 atoms")
        fit_ka = fit_lj_ka(check_oracle=not args.no_cpu_baseline)
        log("  KA loss+grad %.2f ms (mixed), grad vs FD %.2e" % (fit_ka["mixed"]["ms_per_loss_and_grad"],
                                                              fit_ka["mixed_grad_vs_fd_max_rel"]))
        log("force matching: 512 eam clusters")
        replicas = eam_replicas(check_oracle=not args.no_cpu_baseline)
        log("  clusters forward %.2f ms, adjoint %.2f ms (mixed), grad vs FD %.2e"
            % (replicas["mixed"]["ms_forward"], replicas["mixed"]["ms_adjoint"], replicas["mixed_grad_vs_fd_max_rel"]))

    log("per-call pair-style path (host positions in, forces out, every step)")
    plugin = plugin_mode(res["mixed"]["ctx"])
    log("  %.3f ms per step, %.1f Matom-steps/s with PCIe" % (plugin["ms_per_step"], plugin["value"]))
    log("one GPU's share of a 2x2x2 split: octant + ghost shell, pack / unpack")
    share = halo_share(res["mixed"]["ctx"], lj["hi"][0] - lj["lo"][0])
    log("  %d owned + %d ghosts: pack %.3f ms, unpack %.3f ms, force kernel %.3f ms"
        % (share["owned"], share["ghosts"], share["pack_ms"], share["unpack_ms"], share["force_kernel_ms"]))

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
                   "precision": "mixed", "atoms": n, "builds_in_timed_steps": res["mixed"]["builds"],
                   "autotuned_launch": res["mixed"]["launch"]},
        "fp64": {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["ms_per_step"]},
        "mixed_over_fp64": res["mixed"]["value"] / res["fp64"]["value"],
        "force_error_vs_fp64_oracle": errors,
        "oracle_force_eval_s": t_force,
        "step_time_share": shares,
        "nve_drift": drift,
        "eam_256k": eam,
        "force_matching_ka": fit_ka,
        "eam_cluster_replicas": replicas,
        "sharded_share": share,
        "plugin_mode_pcie": plugin,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    for r in res.values():
        r["ctx"].close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
