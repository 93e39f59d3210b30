"""GPU parity at the benchmark's full configuration sizes, against the FP64
oracle (oracle/mdref.c for the MD path, oracle/fitref.c for the batched
forces and gradients) on identical positions.

  C3  lj/cut, fcc 80^3 = 2,048,000 atoms, rho* 0.8442, rc 2.5, skin 0.3,
      Gaussian T* = 1.44 (seed 87287, loop geom), after 100 device NVE steps
  C2  eam, fcc 40^3 = 256,000 atoms, a = 3.615 A, the synthetic funcfl table
      (seed 4951, 500 points, cutoff 4.95), skin 1.0, Gaussian 1600 K, after
      100 device NVE steps (dt 5 fs)
  C1  lj/cut, fcc 20^3 = 32,000 atoms, Gaussian T* = 1.44 (seed 87287, loop
      geom), every 20 delay 0 check yes, 100 NVE steps: the whole run on
      the device against the oracle's run, thermo rows at steps 0, 50, 100
  C4  KA 80:20 lj/cut mixture, 64 snapshots x 8,000 atoms (T* = 0.8 device
      MD), force-matching loss against perturbed parameters
  C5  512 non-periodic fcc clusters of 200-5000 atoms (seed 7), 64-knot
      synthetic eam, force-matching loss against perturbed knots

Bars (stated here):
  * neighbour pair sets: identical to the oracle's half_bin_newton set, in
    both precisions (the mixed build reads the fixed-point mirror), compared
    as sorted (tag_i, tag_j) arrays;
  * forces: FP64 max_rel <= 1e-11, mixed max_rel <= 1e-5, with
    max_rel = max_i |f_i - f_i^ref| / RMS_j |f_j^ref|;
  * energy and virial: relative <= 1e-11 (FP64) / 1e-6 (mixed); EAM rho and
    fp: relative <= 1e-11 (FP64) / 1e-6 (mixed);
  * C1 run: the same builds (and dangerous builds) as the oracle's run;
    thermo rows equal to the oracle's printed digits (FP64); for mixed, the
    reference's regression rule, mean-square error of every thermo column
    over the rows < 1e-6 (regress/benchmark.py:105-124);
  * gradients: against the FP64 oracle gradient <= 1e-9 (FP64) / 1e-6
    (mixed) relative, and against central differences (h = 1e-6 relative)
    of the device FP64 loss <= 1e-6 relative.
"""
import os
import tempfile

import numpy as np
import pytest

from gpu_helpers import force_errors
from oracle import mdref as M

pytestmark = pytest.mark.gpu


def _pair_keys(pairs):
    p = np.asarray(pairs, dtype=np.int64)
    a = np.minimum(p[:, 0], p[:, 1])
    b = np.maximum(p[:, 0], p[:, 1])
    k = (a << 32) | b
    k.sort()
    return k


def _device_ctx(kind, box, x, t, mass, precision, table=None, v=None):
    from lammps_amd import mdx
    ctx = mdx.Context(0)
    ctx.set_box(box[0], box[1], (1, 1, 1))
    ctx.set_atoms(x, t, mass, v=v, tags=np.arange(1, len(t) + 1))
    if kind == "lj":
        ctx.pair_lj_from_coeffs(1, {(1, 1): (1.0, 1.0, 2.5)})
        ctx.neigh_modify(0.3, every=20, delay=0, check=True)
    else:
        mdx.eam_context_tables(ctx, table)
        ctx.neigh_modify(1.0, every=1, delay=5, check=True)
    ctx.precision(precision)
    return ctx


def _thermalised(kind, ncell, table=None, steps=100):
    """Positions after `steps` FP64 device NVE steps from the benchmark's
    seeded lattice and velocities (tag order), and the box."""
    from lammps_amd import mdx
    if kind == "lj":
        lo, hi, x = mdx.lattice("fcc", 0.8442, "lj", (0, 0, 0), (ncell,) * 3)
        mass = np.array([0.0, 1.0])
        t = np.ones(len(x), dtype=np.int32)
        v = mdx.velocity_create(x, t, mass, 1.44, 87287, dist="gaussian", loop="geom")
        ftm2v = 1.0
    else:
        lo, hi, x = mdx.lattice("fcc", 3.615, "metal", (0, 0, 0), (ncell,) * 3)
        mass = np.array([0.0, 63.55])
        t = np.ones(len(x), dtype=np.int32)
        v = mdx.velocity_create(x, t, mass, 1600.0, 376847, dist="gaussian", loop="geom", units="metal")
        ftm2v = mdx.UNIT_CONSTANTS["metal"]["ftm2v"]
    ctx = _device_ctx(kind, (lo, hi), x, t, mass, mdx.FP64, table, v)
    ctx.setup()
    ctx.run(steps, 0.005, ftm2v, 0)
    xs = ctx.get_atoms()[0]
    ctx.close()
    return (lo, hi), xs, t, mass


def _oracle(kind, box, x, t, table=None):
    r = M.Ref(M.LJ if kind == "lj" else M.METAL)
    r.box(box[0], box[1])
    r.add_atoms(x, t)
    if kind == "lj":
        r.mass(1, 1.0)
        r.pair_lj(2.5)
        r.pair_coeff(1, 1, 1.0, 1.0, 2.5)
        r.neighbor(0.3, every=20, delay=0, check=True)
    else:
        r.pair_eam(table)
        r.neighbor(1.0, every=1, delay=5, check=True)
    r.sortfreq(0)
    r.compute(1, 1)
    return r


@pytest.fixture(scope="module")
def c3():
    box, x, t, mass = _thermalised("lj", 80)
    r = _oracle("lj", box, x, t)
    e, w = r.energy_virial()
    out = dict(box=box, x=x, t=t, mass=mass, f=r.state()[2], e=e, w=w, keys=_pair_keys(r.list_pairs()))
    del r
    return out


@pytest.fixture(scope="module")
def c2():
    from lammps_amd import mdx
    d = tempfile.mkdtemp(prefix="mdx_c2_")
    table = os.path.join(d, "synthetic_cu.eam")
    mdx.write_synthetic_funcfl(table, seed=4951)
    box, x, t, mass = _thermalised("eam", 40, table)
    r = _oracle("eam", box, x, t, table)
    e, w = r.energy_virial()
    rho, fp = r.eam_rho_fp()
    out = dict(box=box, x=x, t=t, mass=mass, table=table, f=r.state()[2], e=e, w=w, rho=rho, fp=fp,
               keys=_pair_keys(r.list_pairs()))
    del r
    return out


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c3_lj_2048k_pair_set_identical(c3, prec):
    from lammps_amd import mdx
    assert len(c3["t"]) == 2048000
    ctx = _device_ctx("lj", c3["box"], c3["x"], c3["t"], c3["mass"], mdx.FP64 if prec == "fp64" else mdx.MIXED)
    ctx.neigh_build()
    got = _pair_keys(ctx.neigh_pairs())
    st = ctx.stats()
    ctx.close()
    assert len(got) == len(c3["keys"])
    assert np.array_equal(got, c3["keys"])
    assert st["total_neighbors"] == 2 * len(c3["keys"])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c3_lj_2048k_forces_energy_virial(c3, prec):
    from lammps_amd import mdx
    ctx = _device_ctx("lj", c3["box"], c3["x"], c3["t"], c3["mass"], mdx.FP64 if prec == "fp64" else mdx.MIXED)
    out = ctx.compute(1, 1)
    ctx.close()
    mx, rms = force_errors(out["f"], c3["f"])
    tol, etol = (1e-11, 1e-11) if prec == "fp64" else (1e-5, 1e-6)
    print("C3 lj 2.048M %s: force max_rel=%.3e rms_rel=%.3e energy rel=%.3e" %
          (prec, mx, rms, abs(out["eng"] - c3["e"]) / abs(c3["e"])))
    assert mx <= tol
    assert abs(out["eng"] - c3["e"]) <= etol * abs(c3["e"])
    np.testing.assert_allclose(out["virial"], c3["w"], rtol=etol, atol=etol * np.abs(c3["w"]).max())


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c2_eam_256k_pair_set_identical(c2, prec):
    from lammps_amd import mdx
    assert len(c2["t"]) == 256000
    ctx = _device_ctx("eam", c2["box"], c2["x"], c2["t"], c2["mass"], mdx.FP64 if prec == "fp64" else mdx.MIXED,
                      c2["table"])
    ctx.neigh_build()
    got = _pair_keys(ctx.neigh_pairs())
    ctx.close()
    assert np.array_equal(got, c2["keys"])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c2_eam_256k_forces_energy_virial_rho_fp(c2, prec):
    from lammps_amd import mdx
    ctx = _device_ctx("eam", c2["box"], c2["x"], c2["t"], c2["mass"], mdx.FP64 if prec == "fp64" else mdx.MIXED,
                      c2["table"])
    out = ctx.compute(1, 1)
    rho, fp = ctx.eam_rho_fp()
    ctx.close()
    mx, rms = force_errors(out["f"], c2["f"])
    tol, etol = (1e-11, 1e-11) if prec == "fp64" else (1e-5, 1e-6)
    print("C2 eam 256k %s: force max_rel=%.3e rms_rel=%.3e energy rel=%.3e rho rel=%.3e" %
          (prec, mx, rms, abs(out["eng"] - c2["e"]) / abs(c2["e"]),
           np.max(np.abs(rho - c2["rho"])) / np.max(np.abs(c2["rho"]))))
    assert mx <= tol
    assert abs(out["eng"] - c2["e"]) <= etol * abs(c2["e"])
    np.testing.assert_allclose(out["virial"], c2["w"], rtol=etol, atol=etol * np.abs(c2["w"]).max())
    assert np.max(np.abs(rho - c2["rho"])) <= etol * np.max(np.abs(c2["rho"]))
    assert np.max(np.abs(fp - c2["fp"])) <= etol * np.max(np.abs(c2["fp"]))


# ------------------------------------------------------------ C4 / C5
@pytest.fixture(scope="module")
def c4():
    from lammps_amd import workloads as W
    from lammps_amd.fit import pair_table
    from oracle import fitref as FR
    snaps, types, L = W.ka_snapshots()
    S, N = snaps.shape[0], snaps.shape[1]
    geo = dict(natoms=[N] * S, lo=[[0, 0, 0]] * S, hi=[[L, L, L]] * S, per=[[1, 1, 1]] * S, t=np.tile(types, S),
               x=snaps.reshape(-1, 3))
    eps, sig, cut = W.ka_pair_vectors()
    e_t, s_t = W.ka_target_vectors()
    ref = FR.Batch(geo["natoms"], geo["lo"], geo["hi"], geo["per"], 2, geo["t"], geo["x"])
    tabs = [pair_table(q, 2) for q in (eps, sig, cut)]
    Ft, _ = ref.lj_forces(pair_table(e_t, 2), pair_table(s_t, 2), tabs[2])
    F0, en0 = ref.lj_forces(*tabs)
    G = 2.0 * (F0 - Ft) / (3.0 * F0.shape[0])
    ge, gs = ref.lj_vjp(*tabs, G)
    gref = np.array([ge[1, 1], ge[1, 2], ge[2, 2], gs[1, 1], gs[1, 2], gs[2, 2]])
    return dict(geo=geo, S=S, N=N, eps=eps, sig=sig, cut=cut, tabs=tabs, Ft=Ft, F0=F0, en0=en0, G=G, gref=gref)


def _ka_batch(c, prec):
    from lammps_amd import batch as B
    g = c["geo"]
    b = B.Batch(g["natoms"], g["lo"], g["hi"], g["per"], 2, g["t"])
    b.set_positions(g["x"])
    b.precision(prec)
    b.lj_init(*c["tabs"])
    return b


def test_c4_ka_64x8000_forces_and_gradient(c4):
    from lammps_amd import batch as B
    assert c4["S"] == 64 and c4["N"] == 8000
    for prec, ftol, gtol in ((B.FP64, 1e-11, 1e-9), (B.MIXED, 1e-5, 1e-6)):
        b = _ka_batch(c4, prec)
        f, e = b.lj_forces()
        ge, gs = b.lj_vjp(c4["G"])
        b.close()
        g = np.array([ge[1, 1], ge[1, 2], ge[2, 2], gs[1, 1], gs[1, 2], gs[2, 2]])
        mx, _ = force_errors(f, c4["F0"])
        gerr = np.max(np.abs(g - c4["gref"]) / np.abs(c4["gref"]))
        print("C4 KA 64x8000 prec=%d: force max_rel=%.3e grad rel=%.3e" % (prec, mx, gerr))
        assert mx <= ftol
        assert np.max(np.abs(e - c4["en0"]) / np.abs(c4["en0"])) <= (1e-12 if prec == B.FP64 else 1e-6)
        assert gerr <= gtol


def test_c4_ka_gradient_matches_central_differences(c4):
    from lammps_amd import batch as B
    from lammps_amd.fit import pair_table
    b = _ka_batch(c4, B.FP64)
    ge, gs = b.lj_vjp(c4["G"])
    g = np.array([ge[1, 1], ge[1, 2], ge[2, 2], gs[1, 1], gs[1, 2], gs[2, 2]])
    n3 = 3.0 * c4["F0"].shape[0]
    fd = []
    for which, base in ((0, c4["eps"]), (1, c4["sig"])):
        for k in range(3):
            vals = []
            for sgn in (1, -1):
                pe, ps = c4["eps"].copy(), c4["sig"].copy()
                (pe, ps)[which][k] = base[k] * (1 + sgn * 1e-6)
                b.lj_init(pair_table(pe, 2), pair_table(ps, 2), c4["tabs"][2])
                F, _ = b.lj_forces(energy=False)
                vals.append(np.sum((F - c4["Ft"]) ** 2) / n3)
            fd.append((vals[0] - vals[1]) / (2e-6 * base[k]))
    b.close()
    fd = np.array(fd)
    print("C4 central differences:", fd, "adjoint:", g)
    assert np.max(np.abs(g - fd) / np.abs(fd)) <= 1e-6


@pytest.fixture(scope="module")
def c5():
    from lammps_amd import workloads as W
    from oracle import fitref as FR
    d = W.eam_clusters(nrep=512, thermal_steps=100)
    natoms = [len(q) for q in d["x"]]
    x = np.concatenate(d["x"])
    n = len(x)
    p = d["pot"]
    target = W.perturbed_knots(p)
    per = [[0, 0, 0]] * len(natoms)
    ref = FR.Batch(natoms, d["lo"], d["hi"], per, 1, np.ones(n), x)
    ref.eam_set(p["drho"], p["dr"], p["cut"], *target)
    Ft, _, _, _ = ref.eam_forces()
    ref.eam_set(p["drho"], p["dr"], p["cut"], p["frho"], p["rhor"], p["z2r"])
    F0, e0, rho0, _ = ref.eam_forces()
    G = 2.0 * (F0 - Ft) / (3.0 * n)
    gref = ref.eam_vjp(G)
    return dict(d=d, natoms=natoms, x=x, n=n, p=p, per=per, Ft=Ft, F0=F0, e0=e0, rho0=rho0, G=G, gref=gref)


def _cluster_batch(c, prec):
    from lammps_amd import batch as B
    b = B.Batch(c["natoms"], c["d"]["lo"], c["d"]["hi"], c["per"], 1, np.ones(c["n"], np.int32))
    b.set_positions(c["x"])
    b.precision(prec)
    p = c["p"]
    b.eam_init(p["drho"], p["dr"], p["cut"], p["frho"], p["rhor"], p["z2r"])
    return b


def test_c5_512_clusters_forces_and_knot_gradient(c5):
    from lammps_amd import batch as B
    assert len(c5["natoms"]) == 512 and min(c5["natoms"]) >= 200 and max(c5["natoms"]) <= 5000
    for prec, ftol, gtol in ((B.FP64, 1e-11, 1e-9), (B.MIXED, 1e-5, 1e-6)):
        b = _cluster_batch(c5, prec)
        f, e, rho, _ = b.eam_forces(rho_fp=True)
        g = b.eam_vjp(c5["G"])
        b.close()
        mx, _ = force_errors(f, c5["F0"])
        gerr = max(float(np.max(np.abs(a - r)) / np.max(np.abs(r))) for a, r in zip(g, c5["gref"]))
        print("C5 512 clusters prec=%d: force max_rel=%.3e knot grad rel=%.3e" % (prec, mx, gerr))
        assert mx <= ftol
        assert np.max(np.abs(e - c5["e0"]) / np.abs(c5["e0"])) <= (1e-12 if prec == B.FP64 else 1e-6)
        assert np.max(np.abs(rho - c5["rho0"])) <= (1e-12 if prec == B.FP64 else 1e-6) * np.max(np.abs(c5["rho0"]))
        assert gerr <= gtol


def test_c5_knot_gradient_matches_central_differences(c5):
    from lammps_amd import batch as B
    b = _cluster_batch(c5, B.FP64)
    g = b.eam_vjp(c5["G"])
    p = c5["p"]
    knots = [p["frho"], p["rhor"], p["z2r"]]
    n3 = 3.0 * c5["n"]
    rng = np.random.default_rng(0)
    worst = 0.0
    for which in range(3):
        nz = np.nonzero(np.abs(g[which]) > 1e-3 * np.max(np.abs(g[which])))[0]
        for k in rng.choice(nz, size=min(6, len(nz)), replace=False):
            h = 1e-6 * max(abs(knots[which][k]), 1e-3)
            vals = []
            for sgn in (1, -1):
                kk = [a.copy() for a in knots]
                kk[which][k] += sgn * h
                b.eam_init(p["drho"], p["dr"], p["cut"], *kk)
                F, _ = b.eam_forces(energy=False)
                vals.append(np.sum((F - c5["Ft"]) ** 2) / n3)
            fd = (vals[0] - vals[1]) / (2 * h)
            err = abs(g[which][k] - fd) / np.max(np.abs(g[which]))
            worst = max(worst, err)
            assert err <= 1e-6, (which, int(k), g[which][k], fd)
    b.close()
    print("C5 knot gradient vs central differences: worst %.3e" % worst)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c1_lj_32k_gaussian_run_matches_the_oracle(prec):
    from lammps_amd import mdx
    r = M.Ref(M.LJ)
    r.lattice_box(M.FCC, 0.8442, (0, 0, 0), (20, 20, 20))
    r.mass(1, 1.0)
    r.velocity_create(1.44, 87287, dist="gaussian", loop="geom")
    r.pair_lj(2.5)
    r.pair_coeff(1, 1, 1.0, 1.0, 2.5)
    r.neighbor(0.3, every=20, delay=0, check=True)
    r.timestep(0.005)
    x0, v0, _, t0, _ = r.state()
    r.thermo(50)
    r.setup()
    ref_rows = r.run(100)
    L = 20 * (4.0 / 0.8442) ** (1.0 / 3.0)
    ctx = mdx.Context(0)
    ctx.set_box((0, 0, 0), (L, L, L), (1, 1, 1))
    ctx.set_atoms(x0, t0, np.array([0.0, 1.0]), v=v0, tags=np.arange(1, len(t0) + 1))
    ctx.pair_lj_from_coeffs(1, {(1, 1): (1.0, 1.0, 2.5)})
    ctx.neigh_modify(0.3, every=20, delay=0, check=True)
    ctx.precision(mdx.FP64 if prec == "fp64" else mdx.MIXED)
    rows = ctx.setup() + ctx.run(100, 0.005, 1.0, 50)
    st = ctx.stats()
    ctx.close()
    n = len(t0)
    got = {int(step): mdx.thermo_columns(raw, n, (L, L, L), "lj") for step, raw in rows}
    ref = {int(q[0]): q[1:] for q in ref_rows}
    common = sorted(set(got) & set(ref))
    assert 50 in common and 100 in common
    rs = r.neigh_stats()
    assert st["ncalls"] == rs["builds"] > 0 and st["ndanger"] == rs["dangerous"]
    if prec == "fp64":
        for step in common:
            for a, b in zip(got[step], ref[step]):
                assert ("%12.8g" % a).strip() == ("%12.8g" % b).strip(), (step, a, b)
    else:
        g = np.array([got[k] for k in common])
        q = np.array([ref[k] for k in common])
        mse = np.mean((g - q) ** 2, axis=0)
        assert np.all(mse < 1e-6), mse
