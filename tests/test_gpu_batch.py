"""Batched forces and force-matching adjoints on the device (mdx_batch.h)
against the FP64 oracle (oracle/fitref.c, pinned by finite differences in
tests/test_fit_oracle.py).

Tolerances: FP64 forces max_rel <= 1e-11 and gradients <= 1e-9 relative
(summation order only); mixed forces max_rel <= 1e-5 and mixed gradients
<= 1e-6 relative to the FP64 gradient.  The device FP64 gradient is also
checked against central differences (h = 1e-6) of the device FP64 loss."""
import numpy as np
import pytest

from oracle import fitref as FR
from gpu_helpers import force_errors

pytestmark = pytest.mark.gpu

from lammps_amd import batch as B  # noqa: E402
from lammps_amd import workloads as W  # noqa: E402

KA = W.KA_PAIRS


def ka_tables(pairs=KA):
    nt1 = 3
    e, s, c = np.zeros((nt1, nt1)), np.zeros((nt1, nt1)), np.zeros((nt1, nt1))
    for (a, b), (ee, ss, cc) in pairs.items():
        for i, j in ((a, b), (b, a)):
            e[i, j], s[i, j], c[i, j] = ee, ss, cc
    return e, s, c


def ka_batch_inputs(nrep=3, nside=10, seed=1):
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    x0, t0, L = W.ka_initial(nside=nside, nb=int(0.2 * nside ** 3))
    for r in range(nrep):
        xs.append(x0 + rng.normal(scale=0.06, size=x0.shape))
        ts.append(t0)
    n = len(t0)
    return [n] * nrep, [[0, 0, 0]] * nrep, [[L, L, L]] * nrep, [[1, 1, 1]] * nrep, np.concatenate(ts), np.concatenate(xs)


@pytest.fixture(scope="module")
def ka():
    natoms, lo, hi, per, t, x = ka_batch_inputs()
    ref = FR.Batch(natoms, lo, hi, per, 2, t, x)
    e, s, c = ka_tables()
    f_ref, en_ref = ref.lj_forces(e, s, c)
    G = np.random.default_rng(7).normal(size=x.shape)
    ge_ref, gs_ref = ref.lj_vjp(e, s, c, G)
    return dict(natoms=natoms, lo=lo, hi=hi, per=per, t=t, x=x, f=f_ref, e=en_ref, G=G, ge=ge_ref, gs=gs_ref)


def make_lj(ka, precision):
    b = B.Batch(ka["natoms"], ka["lo"], ka["hi"], ka["per"], 2, ka["t"])
    b.set_positions(ka["x"])
    b.precision(precision)
    b.lj_init(*ka_tables())
    return b


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_lj_batch_forces_fp64_and_mixed(ka):
    for prec, tol in ((B.FP64, 1e-11), (B.MIXED, 1e-5)):
        b = make_lj(ka, prec)
        f, e = b.lj_forces()
        mx, _ = force_errors(f, ka["f"])
        assert mx <= tol, (prec, mx)
        assert np.max(np.abs(e - ka["e"]) / np.abs(ka["e"])) <= (1e-12 if prec == B.FP64 else 1e-6)
        b.close()


def test_lj_batch_gradient_fp64_and_mixed(ka):
    for prec, tol in ((B.FP64, 1e-9), (B.MIXED, 1e-6)):
        b = make_lj(ka, prec)
        ge, gs = b.lj_vjp(ka["G"])
        for g, gr in ((ge, ka["ge"]), (gs, ka["gs"])):
            for a, c in ((1, 1), (1, 2), (2, 2)):
                assert abs(g[a, c] - gr[a, c]) <= tol * abs(gr[a, c]), (prec, a, c, g[a, c], gr[a, c])
        b.close()


def test_lj_batch_gradient_matches_device_finite_differences(ka):
    b = make_lj(ka, B.FP64)
    ge, gs = b.lj_vjp(ka["G"])
    e0, s0, c0 = ka_tables()
    for which, g in ((0, ge), (1, gs)):
        for a, c in ((1, 1), (1, 2), (2, 2)):
            vals = []
            base = (e0, s0)[which][a, c]
            h = 1e-6 * base
            for sgn in (1, -1):
                tabs = [e0.copy(), s0.copy()]
                tabs[which][a, c] += sgn * h
                tabs[which][c, a] = tabs[which][a, c]
                b.lj_init(tabs[0], tabs[1], c0)
                vals.append(np.sum(ka["G"] * b.lj_forces(energy=False)[0]))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(g[a, c] - fd) <= 1e-6 * abs(fd), (which, a, c, g[a, c], fd)
    b.close()


def test_autotuned_launch_gives_the_same_results(ka):
    b = make_lj(ka, B.FP64)
    f0, _ = b.lj_forces()
    ge0, gs0 = b.lj_vjp(ka["G"])
    b.autotune(True)
    f1, _ = b.lj_forces()
    ge1, gs1 = b.lj_vjp(ka["G"])
    assert force_errors(f1, f0)[0] <= 1e-12
    assert rel(ge1[1:, 1:], ge0[1:, 1:]) <= 1e-12 and rel(gs1[1:, 1:], gs0[1:, 1:]) <= 1e-12
    for kind in (B.BK_LJ, B.BK_LJ_VJP):
        sh = b.get_launch(kind)
        assert sh["tpa"] in (2, 4, 8, 16) and sh["block"] in (128, 256, 512)
    # explicit shapes, including a tiny grid that loops over many tiles
    for tpa, block, grid in ((1, 64, 1), (32, 1024, 0), (4, 128, 3)):
        b.set_launch(B.BK_LJ, block, tpa, grid)
        b.set_launch(B.BK_LJ_VJP, block, tpa, grid)
        assert force_errors(b.lj_forces()[0], f0)[0] <= 1e-12
        ge2, _ = b.lj_vjp(ka["G"])
        assert rel(ge2[1:, 1:], ge0[1:, 1:]) <= 1e-12
    b.close()


@pytest.fixture(scope="module")
def eam():
    pot = W.synthetic_knots(64)
    xs = [W.fcc_cluster(n) for n in (200, 431, 1000)]
    rng = np.random.default_rng(3)
    xs = [x + rng.normal(scale=0.12, size=x.shape) for x in xs]
    lo = [x.min(0) - 2.0 for x in xs]
    hi = [x.max(0) + 2.0 for x in xs]
    # plus a periodic crystal replica
    a = 3.615
    ncell = 4
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*(np.arange(ncell),) * 3, indexing="ij"), -1).reshape(-1, 3)
    xp = ((g[:, None, ::-1] + basis[None]) * a).reshape(-1, 3) + rng.normal(scale=0.1, size=(4 * ncell ** 3, 3))
    L = ncell * a
    xs.append(xp)
    lo.append([0, 0, 0])
    hi.append([L, L, L])
    per = [[0, 0, 0]] * 3 + [[1, 1, 1]]
    natoms = [len(x) for x in xs]
    x = np.concatenate(xs)
    ref = FR.Batch(natoms, lo, hi, per, 1, np.ones(len(x)), x)
    ref.eam_set(pot["drho"], pot["dr"], pot["cut"], pot["frho"], pot["rhor"], pot["z2r"])
    f, e, rho, fp = ref.eam_forces()
    G = np.random.default_rng(11).normal(size=x.shape)
    g = ref.eam_vjp(G)
    return dict(pot=pot, natoms=natoms, lo=lo, hi=hi, per=per, x=x, f=f, e=e, rho=rho, fp=fp, G=G, g=g)


def make_eam(d, precision):
    b = B.Batch(d["natoms"], d["lo"], d["hi"], d["per"], 1, np.ones(len(d["x"]), np.int32))
    b.set_positions(d["x"])
    b.precision(precision)
    p = d["pot"]
    b.eam_init(p["drho"], p["dr"], p["cut"], p["frho"], p["rhor"], p["z2r"])
    return b


def test_eam_batch_forces_fp64_and_mixed(eam):
    for prec, tol in ((B.FP64, 1e-11), (B.MIXED, 1e-5)):
        b = make_eam(eam, prec)
        f, e, rho, fp = b.eam_forces(rho_fp=True)
        assert force_errors(f, eam["f"])[0] <= tol, prec
        assert np.max(np.abs(e - eam["e"]) / np.abs(eam["e"])) <= (1e-12 if prec == B.FP64 else 1e-6)
        assert rel(rho, eam["rho"]) <= (1e-12 if prec == B.FP64 else 1e-6)
        b.close()


def test_eam_batch_knot_gradient_fp64_and_mixed(eam):
    for prec, tol in ((B.FP64, 1e-9), (B.MIXED, 1e-6)):
        b = make_eam(eam, prec)
        b.eam_forces()
        g = b.eam_vjp(eam["G"])
        for gd, gr in zip(g, eam["g"]):
            assert np.max(np.abs(gd - gr)) <= tol * np.max(np.abs(gr)), (prec, np.max(np.abs(gd - gr)))
        # a vjp without a preceding forward on these tables recomputes rho/fp
        b2 = make_eam(eam, prec)
        g2 = b2.eam_vjp(eam["G"])
        for gd, gr in zip(g2, eam["g"]):
            assert np.max(np.abs(gd - gr)) <= tol * np.max(np.abs(gr))
        b.close()
        b2.close()


def test_eam_batch_gradient_matches_device_finite_differences(eam):
    b = make_eam(eam, B.FP64)
    g = b.eam_vjp(eam["G"])
    p = eam["pot"]
    knots = [p["frho"], p["rhor"], p["z2r"]]
    rng = np.random.default_rng(2)
    for which in range(3):
        nz = np.nonzero(np.abs(g[which]) > 1e-3 * np.max(np.abs(g[which])))[0]
        for k in rng.choice(nz, size=min(8, len(nz)), replace=False):
            h = 1e-6 * max(abs(knots[which][k]), 1e-3)
            vals = []
            for sgn in (1, -1):
                kk = [a.copy() for a in knots]
                kk[which][k] += sgn * h
                b.eam_init(p["drho"], p["dr"], p["cut"], *kk)
                vals.append(np.sum(eam["G"] * b.eam_forces(energy=False)[0]))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(g[which][k] - fd) <= 1e-6 * np.max(np.abs(g[which])), (which, k, g[which][k], fd)
    b.close()


def test_torch_autograd_functions_match_the_adjoint(ka, eam):
    import torch

    from lammps_amd import fit

    dev = torch.device("cuda", 0)
    eps, sig, cut = W.ka_pair_vectors()
    fm = fit.LJForceMatch(ka["natoms"], ka["lo"], ka["hi"], ka["per"], 2, ka["t"], ka["x"], cut, precision=B.FP64)
    e_t = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    s_t = torch.tensor(sig, dtype=torch.float64, requires_grad=True)
    F = fm(e_t, s_t)
    assert F.is_cuda and F.shape == (len(ka["t"]), 3)
    assert force_errors(F.detach().cpu().numpy(), ka["f"])[0] <= 1e-11
    G = torch.tensor(ka["G"], device=dev)
    (F * G).sum().backward()
    for got, ref in ((e_t.grad.numpy(), ka["ge"]), (s_t.grad.numpy(), ka["gs"])):
        for v, (a, c) in zip(got, ((1, 1), (1, 2), (2, 2))):
            assert abs(v - ref[a, c]) <= 1e-9 * abs(ref[a, c])
    p = eam["pot"]
    em = fit.EAMForceMatch(eam["natoms"], eam["lo"], eam["hi"], eam["per"], eam["x"], p["drho"], p["dr"], p["cut"],
                           precision=B.FP64)
    kn = [torch.tensor(p[k], dtype=torch.float64, requires_grad=True) for k in ("frho", "rhor", "z2r")]
    F2 = em(*kn)
    (F2 * torch.tensor(eam["G"], device=dev)).sum().backward()
    for t, ref in zip(kn, eam["g"]):
        assert np.max(np.abs(t.grad.numpy() - ref)) <= 1e-9 * np.max(np.abs(ref))
    fm.batch.close()
    em.batch.close()


def test_neighbour_counts_match_brute_force_small_boxes():
    # periodic box of 2.1 cutoffs with few bins per dimension, and a cluster
    rng = np.random.default_rng(5)
    L = 2.1 * 2.5
    x1 = rng.uniform(0, L, size=(60, 3))
    x2 = rng.uniform(0, 6.0, size=(40, 3))
    x = np.concatenate([x1, x2])
    t = np.ones(100, np.int32)
    b = B.Batch([60, 40], [[0, 0, 0], [0, 0, 0]], [[L, L, L], [6, 6, 6]], [[1, 1, 1], [0, 0, 0]], 1, t)
    b.set_positions(x)
    one = np.ones((2, 2))
    b.lj_init(one, one * 0.5, one * 2.5)
    total, _ = b.neigh_stats()
    cnt = 0
    for i in range(60):
        d = x1 - x1[i]
        d -= L * np.round(d / L)
        r2 = np.sum(d * d, 1)
        cnt += int(np.sum(r2 < 2.5 ** 2)) - 1
    for i in range(40):
        r2 = np.sum((x2 - x2[i]) ** 2, 1)
        cnt += int(np.sum(r2 < 2.5 ** 2)) - 1
    assert total == cnt
    b.close()
