"""Seeded synthetic inputs for the force-matching workloads.

KA (binary Lennard-Jones, 8,000 atoms): 20^3 simple-cubic sites in a cubic
box at density 1.2, 1,600 sites made type 2 by a Fisher-Yates shuffle
driven by RanPark (seed 31415), Gaussian velocities at T* = 0.8 (seed
27183, zero momentum), then a device MD run: 1,000 steps at dt 0.001,
10,000 steps at dt 0.005 with velocity rescaling to 0.8 every 100 steps,
then 64 snapshots every 500 steps.  Target forces come from parameters
theta_t = theta (1 + u), u ~ U[-0.05, 0.05] drawn with RanPark (seed 5)
in the order eps_11, sigma_11, eps_12, sigma_12, eps_22, sigma_22.

EAM clusters (512): RanPark seed 7 draws, per cluster k in order,
N_k = 200 + floor(4801 u) and T_k = 300 + 900 u; each cluster is the N_k
fcc sites (a = 3.615) nearest the origin (ties by site index), given
Gaussian velocities at T_k (seed 7 + k, zero momentum) and run for a short
NVE thermalisation on the device (non-periodic box with a vacuum margin).
The potential is the synthetic funcfl table (seed 4951) on 64 knots per
function.
"""
import os
import tempfile

import numpy as np

from . import mdx

KA_PAIRS = {(1, 1): (1.0, 1.0, 2.5), (1, 2): (1.5, 0.8, 2.0), (2, 2): (0.5, 0.88, 2.2)}


class RanPark:
    """RanPark::uniform (random_park.cpp:42-54) through the library's host code."""

    def __init__(self, seed):
        self.state = np.array([seed], dtype=np.int32)

    def uniform(self):
        return mdx._host_lib().mdx_host_ranpark_uniform(self.state.ctypes.data_as(mdx._ip))


def ka_initial(nside=20, density=1.2, nb=1600, seed=31415):
    """Positions (z-outer, y, x order), types and box length of the KA start."""
    n = nside ** 3
    L = (n / density) ** (1.0 / 3.0)
    a = L / nside
    k, j, i = np.meshgrid(np.arange(nside), np.arange(nside), np.arange(nside), indexing="ij")
    x = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64) * a
    perm = np.arange(n)
    rng = RanPark(seed)
    for m in range(n - 1, 0, -1):
        r = min(int(rng.uniform() * (m + 1)), m)
        perm[m], perm[r] = perm[r], perm[m]
    types = np.ones(n, dtype=np.int32)
    types[perm[:nb]] = 2
    return x, types, L


def ka_pair_vectors(pairs=KA_PAIRS):
    keys = [(1, 1), (1, 2), (2, 2)]
    eps = np.array([pairs[k][0] for k in keys])
    sig = np.array([pairs[k][1] for k in keys])
    cut = np.array([pairs[k][2] for k in keys])
    return eps, sig, cut


def ka_target_vectors(seed=5, amp=0.05):
    eps, sig, _ = ka_pair_vectors()
    rng = RanPark(seed)
    e_t, s_t = eps.copy(), sig.copy()
    for k in range(3):
        e_t[k] *= 1.0 + amp * (2.0 * rng.uniform() - 1.0)
        s_t[k] *= 1.0 + amp * (2.0 * rng.uniform() - 1.0)
    return e_t, s_t


def ka_snapshots(nsnap=64, every=500, quench_steps=1000, equil_steps=10000, rescale_every=100, temp=0.8,
                 vseed=27183, device=0, nside=20):
    """Runs the KA preparation on the device and returns (snapshots [nsnap, N, 3], types, L)."""
    x, types, L = ka_initial(nside=nside, nb=int(0.2 * nside ** 3))
    mass = np.array([0.0, 1.0, 1.0])
    v = mdx.velocity_create(x, types, mass, temp, vseed, dist="gaussian", loop="all", mom=True)
    ctx = mdx.Context(device)
    ctx.set_box((0, 0, 0), (L, L, L))
    ctx.pair_lj_from_coeffs(2, KA_PAIRS)
    ctx.neigh_modify(0.3, every=1, delay=0, check=True)
    ctx.precision(mdx.FP64)

    def rescale(v):
        ke = 0.5 * np.sum(v ** 2)
        t_now = 2.0 * ke / (3 * len(v) - 3)
        return v * np.sqrt(temp / t_now)

    ctx.set_atoms(x, types, mass, v=v)
    ctx.setup()
    ctx.run(quench_steps, 0.001)
    for _ in range(equil_steps // rescale_every):
        x, v, _, _ = ctx.get_atoms()
        ctx.set_atoms(x, types, mass, v=rescale(v))
        ctx.setup()
        ctx.run(rescale_every, 0.005)
    snaps = np.zeros((nsnap, len(types), 3))
    for s in range(nsnap):
        ctx.run(every, 0.005)
        snaps[s] = ctx.get_atoms()[0]
    ctx.close()
    return snaps, types, L


def synthetic_knots(nknots=64, cut=4.95, a=3.615, seed=4951):
    """Knot values (frho, rhor, z2r), spacings and cutoff of the synthetic
    funcfl potential on nknots points per function (the file holds
    nknots + 1 points; PairEAM keeps n - 1, pair_eam.cpp:497-502)."""
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "synth.eam")
        dr = cut / (nknots - 1)
        params = mdx.write_synthetic_funcfl(path, seed=seed, a=a, nrho=nknots + 1, nr=nknots + 1, dr=dr, cut=cut)
        t = mdx.eam_funcfl(path)
    knots = (t["frho_spline"][0, 1:, 6].copy(), t["rhor_spline"][0, 1:, 6].copy(), t["z2r_spline"][0, 1:, 6].copy())
    return dict(frho=knots[0], rhor=knots[1], z2r=knots[2], drho=t["drho"], dr=t["dr"], cut=t["cutmax"],
                mass=t["mass"], params=params, path_tables=t)


def cluster_sizes(nrep=512, seed=7):
    rng = RanPark(seed)
    n, temp = [], []
    for _ in range(nrep):
        n.append(200 + int(4801 * rng.uniform()))
        temp.append(300.0 + 900.0 * rng.uniform())
    return np.array(n), np.array(temp)


def fcc_cluster(n, a=3.615):
    """The n fcc sites nearest the origin, ties broken by site index (z, y, x, basis order)."""
    m = int(np.ceil((n / 4.0) ** (1.0 / 3.0))) + 2
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    r = np.arange(-m, m + 1)
    k, j, i = np.meshgrid(r, r, r, indexing="ij")
    cells = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    sites = ((cells[:, None, :] + basis[None]) * a).reshape(-1, 3)
    d2 = np.round(np.sum(sites ** 2, axis=1), 9)
    order = np.lexsort((np.arange(len(sites)), d2))
    return sites[order[:n]].copy()


def eam_clusters(nrep=512, thermal_steps=100, dt=0.005, margin=8.0, nknots=64, device=0, seed=7):
    """C5 inputs: list of cluster positions after thermalisation, their boxes,
    and the knot potential."""
    pot = synthetic_knots(nknots)
    sizes, temps = cluster_sizes(nrep, seed)
    tab = pot["path_tables"]
    ctx = mdx.Context(device) if thermal_steps > 0 else None
    xs, lo, hi = [], [], []
    for k in range(nrep):
        x = fcc_cluster(int(sizes[k]))
        types = np.ones(len(x), dtype=np.int32)
        mass = np.array([0.0, pot["mass"]])
        v = mdx.velocity_create(x, types, mass, float(temps[k]), seed + k, dist="gaussian", loop="all", mom=True,
                                units="metal")
        blo, bhi = x.min(0) - margin, x.max(0) + margin
        if ctx is not None:
            ctx.set_box(blo, bhi, periodic=(0, 0, 0))
            ctx.pair_eam(1, tab["nrho"], tab["drho"], tab["nr"], tab["dr"], tab["rhomax"], tab["cutmax"] ** 2,
                         tab["frho_spline"], tab["rhor_spline"], tab["z2r_spline"], np.zeros(2, np.int32),
                         np.zeros(4, np.int32), np.zeros(4, np.int32))
            ctx.neigh_modify(1.0, every=1, delay=0, check=True)
            ctx.set_atoms(x, types, mass, v=v)
            ctx.setup()
            ctx.run(thermal_steps, dt, ftm2v=mdx.UNIT_CONSTANTS["metal"]["ftm2v"])
            x = ctx.get_atoms()[0]
        xs.append(x)
        lo.append(blo)
        hi.append(bhi)
    if ctx is not None:
        ctx.close()
    return dict(x=xs, lo=np.array(lo), hi=np.array(hi), sizes=sizes, temps=temps, pot=pot)


def perturbed_knots(pot, seed=5, amp=0.05):
    """Target potential: every knot scaled by 1 + amp (2u - 1), RanPark(seed), frho then rhor then z2r."""
    rng = RanPark(seed)
    out = []
    for name in ("frho", "rhor", "z2r"):
        k = pot[name].copy()
        for m in range(len(k)):
            k[m] *= 1.0 + amp * (2.0 * rng.uniform() - 1.0)
        out.append(k)
    return out
