"""Pins the CPU oracle (oracle/mdref.c) before anything is compared with it.

Each check is chosen so that a plausible restatement mistake fails it:
golden thermo rows and neighbour statistics from the reference's own logs
(bench/log.1Feb14.{lj,eam}.fixed.linux.1, examples/melt/log.melt.1Feb14.linux.1),
the Park-Miller check value, closed forms, O(N^2) brute force, finite
differences of the energy, conservation laws and exact polynomial
reproduction of the EAM splines.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import mdref as M

with open(os.path.join(GOLDEN, "goldens.json")) as fh:
    GOLD = json.load(fh)
CU_U3 = os.path.join(GOLDEN, "Cu_u3.eam")


def fmt_row(row):
    """LAMMPS thermo format: %8d then %12.8g (thermo.cpp defaults)."""
    return ["%d" % int(row[0])] + [("%12.8g" % v).strip() for v in row[1:]]


def lj_bench(n=20, steps=100, T=1.44):
    r = M.Ref(M.LJ)
    r.lattice_box(M.FCC, 0.8442, (0, 0, 0), (n, n, n))
    r.mass(1, 1.0)
    r.velocity_create(T, 87287, dist="uniform", loop="geom")
    r.pair_lj(2.5)
    r.pair_coeff(1, 1, 1.0, 1.0, 2.5)
    r.neighbor(0.3, every=20, delay=0, check=False)
    r.timestep(0.005)
    r.setup()
    return r, r.run(steps)


def test_lj_bench_golden_rows_and_neighbours():
    g = GOLD["lj_bench_32k"]
    r, rows = lj_bench()
    assert [fmt_row(x) for x in rows] == g["thermo_text"]
    st = r.neigh_stats()
    assert st["total"] == g["total_neighbors"] == 1202833
    assert st["builds"] == g["builds"] == 5
    assert r.nghost == g["nghost"]


def test_eam_bench_golden_rows_and_neighbours():
    g = GOLD["eam_bench_32k"]
    r = M.Ref(M.METAL)
    r.lattice_box(M.FCC, 3.615, (0, 0, 0), (20, 20, 20))
    r.pair_eam(CU_U3)
    r.velocity_create(1600.0, 376847, dist="uniform", loop="geom")
    r.neighbor(1.0, every=1, delay=5, check=True)
    r.timestep(0.005)
    r.thermo(50)
    r.setup()
    rows = r.run(100)
    assert [fmt_row(x) for x in rows] == g["thermo_text"]
    st = r.neigh_stats()
    assert st["total"] == g["total_neighbors"]
    assert st["builds"] == g["builds"]
    assert st["dangerous"] == g["dangerous"]
    assert r.nghost == g["nghost"]


def test_melt_golden_loop_all_velocities():
    g = GOLD["melt_4k"]
    r = M.Ref(M.LJ)
    r.lattice_box(M.FCC, 0.8442, (0, 0, 0), (10, 10, 10))
    r.mass(1, 1.0)
    r.velocity_create(3.0, 87287)
    r.pair_lj(2.5)
    r.pair_coeff(1, 1, 1.0, 1.0, 2.5)
    r.neighbor(0.3, every=20, delay=0, check=False)
    r.thermo(50)
    r.setup()
    rows = r.run(250)
    assert [fmt_row(x) for x in rows] == g["thermo_text"]
    st = r.neigh_stats()
    assert st["total"] == g["total_neighbors"] and st["builds"] == g["builds"]


def test_ranpark_minimal_standard_check_value():
    # Park & Miller (1988): seed 1 -> 1043618065 after 10,000 draws
    u, state = M.ranpark_uniform_seq(1, 10000)
    assert state == 1043618065
    assert u == pytest.approx(1043618065 / 2147483647, rel=1e-15)


def test_ranpark_reset_depends_on_all_coordinate_bytes():
    a = M.ranpark_reset_coord(87287, [1.0, 2.0, 3.0])
    b = M.ranpark_reset_coord(87287, [1.0, 2.0, np.nextafter(3.0, 4.0)])
    c = M.ranpark_reset_coord(87288, [1.0, 2.0, 3.0])
    assert len({a, b, c}) == 3 and all(0 < s < 2 ** 31 - 1 for s in (a, b, c))


def test_velocity_create_temperature_and_momentum():
    for dist in ("uniform", "gaussian"):
        for loop in ("all", "geom"):
            r = M.Ref(M.LJ)
            r.lattice_box(M.FCC, 0.8442, (0, 0, 0), (4, 4, 4))
            r.mass(1, 1.0)
            r.velocity_create(1.44, 87287, dist=dist, loop=loop)
            _, v, _, _, _ = r.state()
            p = v.sum(axis=0)
            assert np.abs(p).max() < 1e-12
            T = (v * v).sum() / (3 * len(v) - 3)
            assert T == pytest.approx(1.44, rel=1e-14)


def random_gas(n, L, seed, rmin=0.9):
    """Random configuration with no pair closer than rmin (min image)."""
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, L, 3)
        ok = True
        for q in pts:
            d = p - q
            d -= L * np.round(d / L)
            if d @ d < rmin * rmin:
                ok = False
                break
        if ok:
            pts.append(p)
    return np.array(pts)


def lj_system(x, L, cut=2.5, skin=0.3, list_style=M.HALF_NEWTON, types=None, coeffs=None, shift=False):
    r = M.Ref(M.LJ)
    ntypes = 1 if types is None else int(max(types))
    r.box((0, 0, 0), (L, L, L), ntypes=ntypes)
    r.add_atoms(x, np.ones(len(x), dtype=np.int32) if types is None else types)
    for t in range(1, ntypes + 1):
        r.mass(t, 1.0)
    r.pair_lj(cut, shift=shift)
    if coeffs is None:
        r.pair_coeff(1, 1, 1.0, 1.0, cut)
    else:
        for (i, j), (e, s, c) in coeffs.items():
            r.pair_coeff(i, j, e, s, c)
    r.neighbor(skin, every=1, delay=0, check=True)
    r.list_style(list_style)
    r.sortfreq(0)
    return r


@pytest.mark.parametrize("style", [M.HALF_NEWTON, M.HALF_NONEWTON, M.FULL])
def test_neighbor_list_equals_brute_force(style):
    L = 7.3
    x = random_gas(220, L, 3, rmin=0.8)
    r = lj_system(x, L, list_style=style)
    r.compute(1, 1)
    pairs = r.list_pairs()
    got = {(min(a, b), max(a, b)) for a, b in pairs}
    cut2 = 2.8 ** 2
    want = set()
    for i in range(len(x)):
        d = x[i + 1:] - x[i]
        d -= L * np.round(d / L)
        rsq = (d * d).sum(axis=1)
        for k in np.nonzero(rsq <= cut2)[0]:
            want.add((i + 1, i + 2 + int(k)))
    assert got == want
    if style == M.FULL:
        assert len(pairs) == 2 * len(want)
    elif style == M.HALF_NEWTON:
        assert len(pairs) == len(want)  # each pair stored exactly once
    else:
        # no newton: owned-ghost pairs are stored by both owners
        assert len(pairs) > len(want)


def test_lj_two_atoms_closed_form():
    for rr in (0.95, 1.12, 1.7, 2.49):
        x = np.array([[5.0, 5.0, 5.0], [5.0 + rr, 5.0, 5.0]])
        r = lj_system(x, 12.0)
        r.compute(1, 1)
        _, _, f, _, _ = r.state()
        e, w = r.energy_virial()
        sr6 = (1.0 / rr) ** 6
        fmag = 24.0 * (2 * sr6 * sr6 - sr6) / rr
        assert f[1, 0] == pytest.approx(fmag, rel=1e-13)
        assert f[0, 0] == pytest.approx(-fmag, rel=1e-13)
        assert e == pytest.approx(4.0 * (sr6 * sr6 - sr6), rel=1e-13)
        assert w[0] == pytest.approx(fmag * rr, rel=1e-12)


def test_lj_force_is_minus_energy_gradient_and_sums_to_zero():
    L = 6.5
    x = random_gas(120, L, 11)
    coeffs = {(1, 1): (1.0, 1.0, 2.5), (1, 2): (1.5, 0.8, 2.0), (2, 2): (0.5, 0.88, 2.2)}
    types = np.where(np.arange(len(x)) % 5 == 0, 2, 1).astype(np.int32)
    r = lj_system(x, L, types=types, coeffs=coeffs, shift=True)
    r.compute(1, 1)
    _, _, f, _, _ = r.state()
    assert np.abs(f.sum(axis=0)).max() < 1e-10
    h = 1e-6
    rng = np.random.default_rng(0)
    for a in rng.choice(len(x), 6, replace=False):
        for d in range(3):
            es = []
            for sgn in (1, -1):
                xp = x.copy()
                xp[a, d] += sgn * h
                rp = lj_system(xp, L, types=types, coeffs=coeffs, shift=True)
                rp.compute(1, 0)
                es.append(rp.energy_virial()[0])
            fd = -(es[0] - es[1]) / (2 * h)
            assert fd == pytest.approx(f[a, d], rel=1e-5, abs=1e-6)


def test_lj_virial_pairwise_equals_fdotr_and_peratom():
    L = 6.5
    x = random_gas(100, L, 5)
    ra = lj_system(x, L)           # newton on: fdotr global virial
    ra.compute(1, 2)
    rb = lj_system(x, L, list_style=M.HALF_NONEWTON)
    rb.newton(False)
    rb.compute(3, 5)               # pairwise global + per-atom
    ea, wa = ra.energy_virial()
    eb, wb = rb.energy_virial()
    eat, vat = rb.peratom()
    assert eb == pytest.approx(ea, rel=1e-12)
    np.testing.assert_allclose(wb, wa, rtol=1e-10, atol=1e-10)
    assert eat.sum() == pytest.approx(ea, rel=1e-12)
    np.testing.assert_allclose(vat.sum(axis=0), wa, rtol=1e-10, atol=1e-10)


def write_funcfl(path, nrho, drho, nr, dr, cut, F, Z, R, mass=63.55):
    with open(path, "w") as fh:
        fh.write("synthetic funcfl\n")
        fh.write("29 %.6f 3.615 FCC\n" % mass)
        fh.write("%d %.16e %d %.16e %.16e\n" % (nrho, drho, nr, dr, cut))
        for arr in (F, Z, R):
            for k in range(0, len(arr), 5):
                fh.write(" ".join("%.16e" % v for v in arr[k:k + 5]) + "\n")


def test_eam_spline_reproduces_cubics_in_the_interior(tmp_path):
    nrho, drho, nr, dr = 200, 0.05, 300, 0.02
    rho = np.arange(nrho) * drho
    rr = np.arange(nr) * dr
    Fp = lambda q: 0.3 - 0.2 * q + 0.01 * q ** 2 - 0.0004 * q ** 3
    Zp = lambda q: 2.0 - 0.5 * q + 0.05 * q ** 2 - 0.003 * q ** 3
    Rp = lambda q: 1.0 + 0.1 * q - 0.02 * q ** 2 + 0.001 * q ** 3
    path = str(tmp_path / "cubic.eam")
    write_funcfl(path, nrho, drho, nr, dr, 4.0, Fp(rho), Zp(rr), Rp(rr))
    r = M.Ref(M.METAL)
    r.box((0, 0, 0), (20, 20, 20))
    r.add_atoms(np.array([[1.0, 1.0, 1.0]]), np.array([1], dtype=np.int32))
    r.pair_eam(path)
    r.neighbor(1.0)
    r.compute(1, 0)
    t = r.eam_tables()
    # file2array regrids onto nr = (nr_file-1)*dr/dr points (pair_eam.cpp:511-517)
    assert t["nr"] == nr - 1 and t["nrho"] == nrho - 1
    rng = np.random.default_rng(1)
    for tab, delta, fn in ((t["frho_spline"][0], drho, Fp), (t["rhor_spline"][0], dr, Rp)):
        for m in range(3, len(tab) - 3):
            p = rng.uniform()
            c = tab[m]
            val = ((c[3] * p + c[4]) * p + c[5]) * p + c[6]
            der = (c[0] * p + c[1]) * p + c[2]
            q = (m - 1 + p) * delta
            h = 1e-6
            assert val == pytest.approx(fn(q), rel=1e-12, abs=1e-12)
            assert der == pytest.approx((fn(q + h) - fn(q - h)) / (2 * h), rel=1e-6, abs=1e-8)
    # z2r = 27.2*0.529*Z^2 is degree 6: only check the node values (file2array
    # is the identity on the file grid for one funcfl file)
    z = t["z2r_spline"][0]
    np.testing.assert_allclose(z[1:, 6], 27.2 * 0.529 * Zp(rr[: nr - 1]) ** 2, rtol=1e-14)


def test_eam_force_is_minus_energy_gradient():
    L = 3.615 * 3
    rng = np.random.default_rng(7)
    base = []
    for k in range(3):
        for j in range(3):
            for i in range(3):
                for b in ((0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)):
                    base.append((np.array([i, j, k]) + b) * 3.615)
    x = np.array(base) + rng.normal(0, 0.08, (len(base), 3))

    def sysm(xx):
        r = M.Ref(M.METAL)
        r.box((0, 0, 0), (L, L, L))
        r.add_atoms(xx, np.ones(len(xx), dtype=np.int32))
        r.pair_eam(CU_U3)
        r.neighbor(1.0)
        r.sortfreq(0)
        return r

    r = sysm(x)
    r.compute(1, 0)
    _, _, f, _, _ = r.state()
    assert np.abs(f.sum(axis=0)).max() < 1e-9
    h = 1e-5
    for a in (0, 17, 50, 99):
        for d in range(3):
            es = []
            for sgn in (1, -1):
                xp = x.copy()
                xp[a, d] += sgn * h
                rp = sysm(xp)
                rp.compute(1, 0)
                es.append(rp.energy_virial()[0])
            assert -(es[0] - es[1]) / (2 * h) == pytest.approx(f[a, d], rel=2e-4, abs=2e-5)


def test_nve_free_flight_is_exact_ballistic():
    # atoms beyond the cutoff of each other: f = 0, x(t) = x0 + n*dt*v
    x = np.array([[1.0, 1.0, 1.0], [6.0, 6.0, 6.0], [11.0, 1.0, 6.0]])
    r = lj_system(x, 15.0)
    v = np.array([[0.1, -0.2, 0.3], [0.0, 0.05, -0.01], [-0.3, 0.2, 0.1]])
    r.set_velocities(v)
    r.sortfreq(0)
    r.setup()
    r.run(50)
    xs, vs, _, _, im = r.state()
    unwrapped = xs + im * 15.0
    np.testing.assert_allclose(vs, v, rtol=0, atol=0)
    np.testing.assert_allclose(unwrapped, x + 50 * 0.005 * v, rtol=0, atol=1e-13)


def _threaded_vs_serial(kind, threads):
    out = []
    for threaded in (False, True):
        if threaded:
            assert M.set_threads(threads) == threads
        r = M.Ref(M.LJ if kind == "lj" else M.METAL, threaded=threaded)
        if kind == "lj":
            r.lattice_box(M.FCC, 0.8442, (0, 0, 0), (8, 8, 8))
            r.mass(1, 1.0)
            r.velocity_create(1.44, 87287, dist="gaussian", loop="geom")
            r.pair_lj(2.5)
            r.pair_coeff(1, 1, 1.0, 1.0, 2.5)
            r.neighbor(0.3, every=2, delay=0, check=True)
        else:
            r.lattice_box(M.FCC, 3.615, (0, 0, 0), (6, 6, 6))
            r.pair_eam(os.path.join(GOLDEN, "Cu_u3.eam"))
            r.velocity_create(1600.0, 376847, dist="uniform", loop="geom")
            r.neighbor(1.0, every=1, delay=2, check=True)
        r.timestep(0.005)
        r.setup()
        rows = r.run(30)
        x, v, f, _, _ = r.state()
        out.append((rows, x, f, r.neigh_stats()))
    return out


@pytest.mark.parametrize("kind", ["lj", "eam"])
def test_threaded_baseline_build_follows_the_serial_oracle(kind):
    """The threaded CPU-baseline build (libmdref_omp.so: -O3, 3 threads,
    per-thread force arrays) runs the serial oracle's trajectory: identical
    neighbour totals and build counts, thermo rows and forces within 1e-10
    relative (summation order of the per-thread arrays only)."""
    (rs, xs, fs, ns), (rt, xt, ft, nt) = _threaded_vs_serial(kind, 3)
    assert ns == nt
    assert np.max(np.abs(rs - rt) / np.maximum(1e-300, np.abs(rs)).clip(1e-12)) <= 1e-10
    scale = np.sqrt(np.mean(np.sum(fs * fs, axis=1)))
    assert np.max(np.linalg.norm(ft - fs, axis=1)) <= 1e-10 * scale
    assert np.max(np.abs(xt - xs)) <= 1e-10


def _lehmer(seed, n):
    """The minimal standard generator (Park and Miller 1988) by its
    definition, x <- 16807 x mod (2^31 - 1), in exact integer arithmetic."""
    out, x = [], seed
    for _ in range(n):
        x = (16807 * x) % 2147483647
        out.append(x / 2147483647.0)
    return out


def _polar_pairs(u):
    """Gaussian draws from uniforms u by the polar method, in the order
    random_park.cpp:55-76 returns them: each accepted pair (v1, v2) yields
    v2 * fac first, then v1 * fac."""
    out, k = [], 0
    while k + 1 < len(u):
        v1, v2 = 2.0 * u[k] - 1.0, 2.0 * u[k + 1] - 1.0
        k += 2
        rsq = v1 * v1 + v2 * v2
        if rsq >= 1.0 or rsq == 0.0:
            continue
        fac = np.sqrt(-2.0 * np.log(rsq) / rsq)
        out += [v2 * fac, v1 * fac]
    return out


def test_gaussian_velocity_create_follows_the_published_generator():
    """velocity create dist gaussian loop all (velocity.cpp:207-240) on two
    atoms of equal mass: after mom yes, atom 1's velocity is parallel to
    (g1 - g4, g2 - g5, g3 - g6), the first six draws of the polar method on
    the Park-Miller sequence from the seed.  The expected draws come from the
    generator's definition (exact integers) and the reference's return order;
    swapping the pair's order gives a different direction, so the test
    separates the two."""
    from lammps_amd import mdx
    g = _polar_pairs(_lehmer(1, 40))[:6]
    want = np.array([g[0] - g[3], g[1] - g[4], g[2] - g[5]])
    want /= np.linalg.norm(want)
    swapped = []
    for a in range(0, 6, 2):
        swapped += [g[a + 1], g[a]]
    alt = np.array([swapped[0] - swapped[3], swapped[1] - swapped[4], swapped[2] - swapped[5]])
    alt /= np.linalg.norm(alt)
    assert np.linalg.norm(want - alt) > 1e-3
    x = np.array([[1.0, 1.0, 1.0], [3.0, 3.0, 3.0]])
    t = np.ones(2, dtype=np.int32)
    mass = np.array([0.0, 1.0])
    v_host = mdx.velocity_create(x, t, mass, 1.0, 1, dist="gaussian", loop="all")
    r = M.Ref(M.LJ)
    r.box((0, 0, 0), (5, 5, 5))
    r.add_atoms(x, t)
    r.mass(1, 1.0)
    r.velocity_create(1.0, 1, dist="gaussian", loop="all")
    v_ref = r.state()[1]
    for v in (v_host, v_ref):
        d = v[0] / np.linalg.norm(v[0])
        assert np.allclose(v[0], -v[1], rtol=0, atol=1e-14)
        assert np.linalg.norm(d - want) <= 1e-12
