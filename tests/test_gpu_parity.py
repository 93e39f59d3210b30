"""GPU parity: the HIP path through the C ABI against the CPU FP64 oracle.

Bars (stated here, in the test):
  * neighbour pair sets: identical to the oracle's half_bin_newton set
    (bit-exact integer work), full-list size exactly twice the half list;
  * FP64 path: per-atom forces max_rel <= 1e-11, energy/virial rel <= 1e-11;
  * mixed path: max_rel <= 1e-5 (the force tolerance of the target), energy
    and virial rel <= 1e-6;
  * NVE runs: thermo rows equal to the reference log's printed digits, and
    the reference's 'Total # of neighbors' and build count.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from gpu_helpers import box_of, device_from_oracle, eam_oracle, force_errors, lj_oracle
from oracle import mdref as M

pytestmark = pytest.mark.gpu

A_LJ = (4.0 / 0.8442) ** (1.0 / 3.0)

with open(os.path.join(GOLDEN, "goldens.json")) as fh:
    GOLD = json.load(fh)


def fmt(row):
    return ["%d" % int(row[0])] + [("%12.8g" % v).strip() for v in row[1:]]


@pytest.fixture(scope="module")
def lj_state():
    r = lj_oracle(ncell=20, steps=37, every=20, check=False)
    r.compute(3, 5)
    return r


@pytest.fixture(scope="module")
def eam_state():
    r = eam_oracle(ncell=20, steps=40)
    r.compute(3, 5)
    return r


def test_lj_neighbor_pairs_identical(lj_state):
    from lammps_amd import mdx
    r = lj_state
    lo, hi = box_of(r, 20, A_LJ)
    ctx = device_from_oracle(r, lo, hi, "lj", mdx.FP64, 0.3, 20, 0, False)
    ctx.neigh_build()
    got = ctx.neigh_pairs()
    ref = r.list_pairs()
    ref_set = {(min(a, b), max(a, b)) for a, b in ref}
    got_set = {(int(a), int(b)) for a, b in got}
    assert len(got) == len(got_set)
    assert got_set == ref_set
    st = ctx.stats()
    assert st["total_neighbors"] == 2 * len(ref_set)


def precision_of(prec, ctx=None):
    """fp64 | mixed (fixed-point neighbour positions, the default) | mixed_fp64pos"""
    from lammps_amd import mdx
    p = mdx.FP64 if prec == "fp64" else mdx.MIXED
    if ctx is not None:
        ctx.position_format(mdx.POS_FP64 if prec == "mixed_fp64pos" else mdx.POS_Q32)
    return p


@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed_fp64pos"])
def test_lj_forces_energy_virial(lj_state, prec):
    r = lj_state
    lo, hi = box_of(r, 20, A_LJ)
    ctx = device_from_oracle(r, lo, hi, "lj", precision_of(prec), 0.3, 20, 0, False)
    precision_of(prec, ctx)
    out = ctx.compute(3, 5, peratom=True)
    _, _, fref, _, _ = r.state()
    eref, wref = r.energy_virial()
    ea_ref, va_ref = r.peratom()
    mx, rms = force_errors(out["f"], fref)
    tol = 1e-11 if prec == "fp64" else 1e-5
    etol = 1e-11 if prec == "fp64" else 1e-6
    print("LJ %s force max_rel=%.3e rms_rel=%.3e" % (prec, mx, rms))
    assert mx <= tol
    assert abs(out["eng"] - eref) <= etol * abs(eref)
    np.testing.assert_allclose(out["virial"], wref, rtol=etol, atol=etol * np.abs(wref).max())
    np.testing.assert_allclose(out["eatom"], ea_ref, rtol=0, atol=100 * etol * np.abs(ea_ref).max())
    np.testing.assert_allclose(out["vatom"], va_ref, rtol=0, atol=100 * etol * np.abs(va_ref).max())


@pytest.mark.parametrize("style", ["lj", "eam"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_global_virial_without_per_atom_virial_matches_the_oracle(lj_state, eam_state, style, prec):
    """Without a per-atom virial output the global virial is formed as the
    reference's virial_fdotr_compute does (sum_i x_i f_i plus the shifted
    images' share), with one: as the per-entry tally.  Both equal the
    oracle's virial (FP64 1e-11, mixed 1e-6 relative) and each other
    (FP64 1e-12)."""
    from lammps_amd import mdx
    r = lj_state if style == "lj" else eam_state
    lo, hi = box_of(r, 20, A_LJ if style == "lj" else 3.615)
    ctx = device_from_oracle(r, lo, hi, style, precision_of(prec), 0.3 if style == "lj" else 1.0, 20, 0, False)
    precision_of(prec, ctx)
    tally = ctx.compute(1, 1, peratom=True)["virial"]
    fdotr = ctx.compute(1, 1)["virial"]
    _, wref = r.energy_virial()
    tol = 1e-11 if prec == "fp64" else 1e-6
    for w in (tally, fdotr):
        np.testing.assert_allclose(w, wref, rtol=tol, atol=tol * np.abs(wref).max())
    if prec == "fp64":
        np.testing.assert_allclose(fdotr, tally, rtol=1e-12, atol=1e-12 * np.abs(tally).max())
    ctx.close()


@pytest.mark.parametrize("tpa", [1, 2, 4, 8, 16, 32])
def test_lj_threads_per_atom_variants_agree(lj_state, tpa):
    """Every pair-kernel shape gives the oracle's forces and pair set."""
    from lammps_amd import mdx
    r = lj_state
    lo, hi = box_of(r, 20, A_LJ)
    ctx = device_from_oracle(r, lo, hi, "lj", mdx.FP64, 0.3, 20, 0, False)
    ctx.set_launch(128 if tpa in (2, 16) else 256, tpa)
    out = ctx.compute(1, 1)
    _, _, fref, _, _ = r.state()
    eref, _ = r.energy_virial()
    mx, _ = force_errors(out["f"], fref)
    assert mx <= 1e-11
    assert abs(out["eng"] - eref) <= 1e-11 * abs(eref)
    got = {(int(a), int(b)) for a, b in ctx.neigh_pairs()}
    ref = {(min(a, b), max(a, b)) for a, b in r.list_pairs()}
    assert got == ref


def test_eam_neighbor_pairs_identical(eam_state):
    from lammps_amd import mdx
    r = eam_state
    lo, hi = box_of(r, 20, 3.615)
    ctx = device_from_oracle(r, lo, hi, "eam", mdx.FP64, 1.0, 1, 5, True)
    ctx.neigh_build()
    got = {(int(a), int(b)) for a, b in ctx.neigh_pairs()}
    ref = {(min(a, b), max(a, b)) for a, b in r.list_pairs()}
    assert got == ref


@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed_fp64pos"])
def test_eam_forces_energy_virial(eam_state, prec):
    r = eam_state
    lo, hi = box_of(r, 20, 3.615)
    ctx = device_from_oracle(r, lo, hi, "eam", precision_of(prec), 1.0, 1, 5, True)
    precision_of(prec, ctx)
    out = ctx.compute(3, 5, peratom=True)
    _, _, fref, _, _ = r.state()
    eref, wref = r.energy_virial()
    rho_ref, fp_ref = r.eam_rho_fp()
    rho, fp = ctx.eam_rho_fp()
    mx, rms = force_errors(out["f"], fref)
    tol = 1e-11 if prec == "fp64" else 1e-5
    etol = 1e-11 if prec == "fp64" else 1e-6
    print("EAM %s force max_rel=%.3e rms_rel=%.3e" % (prec, mx, rms))
    assert mx <= tol
    np.testing.assert_allclose(rho, rho_ref, rtol=etol)
    np.testing.assert_allclose(fp, fp_ref, rtol=etol * 10, atol=etol)
    assert abs(out["eng"] - eref) <= etol * abs(eref)
    np.testing.assert_allclose(out["virial"], wref, rtol=etol, atol=etol * np.abs(wref).max())


@pytest.mark.parametrize("style", ["lj", "eam"])
def test_fused_integration_is_bit_identical_to_separate_kernels(lj_state, eam_state, style):
    """mdx_run with fix nve fused into the force kernel reproduces the
    separate initial/force/final kernels bit for bit over 60 steps with
    rebuilds, and thermo rows agree."""
    from lammps_amd import mdx
    out = {}
    for fused in (True, False):
        if style == "lj":
            r = lj_state
            lo, hi = box_of(r, 20, A_LJ)
            ctx = device_from_oracle(r, lo, hi, "lj", mdx.MIXED, 0.3, 5, 0, True)
            ftm2v = 1.0
        else:
            r = eam_state
            lo, hi = box_of(r, 20, 3.615)
            ctx = device_from_oracle(r, lo, hi, "eam", mdx.MIXED, 1.0, 1, 2, True)
            ftm2v = 1.0 / 1.0364269e-4
        ctx.fused(fused)
        rows = ctx.setup() + ctx.run(37, 0.005, ftm2v, 10) + ctx.run(23, 0.005, ftm2v, 10)
        x, v, f, im = ctx.get_atoms()
        out[fused] = dict(x=x, v=v, f=f, im=im, rows=rows, st=ctx.stats())
        ctx.close()
    for k in ("x", "v", "f", "im"):
        assert np.array_equal(out[True][k], out[False][k]), k
    assert out[True]["st"]["ncalls"] == out[False]["st"]["ncalls"] > 0
    for (s1, r1), (s2, r2) in zip(out[True]["rows"], out[False]["rows"]):
        assert s1 == s2
        np.testing.assert_allclose(r1, r2, rtol=1e-13, atol=1e-13 * max(abs(q) for q in r2))


def test_look_ahead_decide_relaunches_the_steps_it_skipped(eam_state):
    """check yes, every 1, delay 2, no thermo output inside the run: from
    the second step of a run on, a step's kernels are queued before the host
    has read the displacement trigger they depend on.  A step whose trigger
    fired does nothing on the device and is queued again after the rebuild,
    so each of rho and fused force+integrate is launched once per step plus
    once per such rebuild (all but the builds of the run's first and last
    steps), and the run equals the unfused path bit for bit."""
    from lammps_amd import mdx
    r = eam_state
    lo, hi = box_of(r, 20, 3.615)
    out = {}
    for fused in (True, False):
        ctx = device_from_oracle(r, lo, hi, "eam", mdx.FP64, 1.0, 1, 2, True)
        ctx.fused(fused)
        ctx.setup()
        ctx.profiling(True)
        k0 = ctx.stats()["kernel_launches"]
        ctx.run(60, 0.005, 1.0 / 1.0364269e-4, 0)
        st = ctx.stats()
        k = [a - b for a, b in zip(st["kernel_launches"], k0)]
        out[fused] = dict(atoms=ctx.get_atoms(), builds=st["ncalls"], k=k)
        ctx.close()
    for a, b in zip(out[True]["atoms"], out[False]["atoms"]):
        assert np.array_equal(a, b)
    builds = out[True]["builds"]
    assert builds == out[False]["builds"] >= 3
    k = out[True]["k"]
    relaunched = k[mdx.K_FUSED] - 60
    assert builds - 2 <= relaunched <= builds and relaunched >= 1
    assert k[mdx.K_EAM_RHO] == k[mdx.K_FUSED]
    assert out[False]["k"][mdx.K_EAM_FORCE] == 60


@pytest.mark.parametrize("style", ["lj", "eam"])
def test_step_kernel_timing_leaves_the_state_unchanged(lj_state, eam_state, style):
    """mdx_time_step_kernel launches the step kernel into spare buffers: a
    run after it equals, bit for bit, the same run without it."""
    from lammps_amd import mdx
    r = lj_state if style == "lj" else eam_state
    lo, hi = box_of(r, 20, A_LJ if style == "lj" else 3.615)
    ftm2v = 1.0 if style == "lj" else 1.0 / 1.0364269e-4
    out = []
    for probe in (False, True):
        ctx = device_from_oracle(r, lo, hi, style, mdx.MIXED, 0.3 if style == "lj" else 1.0, 5, 0, True)
        ctx.setup()
        ctx.run(7, 0.005, ftm2v, 0)
        if probe:
            ms = ctx.time_step_kernel(5, 0.005, ftm2v)
            assert 0.0 < ms < 100.0
        ctx.run(9, 0.005, ftm2v, 0)
        out.append(ctx.get_atoms())
        ctx.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_lj_bench_run_matches_reference_log():
    """bench/in.lj on the device (FP64): thermo rows at steps 0 and 100 equal
    the log's printed digits, 1202833 neighbours, 5 builds."""
    from lammps_amd import mdx
    g = GOLD["lj_bench_32k"]
    r = M.Ref(M.LJ)
    r.lattice_box(M.FCC, 0.8442, (0, 0, 0), (20, 20, 20))
    r.mass(1, 1.0)
    r.velocity_create(1.44, 87287, dist="uniform", loop="geom")
    lo, hi = box_of(r, 20, A_LJ)
    r.pair_lj(2.5)
    r.pair_coeff(1, 1, 1.0, 1.0, 2.5)
    ctx = device_from_oracle(r, lo, hi, "lj", mdx.FP64, 0.3, 20, 0, False)
    rows = ctx.setup() + ctx.run(100, 0.005, 1.0, 0)
    prd = np.array(hi) - np.array(lo)
    got = [[s] + mdx.thermo_columns(raw, 32000, prd, "lj") for s, raw in rows]
    assert [fmt(x) for x in got] == g["thermo_text"]
    st = ctx.stats()
    assert st["total_neighbors"] == 2 * g["total_neighbors"]
    assert st["ncalls"] == g["builds"]


def test_eam_bench_run_matches_reference_log():
    """bench/in.eam on the device (FP64): rows at 0, 50, 100, builds and neighbours."""
    from lammps_amd import mdx
    g = GOLD["eam_bench_32k"]
    r = M.Ref(M.METAL)
    r.lattice_box(M.FCC, 3.615, (0, 0, 0), (20, 20, 20))
    r.pair_eam(os.path.join(GOLDEN, "Cu_u3.eam"))
    r.velocity_create(1600.0, 376847, dist="uniform", loop="geom")
    r.neighbor(1.0, every=1, delay=5, check=True)
    r.compute(1, 0)  # initialises the spline tables
    lo, hi = box_of(r, 20, 3.615)
    ctx = device_from_oracle(r, lo, hi, "eam", mdx.FP64, 1.0, 1, 5, True)
    rows = ctx.setup() + ctx.run(100, 0.005, 1.0 / 1.0364269e-4, 50)
    prd = np.array(hi) - np.array(lo)
    got = [[s] + mdx.thermo_columns(raw, 32000, prd, "metal") for s, raw in rows]
    assert [fmt(x) for x in got] == g["thermo_text"]
    st = ctx.stats()
    assert st["total_neighbors"] == 2 * g["total_neighbors"]
    assert st["ncalls"] == g["builds"]
    assert st["ndanger"] == g["dangerous"]


@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed_fp64pos"])
def test_lj_nonperiodic_box_forces(lj_state, prec):
    """boundary f f f: the periodic crystal's atoms in a larger closed box (no
    images); the fixed-point mirror then maps the box onto 2^30."""
    from lammps_amd import mdx
    r0 = lj_state
    x, _, _, t, _ = r0.state()
    L = 20 * A_LJ
    lo, hi = (-1.0, -1.5, -2.0), (L + 1.0, L + 1.5, L + 2.0)
    x = np.mod(x, L)
    r = M.Ref(M.LJ)
    r.box(lo, hi, periodic=(0, 0, 0))
    r.add_atoms(x, t)
    r.mass(1, 1.0)
    r.pair_lj(2.5)
    r.pair_coeff(1, 1, 1.0, 1.0, 2.5)
    r.neighbor(0.3)
    r.setup()
    _, _, fref, _, _ = r.state()
    ctx = mdx.Context(0)
    ctx.set_box(lo, hi, (0, 0, 0))
    ctx.set_atoms(x, t, np.array([0.0, 1.0]), tags=np.arange(1, len(t) + 1))
    ctx.pair_lj_from_coeffs(1, {(1, 1): (1.0, 1.0, 2.5)})
    ctx.precision(precision_of(prec))
    precision_of(prec, ctx)
    ctx.neigh_modify(0.3)
    out = ctx.compute(1, 1)
    mx, rms = force_errors(out["f"], fref)
    print("LJ non-periodic %s force max_rel=%.3e rms_rel=%.3e" % (prec, mx, rms))
    assert mx <= (1e-11 if prec == "fp64" else 1e-5)
    eref, _ = r.energy_virial()
    assert abs(out["eng"] - eref) <= (1e-11 if prec == "fp64" else 1e-6) * abs(eref)


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 1), (0, 1, 1)])
def test_build_with_walls_matches_the_oracle(lj_state, periodic):
    """Stencil rows cut by non-periodic walls and wrapped segments on the
    periodic faces: the device list holds the oracle's half_bin_newton pair
    set, and its FP64 forces match the oracle's."""
    from lammps_amd import mdx
    x, _, _, t, _ = lj_state.state()
    L = 20 * A_LJ
    x = np.mod(x, L)
    lo = tuple(0.0 if p else -1.0 for p in periodic)
    hi = tuple(L if p else L + 1.0 for p in periodic)
    ref = M.Ref(M.LJ)
    ref.box(lo, hi, periodic)
    ref.add_atoms(x, t)
    ref.mass(1, 1.0)
    ref.pair_lj(2.5)
    ref.pair_coeff(1, 1, 1.0, 1.0, 2.5)
    ref.neighbor(0.3)
    ref.setup()
    want = {(min(a, b), max(a, b)) for a, b in ref.list_pairs()}
    fref = ref.state()[2]
    ctx = mdx.Context(0)
    ctx.set_box(lo, hi, periodic)
    ctx.set_atoms(x, t, np.array([0.0, 1.0]), tags=np.arange(1, len(t) + 1))
    ctx.pair_lj_from_coeffs(1, {(1, 1): (1.0, 1.0, 2.5)})
    ctx.precision(mdx.FP64)
    ctx.neigh_modify(0.3)
    ctx.neigh_build()
    got = {(int(a), int(b)) for a, b in ctx.neigh_pairs()}
    st = ctx.stats()
    f = ctx.compute(1, 1)["f"]
    ctx.close()
    assert len(want) > 0
    assert got == want
    assert st["total_neighbors"] == 2 * len(want)
    assert force_errors(f, fref)[0] <= 1e-11


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 0, 0), (1, 0, 1)])
def test_fixed_point_build_gives_the_fp64_pair_set(lj_state, periodic):
    """With the Q32 position format (the default, in either precision) the
    list build tests its candidates on the 32-bit fixed-point position
    mirror (FP32 test, FP64 re-test within 2e-6 of the cutoff): pair set,
    total and largest neighbour count equal those of the build on FP64
    positions (MDX_POS_FP64)."""
    from lammps_amd import mdx
    x, _, _, t, _ = lj_state.state()
    L = 20 * A_LJ
    x = np.mod(x, L)
    lo = tuple(0.0 if p else -1.0 for p in periodic)
    hi = tuple(L if p else L + 1.0 for p in periodic)
    out = []
    for fmt in (mdx.POS_Q32, mdx.POS_FP64):
        ctx = mdx.Context(0)
        ctx.set_box(lo, hi, periodic)
        ctx.set_atoms(x, t, np.array([0.0, 1.0]), tags=np.arange(1, len(t) + 1))
        ctx.pair_lj_from_coeffs(1, {(1, 1): (1.0, 1.0, 2.5)})
        ctx.precision(mdx.MIXED)
        ctx.position_format(fmt)
        ctx.neigh_modify(0.3)
        ctx.neigh_build()
        st = ctx.stats()
        out.append(({(int(a), int(b)) for a, b in ctx.neigh_pairs()}, st["total_neighbors"], st["max_neighbors"]))
        ctx.close()
    assert len(out[0][0]) > 0
    assert out[0] == out[1]


def test_fixed_point_build_near_the_cutoff_matches_the_oracle():
    """Pairs placed at cutneigh*(1+d) for d from -3e-6 to 3e-6 (inside and
    outside the FP64 re-test band), some through periodic images: the
    mixed-precision build keeps exactly the oracle's FP64 pair set."""
    from lammps_amd import mdx
    L, cut = 21.0, 2.8
    deltas = [-3e-6, -1e-6, -1e-7, -1e-9, -1e-12, 0.0, 1e-12, 1e-9, 1e-7, 1e-6, 3e-6]
    rng = np.random.default_rng(77)
    g = np.arange(3) * 7.0 + 0.5
    centers = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = []
    for k, c in enumerate(centers):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        if k % 2:
            u = -np.abs(u)  # towards the low faces: partners of the 0.5 row cross the boundary
        pts += [c, np.mod(c + cut * (1.0 + deltas[k % len(deltas)]) * u, L)]
    x = np.array(pts)
    n = len(x)
    ref = M.Ref(M.LJ)
    ref.box((0, 0, 0), (L, L, L))
    ref.add_atoms(x, np.ones(n, dtype=np.int32))
    ref.mass(1, 1.0)
    ref.pair_lj(2.5)
    ref.pair_coeff(1, 1, 1.0, 1.0, 2.5)
    ref.neighbor(0.3)
    ref.setup()
    want = {(min(a, b), max(a, b)) for a, b in ref.list_pairs()}
    ctx = mdx.Context(0)
    ctx.set_box((0, 0, 0), (L, L, L), (1, 1, 1))
    ctx.set_atoms(x, np.ones(n, dtype=np.int32), np.array([0.0, 1.0]), tags=np.arange(1, n + 1))
    ctx.pair_lj_from_coeffs(1, {(1, 1): (1.0, 1.0, 2.5)})
    ctx.precision(mdx.MIXED)
    ctx.neigh_modify(0.3)
    ctx.neigh_build()
    got = {(int(a), int(b)) for a, b in ctx.neigh_pairs()}
    ctx.close()
    assert 0 < len(want) < len(centers)  # about half of the placed pairs are kept
    assert got == want


@pytest.mark.parametrize("style", ["lj", "eam"])
def test_compute_adds_into_incoming_forces(lj_state, eam_state, style):
    """mdx_compute adds the pair forces into the caller's f, as Pair::compute
    adds into atom->f (src/pair.h:33-35, src/pair_lj_cut.cpp:115-117), and
    overwrites eatom/vatom, which ev_setup zeroes at every compute
    (src/pair.cpp:706-725): f_in + f_pair, called through the C ABI."""
    import ctypes
    from lammps_amd import mdx
    r = lj_state if style == "lj" else eam_state
    if style == "lj":
        lo, hi = box_of(r, 20, A_LJ)
        ctx = device_from_oracle(r, lo, hi, "lj", mdx.FP64, 0.3, 20, 0, False)
    else:
        lo, hi = box_of(r, 20, 3.615)
        ctx = device_from_oracle(r, lo, hi, "eam", mdx.FP64, 1.0, 1, 5, True)
    n = ctx.n
    rng = np.random.default_rng(11)
    f_in = rng.normal(size=(n, 3)) * 3.0
    f = f_in.copy()
    ea = np.full(n, 7.0)
    va = np.full((n, 6), -5.0)
    e = ctypes.c_double(0.0)
    w = np.zeros(6)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    assert ctx._L.mdx_compute(ctx.h, 3, 5, dp(f), ctypes.byref(e), dp(w), dp(ea), dp(va)) == 0
    fresh = ctx.compute(3, 5, peratom=True)  # into zeros
    ctx.close()
    fref = r.state()[2]
    assert force_errors(fresh["f"], fref)[0] <= 1e-11
    np.testing.assert_array_equal(f, f_in + fresh["f"])
    np.testing.assert_array_equal(ea, fresh["eatom"])
    np.testing.assert_array_equal(va, fresh["vatom"])
