"""The product's host-side set-up commands (include/mdx_host.h) against the
oracle: bit-identical lattices, velocities and EAM spline tables (CPU only)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import mdref as M

CU_U3 = os.path.join(GOLDEN, "Cu_u3.eam")


def test_fcc_lattice_lj_matches_oracle_bitwise():
    from lammps_amd import mdx
    blo, bhi, x = mdx.lattice("fcc", 0.8442, "lj", (0, 0, 0), (7, 6, 5))
    r = M.Ref(M.LJ)
    r.lattice_box(M.FCC, 0.8442, (0, 0, 0), (7, 6, 5))
    xr, _, _, _, _ = r.state()
    assert x.shape == (4 * 7 * 6 * 5, 3)
    assert np.array_equal(x, xr)
    assert bhi[0] == pytest.approx(7 * (4 / 0.8442) ** (1 / 3), rel=1e-15)


@pytest.mark.parametrize("style,ostyle", [("sc", M.SC), ("bcc", M.BCC), ("fcc", M.FCC)])
def test_metal_lattices_match_oracle(style, ostyle):
    from lammps_amd import mdx
    _, _, x = mdx.lattice(style, 3.615, "metal", (0, 0, 0), (4, 4, 4))
    r = M.Ref(M.METAL)
    r.lattice_box(ostyle, 3.615, (0, 0, 0), (4, 4, 4))
    xr, _, _, _, _ = r.state()
    assert np.array_equal(x, xr)


@pytest.mark.parametrize("dist", ["uniform", "gaussian"])
@pytest.mark.parametrize("loop", ["all", "geom"])
def test_velocity_create_matches_oracle_bitwise(dist, loop):
    from lammps_amd import mdx
    _, _, x = mdx.lattice("fcc", 0.8442, "lj", (0, 0, 0), (5, 5, 5))
    t = np.ones(len(x), dtype=np.int32)
    v = mdx.velocity_create(x, t, np.array([0.0, 1.0]), 1.44, 87287, dist=dist, loop=loop)
    r = M.Ref(M.LJ)
    r.lattice_box(M.FCC, 0.8442, (0, 0, 0), (5, 5, 5))
    r.mass(1, 1.0)
    r.velocity_create(1.44, 87287, dist=dist, loop=loop)
    _, vr, _, _, _ = r.state()
    assert np.array_equal(v, vr)


def test_velocity_create_metal_units_and_mass():
    from lammps_amd import mdx
    _, _, x = mdx.lattice("fcc", 3.615, "metal", (0, 0, 0), (4, 4, 4))
    t = np.ones(len(x), dtype=np.int32)
    v = mdx.velocity_create(x, t, np.array([0.0, 63.55]), 1600.0, 376847, loop="geom", units="metal")
    r = M.Ref(M.METAL)
    r.lattice_box(M.FCC, 3.615, (0, 0, 0), (4, 4, 4))
    r.mass(1, 63.55)
    r.velocity_create(1600.0, 376847, loop="geom")
    _, vr, _, _, _ = r.state()
    assert np.array_equal(v, vr)


def oracle_tables(path):
    r = M.Ref(M.METAL)
    r.box((0, 0, 0), (20, 20, 20))
    r.add_atoms(np.array([[1.0, 1.0, 1.0]]), np.array([1], dtype=np.int32))
    r.pair_eam(path)
    r.neighbor(1.0)
    r.compute(1, 0)
    return r.eam_tables()


def test_funcfl_tables_match_oracle_bitwise():
    from lammps_amd import mdx
    t = mdx.eam_funcfl(CU_U3)
    o = oracle_tables(CU_U3)
    for k in ("nrho", "nr", "drho", "dr", "rhomax", "cutmax"):
        assert t[k] == o[k], k
    for k in ("frho_spline", "rhor_spline", "z2r_spline"):
        assert np.array_equal(t[k], o[k]), k
    assert t["mass"] == 63.55


def test_synthetic_funcfl_is_the_analytic_potential(tmp_path):
    """The generated table is the closed forms of the C2 workload: the spline
    reproduces F, rho and r*phi at off-grid points to interpolation accuracy,
    and the perfect lattice sits at the reference density rho_e."""
    from lammps_amd import mdx
    path = str(tmp_path / "synth.eam")
    par = mdx.write_synthetic_funcfl(path, seed=4951)
    assert 10 <= par["Z0"] <= 12 and 1.4 <= par["beta"] <= 1.6 and 5.5 <= par["chi"] <= 6.5
    assert 2.5 <= par["E0"] <= 3.5
    assert par["re"] == pytest.approx(3.615 / np.sqrt(2), rel=1e-15)
    t = mdx.eam_funcfl(path)
    o = oracle_tables(path)
    for k in ("frho_spline", "rhor_spline", "z2r_spline"):
        assert np.array_equal(t[k], o[k]), k
    rr = t["rhor_spline"][0]
    for r in (2.3, 2.5562, 3.1, 4.4):
        m = int(r / t["dr"] + 1.0)
        p = r / t["dr"] + 1.0 - m
        c = rr[m]
        val = ((c[3] * p + c[4]) * p + c[5]) * p + c[6]
        exact = np.exp(-par["chi"] * (r / par["re"] - 1.0))
        assert val == pytest.approx(exact, rel=1e-6)
    a = 3.615
    shells = [(12, a / np.sqrt(2)), (6, a), (24, a * np.sqrt(1.5))]
    rho_e = sum(k * np.exp(-par["chi"] * (r / par["re"] - 1.0)) for k, r in shells)
    assert par["rho_e"] == pytest.approx(rho_e, rel=1e-14)
