"""eam/alloy and eam/fs (DYNAMO setfl files): the oracle's reader and
file2array (oracle/mdref.c, following pair_eam_alloy.cpp / pair_eam_fs.cpp)
and the library's host reader (mdx_host_eam_setfl), pinned by:

- a one-element setfl holding a funcfl potential's own knots gives the
  funcfl path's splines and forces bit for bit;
- a two-element file whose elements are identical gives the one-element
  forces bit for bit, whatever the type assignment;
- the energy of an A-B dimer equals the EAM definition evaluated by hand
  from the splines (F_A(rho at A) + F_B(rho at B) + z2r_AB(r)/r), which
  pins the density orientation of the type maps (alloy: the neighbour's
  element function; fs: block of the neighbour's element, function of the
  receiving element);
- forces are minus the energy gradient (central differences) for binary
  alloy and fs crystals;
- the host reader's tables and type maps equal the oracle's bit for bit.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import mdref as M
from setfl_util import binary_cu_like, spline_eval, write_setfl

CU_U3 = os.path.join(GOLDEN, "Cu_u3.eam")


def fcc(ncell, a, seed, disp):
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*(np.arange(ncell),) * 3, indexing="ij"), -1).reshape(-1, 3)
    x = ((g[:, None, ::-1] + basis[None]) * a).reshape(-1, 3)
    return x + np.random.default_rng(seed).normal(scale=disp, size=x.shape), ncell * a


def ref_system(x, L, types, ntypes, setfl=None, elements=None, fs=False, funcfl=None, periodic=(1, 1, 1),
               lo=(0, 0, 0), hi=None):
    r = M.Ref(M.METAL)
    r.box(lo, hi if hi is not None else (L, L, L), periodic=periodic, ntypes=ntypes)
    r.add_atoms(x, types)
    if funcfl:
        r.pair_eam(funcfl)
    else:
        r.pair_eam_setfl(setfl, elements, fs=fs)
    r.neighbor(1.0)
    r.setup()
    return r


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("setfl")
    out = {}
    for fs in (False, True):
        p = str(d / ("AB.eam.%s" % ("fs" if fs else "alloy")))
        binary_cu_like(p, fs=fs)
        out[fs] = p
    # Cu_u3 re-expressed on its own knots
    r = M.Ref(M.METAL)
    r.box((0, 0, 0), (10, 10, 10))
    r.add_atoms(np.array([[1.0, 1.0, 1.0]]), np.array([1], dtype=np.int32))
    r.pair_eam(CU_U3)
    r.neighbor(1.0)
    r.setup()
    t = r.eam_tables()
    fk, rk, zk = (t[k][0][1:, 6] for k in ("frho_spline", "rhor_spline", "z2r_spline"))
    p1 = str(d / "Cu1.eam.alloy")
    write_setfl(p1, ["Cu"], [63.55], t["nrho"], t["drho"], t["nr"], t["dr"], t["cutmax"], [fk], [rk],
                {(0, 0): zk})
    out["cu1"] = p1
    p2 = str(d / "CuCu.eam.alloy")
    write_setfl(p2, ["Cu", "Cx"], [63.55, 63.55], t["nrho"], t["drho"], t["nr"], t["dr"], t["cutmax"], [fk, fk],
                [rk, rk], {(0, 0): zk, (1, 0): zk, (1, 1): zk})
    out["cucu"] = p2
    return out


def test_one_element_setfl_equals_funcfl_bitwise(files):
    x, L = fcc(4, 3.615, 1, 0.12)
    t = np.ones(len(x), dtype=np.int32)
    a = ref_system(x, L, t, 1, funcfl=CU_U3)
    b = ref_system(x, L, t, 1, setfl=files["cu1"], elements=["Cu"])
    ta, tb = a.eam_tables(), b.eam_tables()
    for k in ("frho_spline", "rhor_spline", "z2r_spline"):
        np.testing.assert_array_equal(ta[k][0], tb[k][0])
    # funcfl keeps rhomax = (file nrho - 1) drho, one step past its last knot
    # (pair_eam.cpp:497-502); setfl has rhomax = (nrho - 1) drho: they differ
    # only above the tables, which these configurations never reach
    assert ta["cutmax"] == tb["cutmax"] and tb["rhomax"] == (tb["nrho"] - 1) * tb["drho"]
    np.testing.assert_array_equal(a.state()[2], b.state()[2])
    assert a.energy_virial()[0] == b.energy_virial()[0]


def test_two_identical_elements_equal_one_element_bitwise(files):
    x, L = fcc(4, 3.615, 2, 0.12)
    one = ref_system(x, L, np.ones(len(x), dtype=np.int32), 1, setfl=files["cu1"], elements=["Cu"])
    types = np.random.default_rng(3).integers(1, 3, size=len(x)).astype(np.int32)
    two = ref_system(x, L, types, 2, setfl=files["cucu"], elements=["Cu", "Cx"])
    np.testing.assert_array_equal(one.state()[2], two.state()[2])
    assert one.energy_virial()[0] == two.energy_virial()[0]


def spline_at(sp, delta, n, v):
    p = v / delta + 1.0
    m = min(max(int(p), 1), n - 1)
    return spline_eval(sp, m, min(p - m, 1.0))


@pytest.mark.parametrize("fs", [False, True])
def test_dimer_energy_matches_the_eam_definition(files, fs):
    r_ab = 2.61
    x = np.array([[10.0, 10.0, 10.0], [10.0 + r_ab, 10.0, 10.0]])
    types = np.array([1, 2], dtype=np.int32)
    ref = ref_system(x, None, types, 2, setfl=files[fs], elements=["A", "B"], fs=fs, periodic=(0, 0, 0),
                     lo=(0, 0, 0), hi=(20, 20, 20))
    t = ref.eam_tables()
    n_r, dr, n_rho, drho = t["nr"], t["dr"], t["nrho"], t["drho"]
    rhor, frho, z2r = t["rhor_spline"], t["frho_spline"], t["z2r_spline"]
    # density received by A (element 0) from B (element 1) and vice versa
    if fs:
        rho_at_a = spline_at(rhor[1 * 2 + 0], dr, n_r, r_ab)[0]  # block of B, function for A
        rho_at_b = spline_at(rhor[0 * 2 + 1], dr, n_r, r_ab)[0]
    else:
        rho_at_a = spline_at(rhor[1], dr, n_r, r_ab)[0]  # B's density function
        rho_at_b = spline_at(rhor[0], dr, n_r, r_ab)[0]
    e = spline_at(frho[0], drho, n_rho, rho_at_a)[0] + spline_at(frho[1], drho, n_rho, rho_at_b)[0]
    e += spline_at(z2r[1], dr, n_r, r_ab)[0] / r_ab  # lower triangle (1,0) -> index 1
    rho, _ = ref.eam_rho_fp()
    assert rho[0] == pytest.approx(rho_at_a, rel=1e-14) and rho[1] == pytest.approx(rho_at_b, rel=1e-14)
    assert ref.energy_virial()[0] == pytest.approx(e, rel=1e-13)
    # the two densities differ, so a swapped orientation would fail the checks above
    assert abs(rho_at_a - rho_at_b) > 1e-3 * abs(rho_at_a)


@pytest.mark.parametrize("fs", [False, True])
def test_binary_forces_are_minus_energy_gradient(files, fs):
    x, L = fcc(3, 3.615, 4, 0.15)
    types = np.random.default_rng(5).integers(1, 3, size=len(x)).astype(np.int32)
    r = ref_system(x, L, types, 2, setfl=files[fs], elements=["A", "B"], fs=fs)
    f = r.state()[2]
    h = 1e-6
    for i in (0, 17, 50):
        for d in range(3):
            es = []
            for sgn in (1, -1):
                xx = x.copy()
                xx[i, d] += sgn * h
                es.append(ref_system(xx, L, types, 2, setfl=files[fs], elements=["A", "B"], fs=fs).energy_virial()[0])
            assert abs(-(es[0] - es[1]) / (2 * h) - f[i, d]) < 1e-5 * max(1.0, abs(f[i, d]))


@pytest.mark.parametrize("fs", [False, True])
def test_host_reader_tables_equal_the_oracle(files, fs):
    from lammps_amd import mdx
    x, L = fcc(2, 3.615, 6, 0.1)
    elements = ["B", "A", "B"]  # three types on two elements, reversed order
    types = np.array([1, 2, 3] * (len(x) // 3) + [1] * (len(x) % 3), dtype=np.int32)
    r = ref_system(x, L, types, 3, setfl=files[fs], elements=elements, fs=fs)
    t_ref = r.eam_tables()
    t = mdx.eam_setfl(files[fs], elements, fs=fs)
    for k in ("nrho", "nr", "drho", "dr", "rhomax", "cutmax"):
        assert t[k] == t_ref[k], k
    for k in ("frho_spline", "rhor_spline", "z2r_spline"):
        np.testing.assert_array_equal(t[k], t_ref[k])
    nt1 = 4
    t2f = np.zeros(9, dtype=np.int32)
    t2r = np.zeros(81, dtype=np.int32)
    t2z = np.zeros(81, dtype=np.int32)
    r._ck(r._L.mdref_eam_splines(r.h, M._ptr(np.zeros_like(t_ref["frho_spline"])),
                                 M._ptr(np.zeros_like(t_ref["rhor_spline"])),
                                 M._ptr(np.zeros_like(t_ref["z2r_spline"])), M._ptr(t2f, M._ip),
                                 M._ptr(t2r, M._ip), M._ptr(t2z, M._ip)))
    np.testing.assert_array_equal(t["type2frho"][1:], t2f[1:nt1])
    got_r = t["type2rhor"].reshape(nt1, nt1)[1:, 1:]
    got_z = t["type2z2r"].reshape(nt1, nt1)[1:, 1:]
    # the oracle packs (ntypes+1)^2 maps with stride ntypes+1
    np.testing.assert_array_equal(got_r, t2r[: nt1 * nt1].reshape(nt1, nt1)[1:, 1:])
    np.testing.assert_array_equal(got_z, t2z[: nt1 * nt1].reshape(nt1, nt1)[1:, 1:])
    masses = {"A": 63.55, "B": 58.69}
    np.testing.assert_array_equal(t["mass"][1:], [masses[e] for e in elements])
