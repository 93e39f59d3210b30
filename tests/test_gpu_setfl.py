"""eam/alloy and eam/fs on the device against the oracle (setfl readers and
type maps pinned in tests/test_eam_setfl.py): a displaced binary fcc
crystal with random types; FP64 forces max_rel <= 1e-11, mixed <= 1e-5;
an FP64 NVE run reproduces the oracle's thermo to the printed digits."""
import numpy as np
import pytest

from gpu_helpers import force_errors
from oracle import mdref as M
from setfl_util import binary_cu_like

pytestmark = pytest.mark.gpu


def fcc(ncell, a, seed, disp):
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*(np.arange(ncell),) * 3, indexing="ij"), -1).reshape(-1, 3)
    x = ((g[:, None, ::-1] + basis[None]) * a).reshape(-1, 3)
    return x + np.random.default_rng(seed).normal(scale=disp, size=x.shape), ncell * a


@pytest.fixture(scope="module", params=[False, True], ids=["alloy", "fs"])
def system(request, tmp_path_factory):
    fs = request.param
    path = str(tmp_path_factory.mktemp("setfl") / "AB.setfl")
    binary_cu_like(path, fs=fs)
    x, L = fcc(8, 3.615, 7, 0.1)
    types = np.random.default_rng(8).integers(1, 3, size=len(x)).astype(np.int32)
    r = M.Ref(M.METAL)
    r.box((0, 0, 0), (L, L, L), ntypes=2)
    r.add_atoms(x, types)
    r.pair_eam_setfl(path, ["A", "B"], fs=fs)
    r.neighbor(1.0, every=1, delay=0, check=True)
    r.timestep(0.002)
    r.velocity_create(800.0, 4711, dist="gaussian", loop="geom")
    r.setup()
    return dict(fs=fs, path=path, x=x, L=L, types=types, ref=r)


def device(sysd, precision):
    from lammps_amd import mdx
    r = sysd["ref"]
    x, v, _, t, _ = r.state()
    L = sysd["L"]
    ctx = mdx.Context(0)
    ctx.set_box((0, 0, 0), (L, L, L))
    tab = mdx.eam_setfl(sysd["path"], ["A", "B"], fs=sysd["fs"])
    ctx.set_atoms(x, t, tab["mass"], v=v, tags=np.arange(1, len(t) + 1))
    mdx.eam_context_setfl(ctx, sysd["path"], ["A", "B"], fs=sysd["fs"])
    ctx.precision(precision)
    ctx.neigh_modify(1.0, every=1, delay=0, check=True)
    return ctx


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_setfl_forces_energy(system, prec):
    from lammps_amd import mdx
    ctx = device(system, mdx.FP64 if prec == "fp64" else mdx.MIXED)
    out = ctx.compute(1, 1)
    r = system["ref"]
    fref = r.state()[2]
    eref, wref = r.energy_virial()
    mx, rms = force_errors(out["f"], fref)
    print("setfl %s %s max_rel=%.3e rms_rel=%.3e" % ("fs" if system["fs"] else "alloy", prec, mx, rms))
    assert mx <= (1e-11 if prec == "fp64" else 1e-5)
    assert abs(out["eng"] - eref) <= (1e-11 if prec == "fp64" else 1e-6) * abs(eref)
    ctx.close()


def test_setfl_nve_run_matches_oracle_thermo(system, tmp_path):
    from lammps_amd import mdx
    ctx = device(system, mdx.FP64)
    ctx.setup()
    rows = ctx.run(40, 0.002, mdx.UNIT_CONSTANTS["metal"]["ftm2v"], 20)
    ctx.close()
    r = M.Ref(M.METAL)
    L = system["L"]
    x0, v0, _, t0, _ = system["ref"].state()
    r.box((0, 0, 0), (L, L, L), ntypes=2)
    r.add_atoms(x0, t0)
    r.set_velocities(v0)
    r.pair_eam_setfl(system["path"], ["A", "B"], fs=system["fs"])
    r.neighbor(1.0, every=1, delay=0, check=True)
    r.timestep(0.002)
    r.thermo(20)
    r.setup()
    ref_rows = r.run(40)
    n = len(t0)
    got = {int(step): mdx.thermo_columns(raw, n, (L, L, L), "metal") for step, raw in rows}
    ref = {int(q[0]): q[1:] for q in ref_rows}
    common = sorted(set(got) & set(ref))
    assert 20 in common and 40 in common
    for step in common:
        for a, b in zip(got[step], ref[step]):
            assert ("%12.8g" % a).strip() == ("%12.8g" % b).strip(), (step, a, b)
