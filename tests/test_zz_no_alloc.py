"""mdx_compute_noalloc (the no-allocation variant of mdx_compute, whose work
arrays live in a file-scope device array) against mdx_compute on the same
positions, through the C ABI.

Bars (stated here): the same kernels run in the same order on the same
sorted atoms, so FP64 forces, energy and virial equal mdx_compute's to
1e-12 relative (max_rel over atoms for forces) and mixed ones to 1e-10;
two calls in a row, the second on fewer atoms, each against a fresh
mdx_compute; forces are added into the caller's array (f += pair forces).
"""
import numpy as np
import pytest

from gpu_helpers import CU_U3, force_errors

pytestmark = pytest.mark.gpu


def _lj_system(ncell, seed):
    from lammps_amd import mdx
    lo, hi, x = mdx.lattice("fcc", 0.8442, "lj", (0, 0, 0), (ncell,) * 3)
    rng = np.random.default_rng(seed)
    x = x + rng.uniform(-0.08, 0.08, x.shape)
    return np.asarray(lo), np.asarray(hi), x


def _ctx(kind, lo, hi, prec, x=None):
    from lammps_amd import mdx
    ctx = mdx.Context(0)
    ctx.set_box(lo, hi, (1, 1, 1))
    mass = np.array([0.0, 1.0 if kind == "lj" else 63.55])
    if x is not None:
        ctx.set_atoms(x, np.ones(len(x), dtype=np.int32), mass)
    if kind == "lj":
        ctx.pair_lj_from_coeffs(1, {(1, 1): (1.0, 1.0, 2.5)})
        ctx.neigh_modify(0.3)
    else:
        mdx.eam_context_tables(ctx, CU_U3)
        ctx.neigh_modify(1.0)
    ctx.precision(prec)
    return ctx


def _noalloc(ctx, x, f0):
    import torch
    dev = torch.device("cuda", 0)
    xd = torch.tensor(x, dtype=torch.float64, device=dev).contiguous()
    td = torch.ones(len(x), dtype=torch.int32, device=dev)
    fd = torch.tensor(f0, dtype=torch.float64, device=dev).contiguous()
    torch.cuda.synchronize()
    e, w = ctx.compute_noalloc(len(x), xd.data_ptr(), td.data_ptr(), fd.data_ptr(), 1, 1)
    return fd.cpu().numpy(), e, w


@pytest.mark.parametrize("kind,prec", [("lj", "fp64"), ("lj", "mixed"), ("eam", "fp64"), ("eam", "mixed")])
def test_noalloc_matches_compute_on_two_calls(kind, prec):
    from lammps_amd import mdx
    p = mdx.FP64 if prec == "fp64" else mdx.MIXED
    tol = 1e-12 if prec == "fp64" else 1e-10
    if kind == "lj":
        lo, hi, x = _lj_system(14, 1)  # 10,976 atoms
    else:
        lo, hi, x = mdx.lattice("fcc", 3.615, "metal", (0, 0, 0), (12, 12, 12))  # 6,912 atoms
        lo, hi = np.asarray(lo), np.asarray(hi)
        x = x + np.random.default_rng(2).uniform(-0.1, 0.1, x.shape)
    na = _ctx(kind, lo, hi, p)  # parameters only: no atoms are set on it
    rng = np.random.default_rng(7)
    # call 1: every atom, f starts non-zero (the call adds into it)
    f0 = rng.normal(size=x.shape)
    ref = _ctx(kind, lo, hi, p, x)
    want = ref.compute(1, 1)
    ref.close()
    got, e, w = _noalloc(na, x, f0)
    mx, _ = force_errors(got - f0, want["f"])
    print("noalloc %s %s call 1 (%d atoms): max_rel %.3e" % (kind, prec, len(x), mx))
    assert mx <= tol
    assert abs(e - want["eng"]) <= tol * abs(want["eng"])
    assert np.max(np.abs(w - want["virial"])) <= tol * np.max(np.abs(want["virial"]))
    # call 2: fewer atoms (a random subset, shuffled), f starts at zero
    sub = rng.permutation(len(x))[: len(x) // 3]
    xs = x[sub]
    ref = _ctx(kind, lo, hi, p, xs)
    want = ref.compute(1, 1)
    ref.close()
    got, e, w = _noalloc(na, xs, np.zeros_like(xs))
    mx, _ = force_errors(got, want["f"])
    print("noalloc %s %s call 2 (%d atoms): max_rel %.3e" % (kind, prec, len(xs), mx))
    assert mx <= tol
    assert abs(e - want["eng"]) <= tol * abs(want["eng"])
    assert np.max(np.abs(w - want["virial"])) <= tol * np.max(np.abs(want["virial"]))
    na.close()


def test_noalloc_limits_are_reported_without_computing():
    from lammps_amd import mdx
    lo, hi, x = _lj_system(4, 3)
    na = _ctx("lj", lo, hi, mdx.FP64)
    with pytest.raises(mdx.MdxError) as e:
        na.compute_noalloc(mdx.NOALLOC_MAX_ATOMS + 1, 0, 0, 0)
    assert e.value.code == mdx.ERR_ARG
    with pytest.raises(mdx.MdxError) as e:  # device arrays are required
        na.compute_noalloc(10, 0, 0, 0)
    assert e.value.code == mdx.ERR_ARG
    e, w = na.compute_noalloc(0, 0, 0, 0)
    assert e == 0.0 and not np.any(w)
    na.close()
