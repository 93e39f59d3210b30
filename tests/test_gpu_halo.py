"""Ghost-shell exchange of a spatially decomposed run on one GPU's share:
send lists, pack with the image shift and unpack (mdx_halo_*; the
reference's CommBrick::borders / forward_comm, src/comm_brick.cpp:398-460,
696-864, and AtomVecAtomic::pack_comm with pbc, src/atom_vec_atomic.cpp).

A periodic system G plays every neighbouring rank.  One octant of its box
becomes a sub-domain D: the octant's atoms are D's owned atoms, and the atoms
of G within rc+skin outside the octant, packed from G with their image shift
and unpacked into D, are D's ghosts.  D's box is closed (no periodic images),
so every neighbour of an owned atom that lies across a face reaches it only
as a ghost.

Bars (stated here, in the test):
  * the packed ghost positions equal x + shift bit for bit (the reference's
    ghost coordinate);
  * forces and per-atom energies of D's owned atoms equal G's for the same
    atoms: FP64 max_rel <= 1e-12 (summation order only), mixed <= 1e-5
    against G in FP64;
  * bad arguments (host buffer, out-of-range atoms or swap) are errors that
    launch nothing.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A_LJ = (4.0 / 0.8442) ** (1.0 / 3.0)
CUT = 2.5 + 0.3  # neighbour cutoff: the ghost shell width


def perturbed_lattice(ncell, seed):
    from lammps_amd import mdx
    _, hi, x = mdx.lattice("fcc", 0.8442, "lj", (0, 0, 0), (ncell,) * 3)
    L = hi[0]
    rng = np.random.default_rng(seed)
    x = np.mod(x + rng.uniform(-0.12, 0.12, x.shape), L)
    return x, L


def lj_ctx(lo, hi, periodic, x, prec):
    from lammps_amd import mdx
    ctx = mdx.Context(0)
    ctx.set_box(lo, hi, periodic)
    ctx.set_atoms(x, np.ones(len(x), dtype=np.int32), np.array([0.0, 1.0]), tags=np.arange(1, len(x) + 1))
    ctx.pair_lj_from_coeffs(1, {(1, 1): (1.0, 1.0, 2.5)})
    ctx.precision(mdx.FP64 if prec == "fp64" else mdx.MIXED)
    ctx.neigh_modify(0.3)
    return ctx


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_octant_with_packed_ghosts_reproduces_the_periodic_forces(prec):
    import torch
    x, L = perturbed_lattice(12, seed=5)
    n = len(x)
    G = lj_ctx((0, 0, 0), (L, L, L), (1, 1, 1), x, "fp64")
    ref = G.compute(1, 1, peratom=True)
    xg = G.get_atoms()[0]  # positions as G holds them after its build (wrapped into the box)
    h = L / 2.0
    owned = np.where(np.all((xg >= 0.0) & (xg < h), axis=1))[0]
    from lammps_amd import mdx
    regions = mdx.octant_shell_regions(L, CUT)
    counts = [G.halo_sendlist(k, lo, hi, sh) for k, (lo, hi, sh) in enumerate(regions)]
    ng = sum(counts)
    expect = []
    for (lo, hi, sh) in regions:
        sel = np.all((xg >= lo) & (xg < hi), axis=1)
        expect.append(xg[sel] + np.array(sh))
    expect = np.concatenate(expect)
    assert len(expect) == ng and ng > 0
    buf = torch.zeros(3 * ng, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()  # torch fills on its own stream; the library runs on another
    off = 0
    for k, c in enumerate(counts):
        G.halo_pack(k, buf.data_ptr() + 8 * 3 * off)
        off += c
    torch.cuda.synchronize()
    got = buf.cpu().numpy().reshape(-1, 3)
    # all swaps in one launch: the same rows in the same order
    buf_all = torch.zeros(3 * ng, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    G.halo_pack(-1, buf_all.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(buf_all.cpu().numpy().reshape(-1, 3), got)
    # all send lists in one pass: the same lists (the packed rows match in order)
    assert G.halo_sendlists(regions) == counts
    buf_all.zero_()
    torch.cuda.synchronize()
    G.halo_pack(-1, buf_all.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(buf_all.cpu().numpy().reshape(-1, 3), got)
    # the same rows bit for bit (pack order is G's device order)
    key = lambda a: a[np.lexsort(a.T[::-1])]
    assert np.array_equal(key(got), key(expect))
    # D: owned atoms first, then placeholders for the ghosts, filled by unpack
    lo_d, hi_d = (-CUT - 0.5,) * 3, (h + CUT + 0.5,) * 3
    place = np.full((ng, 3), h / 2.0) + np.linspace(0.0, 1.0, ng)[:, None] * 0.01
    D = lj_ctx(lo_d, hi_d, (0, 0, 0), np.vstack([xg[owned], place]), prec)
    D.halo_unpack(len(owned), ng, buf.data_ptr())
    out = D.compute(1, 1, peratom=True)
    xd = D.get_atoms()[0]
    assert np.array_equal(xd[len(owned):], got)
    fo, fr = out["f"][: len(owned)], ref["f"][owned]
    e = np.linalg.norm(fo - fr, axis=1)
    scale = np.sqrt(np.mean(np.sum(fr * fr, axis=1)))
    tol = 1e-12 if prec == "fp64" else 1e-5
    assert e.max() / scale <= tol
    eo, er = out["eatom"][: len(owned)], ref["eatom"][owned]
    assert np.max(np.abs(eo - er)) <= tol * np.max(np.abs(er))
    D.close()
    G.close()


def test_overlapping_send_regions_in_one_pass():
    """regions that overlap (an atom sent to several neighbours): the one-pass
    lists equal the per-region ones"""
    import torch
    from lammps_amd import mdx
    x, L = perturbed_lattice(8, seed=9)
    G = lj_ctx((0, 0, 0), (L, L, L), (1, 1, 1), x, "fp64")
    G.neigh_build()
    rng = np.random.default_rng(4)
    regions = []
    for _ in range(27):
        a = rng.uniform(0, L, 3); b = rng.uniform(0, L, 3)
        regions.append((np.minimum(a, b), np.maximum(a, b), rng.choice([-L, 0.0, L], 3)))
    one = [G.halo_sendlist(k, lo, hi, sh) for k, (lo, hi, sh) in enumerate(regions)]
    tot = sum(one)
    b1 = torch.zeros(3 * max(tot, 1), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    G.halo_pack(-1, b1.data_ptr())
    assert G.halo_sendlists(regions) == one
    b2 = torch.zeros(3 * max(tot, 1), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    G.halo_pack(-1, b2.data_ptr())
    torch.cuda.synchronize()
    assert tot > len(x) // 4 and np.array_equal(b1.cpu().numpy(), b2.cpu().numpy())
    G.close()


def test_halo_argument_errors_launch_nothing():
    from lammps_amd import mdx
    x, L = perturbed_lattice(6, seed=2)
    G = lj_ctx((0, 0, 0), (L, L, L), (1, 1, 1), x, "fp64")
    G.neigh_build()
    c = G.halo_sendlist(0, (0, 0, 0), (L / 2, L / 2, L / 2), (0, 0, 0))
    assert c > 0
    host = np.zeros(3 * c)
    with pytest.raises(mdx.MdxError) as e:
        G.halo_pack(0, host.ctypes.data)
    assert e.value.code == mdx.ERR_ARG and "device memory" in str(e.value)
    with pytest.raises(mdx.MdxError) as e:
        G.halo_pack(-1, host.ctypes.data)
    assert e.value.code == mdx.ERR_ARG and "device memory" in str(e.value)
    with pytest.raises(mdx.MdxError) as e:
        G.halo_sendlist(27, (0, 0, 0), (1, 1, 1), (0, 0, 0))
    assert e.value.code == mdx.ERR_ARG
    with pytest.raises(mdx.MdxError) as e:
        G.halo_unpack(len(x) - 1, 2, host.ctypes.data)
    assert e.value.code == mdx.ERR_ARG and "must exist" in str(e.value)
    G.close()


def test_stale_and_replaced_send_lists():
    """A list build permutes the atoms, so a send list made before it names
    the wrong atoms: packing it is an error until the list is rebuilt (as
    borders rebuilds the send lists at every reneighbouring).  One-pass
    lists replace every earlier list, not only the first nswap."""
    import torch
    from lammps_amd import mdx
    x, L = perturbed_lattice(6, seed=4)
    G = lj_ctx((0, 0, 0), (L, L, L), (1, 1, 1), x, "fp64")
    G.neigh_build()
    h = L / 2
    c3 = G.halo_sendlist(3, (0, 0, 0), (h, h, h), (0, 0, 0))
    regions = [((0, 0, 0), (h, L, L), (0, 0, 0)), ((h, 0, 0), (L, L, L), (-L, 0, 0))]
    counts = G.halo_sendlists(regions)
    assert c3 > 0 and sum(counts) == len(x)
    buf = torch.zeros(3 * len(x) + 3 * c3, dtype=torch.float64, device="cuda")
    G.halo_pack(-1, buf.data_ptr())  # swap 3's old list is gone: the two new lists only
    torch.cuda.synchronize()
    packed = buf.cpu().numpy().reshape(-1, 3)
    assert np.all(packed[len(x):] == 0.0)
    G.neigh_build()
    for swap in (0, -1):
        with pytest.raises(mdx.MdxError) as e:
            G.halo_pack(swap, buf.data_ptr())
        assert e.value.code == mdx.ERR_ARG and "last list build" in str(e.value)
    G.halo_sendlists(regions)
    G.halo_pack(-1, buf.data_ptr())
    G.close()
