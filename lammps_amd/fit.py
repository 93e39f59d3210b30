"""Differentiable forces for force matching (torch autograd over libmdx).

    fm = LJForceMatch(batch_geometry..., positions, cut_pairs)
    F = fm(eps_pairs, sigma_pairs)          # [N, 3] float64 device tensor
    loss = ((F - F_target) ** 2).sum() / (3 * N)
    loss.backward()                          # eps_pairs.grad, sigma_pairs.grad

The autograd Functions are C++ (torch::autograd::Function in
lammps_amd/csrc/torch_ext.cpp, built in-tree as _mdx_torch.so); their
backward runs the hand-written adjoint kernels of libmdx.  Pair parameters
are 1-D CPU float64 tensors over type pairs (a <= b) in row order.
"""
import importlib.util
import os

import numpy as np

from .batch import FP64, MIXED, Batch

HERE = os.path.dirname(os.path.abspath(__file__))
EXT_PATH = os.path.join(HERE, "_mdx_torch.so")
_ext = None


def ext():
    """The in-tree torch extension (fails loudly when it is not built)."""
    global _ext
    if _ext is None:
        import torch  # noqa: F401  (libtorch must be loaded first)

        from lammps_amd import mdx

        mdx.lib()
        if mdx.LOADED_FLAVOUR != "torch":
            raise RuntimeError("libmdx was loaded without torch in the process (MDX_HIP_RUNTIME=rocm), so it "
                               "runs on ROCm's HIP runtime, not torch's; the torch extension needs one runtime")
        if not os.path.exists(EXT_PATH):
            raise RuntimeError("_mdx_torch.so is not built (run __graft_entry__.build())")
        spec = importlib.util.spec_from_file_location("_mdx_torch", EXT_PATH)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _ext = mod
    return _ext


def pair_index(ntypes):
    """[(a, b)] for a <= b in the order the pair vectors use."""
    return [(a, b) for a in range(1, ntypes + 1) for b in range(a, ntypes + 1)]


def pair_table(values, ntypes):
    """Symmetric (ntypes+1)^2 table from a pair vector."""
    t = np.zeros((ntypes + 1, ntypes + 1))
    for v, (a, b) in zip(np.asarray(values, dtype=np.float64), pair_index(ntypes)):
        t[a, b] = t[b, a] = v
    return t


class LJForceMatch:
    """lj/cut forces of a batch of configurations as a differentiable
    function of the (eps, sigma) pair values; cutoffs fixed."""

    def __init__(self, natoms, lo, hi, periodic, ntypes, types, x, cut_pairs, precision=MIXED, device=0):
        import torch

        self.torch = torch
        self.batch = Batch(natoms, lo, hi, periodic, ntypes, types, device=device)
        self.batch.set_positions(x)
        self.batch.precision(precision)
        self.ntypes = ntypes
        self.device = device
        self.cut = torch.tensor(np.asarray(cut_pairs, dtype=np.float64))

    def __call__(self, eps_pairs, sigma_pairs):
        return ext().lj_forces(eps_pairs, sigma_pairs, self.cut, self.batch.h.value, self.ntypes, self.batch.n,
                               self.device, self.torch.cuda.current_stream(self.device).cuda_stream)


class EAMForceMatch:
    """eam forces of a batch of configurations as a differentiable function
    of the frho, rhor and z2r knot values."""

    def __init__(self, natoms, lo, hi, periodic, x, drho, dr, cut, precision=MIXED, device=0, ntypes=1,
                 types=None):
        import torch

        self.torch = torch
        n = int(np.sum(natoms))
        self.batch = Batch(natoms, lo, hi, periodic, ntypes, np.ones(n, np.int32) if types is None else types,
                           device=device)
        self.batch.set_positions(x)
        self.batch.precision(precision)
        self.device = device
        self.drho, self.dr, self.cut = float(drho), float(dr), float(cut)

    def __call__(self, frho, rhor, z2r):
        return ext().eam_forces(frho, rhor, z2r, self.batch.h.value, self.batch.n, self.device, self.drho, self.dr,
                                self.cut, self.torch.cuda.current_stream(self.device).cuda_stream)


__all__ = ["FP64", "MIXED", "LJForceMatch", "EAMForceMatch", "pair_index", "pair_table", "ext"]
