"""ctypes binding of the batched force / force-matching entry points
(include/mdx_batch.h).

Arrays may be numpy arrays (host) or torch CUDA tensors on the batch's
device (float64, contiguous); device tensors are read and written in place
without a host round trip.  There is no CPU fallback: the in-tree
libmdx.so must be built and a HIP device present.
"""
import ctypes

import numpy as np

from . import mdx as _mdx

FP64, MIXED = _mdx.FP64, _mdx.MIXED
BK_BUILD, BK_LJ, BK_LJ_VJP, BK_EAM_RHO, BK_EAM_FORCE, BK_EAM_ADJ_A, BK_EAM_ADJ_B = range(7)
KIND_NAMES = ["build", "lj", "lj_vjp", "eam_rho", "eam_force", "eam_adj_a", "eam_adj_b"]

_p = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_configured = False


def lib():
    global _configured
    L = _mdx.lib()
    if not _configured:
        sig = {
            "mdx_batch_create": [_i, _i, _p, _p, _p, _p, _i, _p, ctypes.POINTER(_p)],
            "mdx_batch_destroy": [_p],
            "mdx_batch_set_positions": [_p, _p],
            "mdx_batch_set_precision": [_p, _i],
            "mdx_batch_set_skin": [_p, _d],
            "mdx_batch_lj_init": [_p, _p, _p, _p],
            "mdx_batch_lj_forces": [_p, _p, _p],
            "mdx_batch_lj_vjp": [_p, _p, _p, _p],
            "mdx_batch_eam_init": [_p, _i, _d, _i, _d, _d, _p, _p, _p],
            "mdx_batch_eam_forces": [_p, _p, _p, _p, _p],
            "mdx_batch_eam_vjp": [_p, _p, _p, _p, _p],
            "mdx_batch_set_launch": [_p, _i, _i, _i, _i],
            "mdx_batch_get_launch": [_p, _i, _p, _p, _p],
            "mdx_batch_autotune": [_p, _i],
            "mdx_batch_kernel_time": [_p, _i, _p, _p],
            "mdx_batch_neigh_stats": [_p, _p, _p],
            "mdx_batch_set_stream": [_p, _i, _p],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _i
        _configured = True
    return L


def _addr(a):
    """Address of a numpy array or a torch tensor (None passes NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous() or str(a.dtype) != "torch.float64":
            raise ValueError("tensors must be contiguous float64")
        return a.data_ptr()
    return a.ctypes.data


class Batch:
    """Independent configurations in one device object.

    natoms [nrep]; lo, hi, periodic [nrep, 3]; types [N] (1-based)."""

    def __init__(self, natoms, lo, hi, periodic, ntypes, types, device=0):
        self._L = lib()
        self.natoms = np.ascontiguousarray(natoms, dtype=np.int32)
        self.nrep = len(self.natoms)
        self.n = int(self.natoms.sum())
        self.ntypes = ntypes
        self._lo = np.ascontiguousarray(lo, dtype=np.float64).reshape(self.nrep, 3)
        self._hi = np.ascontiguousarray(hi, dtype=np.float64).reshape(self.nrep, 3)
        self._per = np.ascontiguousarray(periodic, dtype=np.int32).reshape(self.nrep, 3)
        self._types = np.ascontiguousarray(types, dtype=np.int32)
        h = _p()
        self._ck(self._L.mdx_batch_create(device, self.nrep, self.natoms.ctypes.data, self._lo.ctypes.data,
                                          self._hi.ctypes.data, self._per.ctypes.data, ntypes,
                                          self._types.ctypes.data, ctypes.byref(h)))
        self.h = h
        self.eam_sizes = None

    def close(self):
        if getattr(self, "h", None):
            self._L.mdx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ck(self, rc):
        if rc != 0:
            raise _mdx.MdxError(rc, self._L.mdx_last_error().decode())

    def set_positions(self, x):
        if not hasattr(x, "data_ptr"):
            x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.n, 3)
        self._keep_x = x
        self._ck(self._L.mdx_batch_set_positions(self.h, _addr(x)))

    def set_stream(self, stream=None):
        """Queue the batch's work on `stream` (a torch.cuda.Stream, or an
        integer handle; None: the batch's own stream)."""
        if stream is None:
            self._ck(self._L.mdx_batch_set_stream(self.h, 0, None))
        else:
            handle = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)
            self._ck(self._L.mdx_batch_set_stream(self.h, 1, ctypes.c_void_p(handle)))

    def precision(self, p):
        self._ck(self._L.mdx_batch_set_precision(self.h, p))

    def skin(self, s):
        self._ck(self._L.mdx_batch_set_skin(self.h, float(s)))

    def _tab(self, a):
        nt1 = self.ntypes + 1
        return np.ascontiguousarray(a, dtype=np.float64).reshape(nt1, nt1)

    def lj_init(self, eps, sigma, cut):
        e, s, c = self._tab(eps), self._tab(sigma), self._tab(cut)
        self._ck(self._L.mdx_batch_lj_init(self.h, e.ctypes.data, s.ctypes.data, c.ctypes.data))

    def lj_forces(self, out=None, energy=True):
        f = np.zeros((self.n, 3)) if out is None else out
        e = np.zeros(self.nrep) if energy else None
        self._ck(self._L.mdx_batch_lj_forces(self.h, _addr(f), _addr(e)))
        return f, e

    def lj_vjp(self, gf):
        nt1 = self.ntypes + 1
        ge, gs = np.zeros((nt1, nt1)), np.zeros((nt1, nt1))
        if not hasattr(gf, "data_ptr"):
            gf = np.ascontiguousarray(gf, dtype=np.float64).reshape(self.n, 3)
        self._ck(self._L.mdx_batch_lj_vjp(self.h, _addr(gf), ge.ctypes.data, gs.ctypes.data))
        return ge, gs

    def eam_init(self, drho, dr, cut, frho, rhor, z2r):
        self._eam = [np.ascontiguousarray(a, dtype=np.float64) for a in (frho, rhor, z2r)]
        fr, rr, zr = self._eam
        if len(rr) != len(zr):
            raise ValueError("rhor and z2r need the same number of knots")
        self.eam_sizes = (len(fr), len(rr))
        self._ck(self._L.mdx_batch_eam_init(self.h, len(fr), float(drho), len(rr), float(dr), float(cut),
                                            fr.ctypes.data, rr.ctypes.data, zr.ctypes.data))

    def eam_forces(self, out=None, energy=True, rho_fp=False):
        f = np.zeros((self.n, 3)) if out is None else out
        e = np.zeros(self.nrep) if energy else None
        rho = np.zeros(self.n) if rho_fp else None
        fp = np.zeros(self.n) if rho_fp else None
        self._ck(self._L.mdx_batch_eam_forces(self.h, _addr(f), _addr(e), _addr(rho), _addr(fp)))
        return (f, e, rho, fp) if rho_fp else (f, e)

    def eam_vjp(self, gf):
        nrho, nr = self.eam_sizes
        a, b, c = np.zeros(nrho), np.zeros(nr), np.zeros(nr)
        if not hasattr(gf, "data_ptr"):
            gf = np.ascontiguousarray(gf, dtype=np.float64).reshape(self.n, 3)
        self._ck(self._L.mdx_batch_eam_vjp(self.h, _addr(gf), a.ctypes.data, b.ctypes.data, c.ctypes.data))
        return a, b, c

    def set_launch(self, kind, block, tpa, grid=0):
        self._ck(self._L.mdx_batch_set_launch(self.h, kind, block, tpa, grid))

    def get_launch(self, kind):
        v = np.zeros(3, dtype=np.int32)
        self._ck(self._L.mdx_batch_get_launch(self.h, kind, v[0:].ctypes.data, v[1:].ctypes.data,
                                              v[2:].ctypes.data))
        return dict(block=int(v[0]), tpa=int(v[1]), grid=int(v[2]))

    def autotune(self, on=True):
        self._ck(self._L.mdx_batch_autotune(self.h, 1 if on else 0))

    def kernel_time(self, kind):
        ms = ctypes.c_double(0)
        n = ctypes.c_long(0)
        self._ck(self._L.mdx_batch_kernel_time(self.h, kind, ctypes.addressof(ms), ctypes.addressof(n)))
        return ms.value, n.value

    def neigh_stats(self):
        t = ctypes.c_long(0)
        m = ctypes.c_int(0)
        self._ck(self._L.mdx_batch_neigh_stats(self.h, ctypes.addressof(t), ctypes.addressof(m)))
        return t.value, m.value
