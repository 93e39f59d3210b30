"""TEST INFRASTRUCTURE ONLY: ctypes binding of the FP64 force-matching
oracle (oracle/fitref.c).  Only tests, smoke() and bench.py's CPU baseline
may use it."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libfitref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB)
        L.fitref_new.restype = ctypes.c_void_p
        L.fitref_new.argtypes = [ctypes.c_int, _ip, _dp, _dp, _ip, ctypes.c_int, _ip, _dp]
        L.fitref_free.argtypes = [ctypes.c_void_p]
        L.fitref_error.restype = ctypes.c_char_p
        L.fitref_error.argtypes = [ctypes.c_void_p]
        L.fitref_lj_forces.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.fitref_lj_vjp.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.fitref_eam_set.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, _dp, _dp, _dp]
        L.fitref_eam_forces.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.fitref_eam_vjp.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.fitref_spline.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def spline(knots, delta):
    """7-coefficient spline table (n+1, 7) of the knot values."""
    k = np.ascontiguousarray(knots, dtype=np.float64)
    out = np.zeros((len(k) + 1, 7))
    lib().fitref_spline(len(k), float(delta), _d(k), _d(out))
    return out


class Batch:
    """Independent configurations: natoms[r] atoms each, box lo/hi [nrep,3],
    periodic [nrep,3], types [N] (1-based), positions x [N,3]."""

    def __init__(self, natoms, lo, hi, periodic, ntypes, types, x):
        self.natoms = np.ascontiguousarray(natoms, dtype=np.int32)
        self.nrep = len(self.natoms)
        self.n = int(self.natoms.sum())
        self.ntypes = ntypes
        lo = np.ascontiguousarray(lo, dtype=np.float64).reshape(self.nrep, 3)
        hi = np.ascontiguousarray(hi, dtype=np.float64).reshape(self.nrep, 3)
        per = np.ascontiguousarray(periodic, dtype=np.int32).reshape(self.nrep, 3)
        t = np.ascontiguousarray(types, dtype=np.int32)
        xx = np.ascontiguousarray(x, dtype=np.float64).reshape(self.n, 3)
        self.h = lib().fitref_new(self.nrep, _i(self.natoms), _d(lo), _d(hi), _i(per), ntypes, _i(t), _d(xx))
        if not self.h:
            raise ValueError("fitref_new failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().fitref_free(self.h)
            self.h = None

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(lib().fitref_error(self.h).decode())

    def _tab(self, a):
        nt1 = self.ntypes + 1
        return np.ascontiguousarray(a, dtype=np.float64).reshape(nt1, nt1)

    def lj_forces(self, eps, sigma, cut):
        f = np.zeros((self.n, 3))
        e = np.zeros(self.nrep)
        self._ok(lib().fitref_lj_forces(self.h, _d(self._tab(eps)), _d(self._tab(sigma)), _d(self._tab(cut)),
                                        _d(f), _d(e)))
        return f, e

    def lj_vjp(self, eps, sigma, cut, gf):
        nt1 = self.ntypes + 1
        ge = np.zeros((nt1, nt1))
        gs = np.zeros((nt1, nt1))
        g = np.ascontiguousarray(gf, dtype=np.float64).reshape(self.n, 3)
        self._ok(lib().fitref_lj_vjp(self.h, _d(self._tab(eps)), _d(self._tab(sigma)), _d(self._tab(cut)), _d(g),
                                     _d(ge), _d(gs)))
        return ge, gs

    def eam_set(self, drho, dr, cut, frho, rhor, z2r):
        self.eam_knots = [np.ascontiguousarray(a, dtype=np.float64) for a in (frho, rhor, z2r)]
        fr, rr, zr = self.eam_knots
        self._ok(lib().fitref_eam_set(self.h, len(fr), float(drho), len(rr), float(dr), float(cut), _d(fr), _d(rr),
                                      _d(zr)))

    def eam_forces(self):
        f = np.zeros((self.n, 3))
        e = np.zeros(self.nrep)
        rho = np.zeros(self.n)
        fp = np.zeros(self.n)
        self._ok(lib().fitref_eam_forces(self.h, _d(f), _d(e), _d(rho), _d(fp)))
        return f, e, rho, fp

    def eam_vjp(self, gf):
        fr, rr, zr = self.eam_knots
        g = np.ascontiguousarray(gf, dtype=np.float64).reshape(self.n, 3)
        a, b, c = np.zeros(len(fr)), np.zeros(len(rr)), np.zeros(len(zr))
        self._ok(lib().fitref_eam_vjp(self.h, _d(g), _d(a), _d(b), _d(c)))
        return a, b, c
