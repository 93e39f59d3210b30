"""ctypes wrapper over the CPU FP64 oracle (oracle/mdref.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

``Ref`` mirrors the LAMMPS input-script commands the oracle restates
(units, lattice+region+create_box+create_atoms, mass, velocity create,
pair_style/pair_coeff, neighbor/neigh_modify, timestep, run).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmdref.so")
_lib = None

LJ, METAL = 0, 1
HALF_NEWTON, HALF_NONEWTON, FULL = 0, 1, 2
SC, BCC, FCC = 0, 1, 2

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


_LIB_OMP_PATH = os.path.join(_HERE, "_build", "libmdref_omp.so")
_lib_omp = None


def lib_threaded():
    """The threaded CPU-baseline build (-O3 -fopenmp -DMDREF_OMP): the same
    functions; force passes without tallies and the list build on threads."""
    global _lib_omp
    if _lib_omp is None:
        if not os.path.exists(_LIB_OMP_PATH):
            build()
        _lib_omp = _setup(ctypes.CDLL(_LIB_OMP_PATH))
    return _lib_omp


def set_threads(n):
    """Threads of the threaded build (returns the count in effect)."""
    return lib_threaded().mdref_threads(int(n))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _setup(ctypes.CDLL(_LIB_PATH))
    return _lib


def _setup(L):
    if True:
        L.mdref_new.restype = _p
        L.mdref_free.argtypes = [_p]
        L.mdref_error.restype = ctypes.c_char_p
        L.mdref_error.argtypes = [_p]
        for name, args in {
            "mdref_units": [_p, _i],
            "mdref_box": [_p, _dp, _dp, _ip, _i],
            "mdref_lattice_box": [_p, _i, _d, _ip, _ip, _ip, _i],
            "mdref_add_atoms": [_p, _i, _dp, _ip],
            "mdref_set_type": [_p, _i, _ip],
            "mdref_mass": [_p, _i, _d],
            "mdref_velocity_create": [_p, _d, _i, _i, _i, _i],
            "mdref_set_velocities": [_p, _i, _dp],
            "mdref_pair_lj": [_p, _d, _i],
            "mdref_pair_lj_coeff": [_p, _i, _i, _d, _d, _d],
            "mdref_pair_eam_file": [_p, ctypes.c_char_p],
            "mdref_pair_eam_setfl": [_p, ctypes.c_char_p, _i, ctypes.c_void_p],
            "mdref_neighbor": [_p, _d, _i, _i, _i, _i],
            "mdref_newton": [_p, _i],
            "mdref_list_style": [_p, _i],
            "mdref_sortfreq": [_p, _i],
            "mdref_timestep": [_p, _d],
            "mdref_thermo_every": [_p, _i],
            "mdref_setup": [_p],
            "mdref_run": [_p, _i, _dp, _i, _ip],
            "mdref_compute": [_p, _i, _i],
            "mdref_natoms": [_p],
            "mdref_nghost": [_p],
            "mdref_get_by_tag": [_p, _dp, _dp, _dp, _ip, _ip],
            "mdref_get_peratom_by_tag": [_p, _dp, _dp],
            "mdref_get_eam_by_tag": [_p, _dp, _dp],
            "mdref_energy_virial": [_p, _dp, _dp],
            "mdref_thermo_now": [_p, _dp],
            "mdref_neigh_stats": [_p, ctypes.POINTER(ctypes.c_long), _ip, _ip, _ip],
            "mdref_eam_info": [_p, _ip, _dp, _ip, _dp, _dp, _dp, _ip, _ip, _ip],
            "mdref_eam_splines": [_p, _dp, _dp, _dp, _ip, _ip, _ip],
            "mdref_ranpark_reset_coord": [_i, _dp],
        }.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _i
        L.mdref_mass_of.argtypes = [_p, _i]
        L.mdref_mass_of.restype = _d
        L.mdref_list_pairs.argtypes = [_p, _ip, ctypes.c_long]
        L.mdref_list_pairs.restype = ctypes.c_long
        L.mdref_ntimestep.argtypes = [_p]
        L.mdref_ntimestep.restype = ctypes.c_long
        L.mdref_ranpark_uniform_seq.argtypes = [_i, _i, _ip]
        L.mdref_ranpark_uniform_seq.restype = _d
        L.mdref_threads.argtypes = [_i]
        L.mdref_threads.restype = _i
    return L


def _ptr(a, kind=_dp):
    return a.ctypes.data_as(kind)


class RefError(RuntimeError):
    pass


class Ref:
    """One oracle system (one LAMMPS 'world' on one process)."""

    def __init__(self, units=LJ, threaded=False):
        self._L = lib_threaded() if threaded else lib()
        self.h = self._L.mdref_new()
        self._ck(self._L.mdref_units(self.h, units))

    def __del__(self):
        try:
            self._L.mdref_free(self.h)
        except Exception:
            pass

    def _ck(self, rc):
        if rc < 0:
            raise RefError(self._L.mdref_error(self.h).decode())
        return rc

    # --- script commands ---
    def lattice_box(self, style, arg, lo, hi, periodic=(1, 1, 1), ntypes=1):
        lo = np.asarray(lo, dtype=np.int32)
        hi = np.asarray(hi, dtype=np.int32)
        per = np.asarray(periodic, dtype=np.int32)
        self._ck(self._L.mdref_lattice_box(self.h, style, arg, _ptr(lo, _ip), _ptr(hi, _ip),
                                           _ptr(per, _ip), ntypes))

    def box(self, lo, hi, periodic=(1, 1, 1), ntypes=1):
        lo = np.asarray(lo, dtype=np.float64)
        hi = np.asarray(hi, dtype=np.float64)
        per = np.asarray(periodic, dtype=np.int32)
        self._ck(self._L.mdref_box(self.h, _ptr(lo), _ptr(hi), _ptr(per, _ip), ntypes))

    def add_atoms(self, x, types):
        x = np.ascontiguousarray(x, dtype=np.float64)
        t = np.ascontiguousarray(types, dtype=np.int32)
        self._ck(self._L.mdref_add_atoms(self.h, len(t), _ptr(x), _ptr(t, _ip)))

    def set_types(self, types):
        t = np.ascontiguousarray(types, dtype=np.int32)
        self._ck(self._L.mdref_set_type(self.h, len(t), _ptr(t, _ip)))

    def mass(self, t, m):
        self._ck(self._L.mdref_mass(self.h, t, m))

    def velocity_create(self, temp, seed, dist="uniform", loop="all", mom=True):
        self._ck(self._L.mdref_velocity_create(self.h, temp, seed, 1 if dist == "gaussian" else 0,
                                               1 if loop == "geom" else 0, 1 if mom else 0))

    def set_velocities(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._ck(self._L.mdref_set_velocities(self.h, v.shape[0], _ptr(v)))

    def pair_lj(self, cut, shift=False):
        self._ck(self._L.mdref_pair_lj(self.h, cut, 1 if shift else 0))

    def pair_coeff(self, i, j, eps, sigma, cut=-1.0):
        self._ck(self._L.mdref_pair_lj_coeff(self.h, i, j, eps, sigma, cut))

    def pair_eam(self, path):
        self._ck(self._L.mdref_pair_eam_file(self.h, path.encode()))

    def pair_eam_setfl(self, path, elements, fs=False):
        """pair_style eam/alloy (or eam/fs) + pair_coeff * * path elements..."""
        arr = (ctypes.c_char_p * len(elements))(*[e.encode() for e in elements])
        self._ck(self._L.mdref_pair_eam_setfl(self.h, path.encode(), 1 if fs else 0, ctypes.cast(arr, ctypes.c_void_p)))

    def neighbor(self, skin, every=1, delay=10, check=True, one=2000):
        self._ck(self._L.mdref_neighbor(self.h, skin, every, delay, 1 if check else 0, one))

    def newton(self, on):
        self._ck(self._L.mdref_newton(self.h, 1 if on else 0))

    def list_style(self, style):
        self._ck(self._L.mdref_list_style(self.h, style))

    def sortfreq(self, n):
        self._ck(self._L.mdref_sortfreq(self.h, n))

    def timestep(self, dt):
        self._ck(self._L.mdref_timestep(self.h, dt))

    def thermo(self, n):
        self._ck(self._L.mdref_thermo_every(self.h, n))

    def setup(self):
        self._ck(self._L.mdref_setup(self.h))

    def run(self, nsteps, max_rows=10000):
        buf = np.zeros((max_rows, 6))
        n = ctypes.c_int(0)
        self._ck(self._L.mdref_run(self.h, nsteps, _ptr(buf), max_rows, ctypes.byref(n)))
        return buf[: n.value].copy()

    def compute(self, eflag=1, vflag=1):
        self._ck(self._L.mdref_compute(self.h, eflag, vflag))

    # --- accessors ---
    @property
    def natoms(self):
        return self._L.mdref_natoms(self.h)

    @property
    def nghost(self):
        return self._L.mdref_nghost(self.h)

    @property
    def ntimestep(self):
        return self._L.mdref_ntimestep(self.h)

    def state(self):
        """(x, v, f, type, image) ordered by atom ID (row = tag-1)."""
        n = self.natoms
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        t = np.zeros(n, dtype=np.int32); im = np.zeros((n, 3), dtype=np.int32)
        self._ck(self._L.mdref_get_by_tag(self.h, _ptr(x), _ptr(v), _ptr(f), _ptr(t, _ip), _ptr(im, _ip)))
        return x, v, f, t, im

    def peratom(self):
        n = self.natoms
        e = np.zeros(n); w = np.zeros((n, 6))
        self._ck(self._L.mdref_get_peratom_by_tag(self.h, _ptr(e), _ptr(w)))
        return e, w

    def eam_rho_fp(self):
        n = self.natoms
        r = np.zeros(n); fp = np.zeros(n)
        self._ck(self._L.mdref_get_eam_by_tag(self.h, _ptr(r), _ptr(fp)))
        return r, fp

    def energy_virial(self):
        e = ctypes.c_double(0)
        w = np.zeros(6)
        self._ck(self._L.mdref_energy_virial(self.h, ctypes.byref(e), _ptr(w)))
        return e.value, w

    def thermo_now(self):
        row = np.zeros(6)
        self._ck(self._L.mdref_thermo_now(self.h, _ptr(row)))
        return row

    def neigh_stats(self):
        t = ctypes.c_long(0); b = ctypes.c_int(0); d = ctypes.c_int(0); m = ctypes.c_int(0)
        self._ck(self._L.mdref_neigh_stats(self.h, ctypes.byref(t), ctypes.byref(b), ctypes.byref(d),
                                           ctypes.byref(m)))
        return {"total": t.value, "builds": b.value, "dangerous": d.value, "max": m.value}

    def list_pairs(self):
        """Current neighbour list as an (npairs, 2) array of (tag_i, tag_j)."""
        n = self._L.mdref_list_pairs(self.h, None, 0)
        buf = np.zeros((n, 2), dtype=np.int32)
        self._L.mdref_list_pairs(self.h, _ptr(buf, _ip), n)
        return buf

    def eam_tables(self):
        nrho = ctypes.c_int(); nr = ctypes.c_int(); nfrho = ctypes.c_int(); nrhor = ctypes.c_int()
        nz2r = ctypes.c_int()
        drho = ctypes.c_double(); dr = ctypes.c_double(); rhomax = ctypes.c_double(); cutmax = ctypes.c_double()
        self._ck(self._L.mdref_eam_info(self.h, ctypes.byref(nrho), ctypes.byref(drho), ctypes.byref(nr),
                                        ctypes.byref(dr), ctypes.byref(rhomax), ctypes.byref(cutmax),
                                        ctypes.byref(nfrho), ctypes.byref(nrhor), ctypes.byref(nz2r)))
        fr = np.zeros((nfrho.value, nrho.value + 1, 7))
        rr = np.zeros((nrhor.value, nr.value + 1, 7))
        zr = np.zeros((nz2r.value, nr.value + 1, 7))
        nt = 9
        t2f = np.zeros(nt, dtype=np.int32)
        t2r = np.zeros(nt * nt, dtype=np.int32)
        t2z = np.zeros(nt * nt, dtype=np.int32)
        self._ck(self._L.mdref_eam_splines(self.h, _ptr(fr), _ptr(rr), _ptr(zr), _ptr(t2f, _ip),
                                           _ptr(t2r, _ip), _ptr(t2z, _ip)))
        return dict(nrho=nrho.value, drho=drho.value, nr=nr.value, dr=dr.value, rhomax=rhomax.value,
                    cutmax=cutmax.value, frho_spline=fr, rhor_spline=rr, z2r_spline=zr)


def ranpark_uniform_seq(seed, n):
    st = ctypes.c_int(0)
    u = lib().mdref_ranpark_uniform_seq(seed, n, ctypes.byref(st))
    return u, st.value


def ranpark_reset_coord(seed, coord):
    c = np.ascontiguousarray(coord, dtype=np.float64)
    return lib().mdref_ranpark_reset_coord(seed, _ptr(c))
