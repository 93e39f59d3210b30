"""ctypes binding of libmdx (include/mdx.h) for tests, bench.py and scripts.

Loads the in-tree lammps_amd/libmdx.so and fails loudly when it is missing:
there is no CPU fallback for the device path.

A process holds one HIP runtime.  PyTorch's ROCm wheel carries its own copy
of the runtime under the soname libmdx.so needs (libamdhip64.so.7); two
copies in one process do not both see the GPU.  So torch, when installed, is
imported before the library is loaded: the loader then binds libmdx.so to
torch's copy and device tensors and the library share one runtime.
MDX_HIP_RUNTIME=rocm skips the import (ROCm's runtime; no torch in the
process).
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdx.so")


def hip_runtime_paths():
    """Files of the HIP runtime(s) mapped into this process (Linux)."""
    try:
        with open("/proc/self/maps") as fh:
            return sorted({ln.split()[-1] for ln in fh if "libamdhip64" in ln and "/" in ln})
    except OSError:
        return []


FP64, MIXED = 0, 1
POS_FP64, POS_Q32 = 0, 1
OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_OVERFLOW = 0, -1, -2, -3, -4
NOALLOC_MAX_ATOMS = 32768  # MDX_NOALLOC_MAX_ATOMS (include/mdx.h)

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None
LOADED_FLAVOUR = None  # at the first load: 'torch' (torch's HIP runtime) or 'rocm'

THERMO_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_long, ctypes.POINTER(ctypes.c_double))


class Stats(ctypes.Structure):
    _fields_ = [("ncalls", ctypes.c_long), ("ndanger", ctypes.c_long), ("total_neighbors", ctypes.c_long),
                ("max_neighbors", ctypes.c_int), ("capacity", ctypes.c_int), ("nbins", ctypes.c_int * 3),
                ("kernel_ms", ctypes.c_double * 8), ("kernel_launches", ctypes.c_long * 8)]


K_BUILD, K_LJ, K_EAM_RHO, K_EAM_FORCE, K_NVE_INITIAL, K_NVE_FINAL, K_FUSED, K_HALO = range(8)


class MdxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mdx error %d: %s" % (code, msg))
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if os.environ.get("MDX_HIP_RUNTIME", "") != "rocm":
            try:
                import torch  # noqa: F401  (its HIP runtime becomes the process's)
            except ImportError:
                pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmdx.so is not built (run __graft_entry__.build() or python -m lammps_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        global LOADED_FLAVOUR
        LOADED_FLAVOUR = "torch" if "torch" in sys.modules else "rocm"
        sig = {
            "mdx_version": [],
            "mdx_create": [_i, ctypes.POINTER(_p)],
            "mdx_destroy": [_p],
            "mdx_set_box": [_p, _dp, _dp, _ip],
            "mdx_set_atoms": [_p, _i, _i, _dp, _dp, _ip, _ip, _dp],
            "mdx_set_positions": [_p, _dp],
            "mdx_get_atoms": [_p, _dp, _dp, _dp, _ip],
            "mdx_pair_lj_init": [_p, _i, _dp, _dp, _dp, _dp, _dp, _dp],
            "mdx_pair_eam_init": [_p, _i, _i, _d, _i, _d, _d, _d, _i, _dp, _i, _dp, _i, _dp, _ip, _ip, _ip],
            "mdx_set_precision": [_p, _i],
            "mdx_neigh_modify": [_p, _d, _i, _i, _i, _i],
            "mdx_neigh_build": [_p],
            "mdx_neigh_pairs": [_p, _ip, ctypes.c_long, ctypes.POINTER(ctypes.c_long)],
            "mdx_compute": [_p, _i, _i, _dp, _dp, _dp, _dp, _dp],
            "mdx_eam_rho_fp": [_p, _dp, _dp],
            "mdx_setup": [_p, THERMO_FN, _p],
            "mdx_run": [_p, _i, _d, _d, _i, THERMO_FN, _p],
            "mdx_get_stats": [_p, ctypes.POINTER(Stats)],
            "mdx_set_launch": [_p, _i, _i],
            "mdx_copy_bandwidth": [_p, ctypes.c_long, _i, _dp],
            "mdx_set_fused": [_p, _i],
            "mdx_set_profiling": [_p, _i],
            "mdx_set_position_format": [_p, _i],
            "mdx_get_launch": [_p, _ip, _ip],
            "mdx_autotune": [_p, _i],
            "mdx_time_step_kernel": [_p, _i, _d, _d, _p],
            "mdx_halo_sendlist": [_p, _i, _dp, _dp, _dp, _ip],
            "mdx_halo_sendlists": [_p, _i, _dp, _dp, _dp, _ip],
            "mdx_halo_pack": [_p, _i, ctypes.c_void_p],
            "mdx_halo_unpack": [_p, _i, _i, ctypes.c_void_p],
            "mdx_compute_noalloc": [_p, _i, ctypes.c_void_p, ctypes.c_void_p, _i, _i, ctypes.c_void_p, _dp, _dp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _i
        L.mdx_last_error.restype = ctypes.c_char_p
        L.mdx_last_error.argtypes = []
        _lib = L
    return _lib


def exported_symbols():
    """Names every entry point that include/mdx.h declares (for load tests)."""
    return ["mdx_last_error", "mdx_version", "mdx_create", "mdx_destroy", "mdx_set_box", "mdx_set_atoms",
            "mdx_set_positions", "mdx_get_atoms", "mdx_pair_lj_init", "mdx_pair_eam_init", "mdx_set_precision",
            "mdx_neigh_modify", "mdx_neigh_build", "mdx_neigh_pairs", "mdx_compute", "mdx_eam_rho_fp",
            "mdx_setup", "mdx_run", "mdx_get_stats", "mdx_set_profiling", "mdx_set_launch"]


def _ptr(a, kind=_dp):
    return None if a is None else a.ctypes.data_as(kind)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    """One device context (all device memory of one simulated system)."""

    def __init__(self, device=0):
        self._L = lib()
        h = _p()
        self._ck(self._L.mdx_create(device, ctypes.byref(h)))
        self.h = h
        self.n = 0
        self._thermo_rows = []
        self._cb = THERMO_FN(self._on_thermo)

    def close(self):
        if getattr(self, "h", None):
            self._L.mdx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc != 0:
            raise MdxError(rc, self._L.mdx_last_error().decode())
        return rc

    def _on_thermo(self, user, step, row):
        self._thermo_rows.append((int(step), [row[k] for k in range(8)]))

    # --- state ---
    def set_box(self, lo, hi, periodic=(1, 1, 1)):
        lo, hi, per = _f64(lo), _f64(hi), _i32(periodic)
        self._ck(self._L.mdx_set_box(self.h, _ptr(lo), _ptr(hi), _ptr(per, _ip)))

    def set_atoms(self, x, types, mass, v=None, tags=None):
        x = _f64(x)
        t = _i32(types)
        n = len(t)
        mass = _f64(mass)
        ntypes = len(mass) - 1
        v = None if v is None else _f64(v)
        tags = None if tags is None else _i32(tags)
        self._ck(self._L.mdx_set_atoms(self.h, n, ntypes, _ptr(x), _ptr(v), _ptr(t, _ip), _ptr(tags, _ip),
                                       _ptr(mass)))
        self.n = n

    def set_positions(self, x):
        x = _f64(x)
        self._ck(self._L.mdx_set_positions(self.h, _ptr(x)))

    def get_atoms(self):
        n = self.n
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3)); im = np.zeros((n, 3), dtype=np.int32)
        self._ck(self._L.mdx_get_atoms(self.h, _ptr(x), _ptr(v), _ptr(f), _ptr(im, _ip)))
        return x, v, f, im

    # --- styles ---
    def pair_lj(self, ntypes, cutsq, lj1, lj2, lj3, lj4, offset):
        a = [_f64(q).ravel() for q in (cutsq, lj1, lj2, lj3, lj4, offset)]
        self._ck(self._L.mdx_pair_lj_init(self.h, ntypes, *[_ptr(q) for q in a]))

    def pair_lj_from_coeffs(self, ntypes, coeffs, shift=False):
        """coeffs {(i,j): (eps, sigma, cut)} for i <= j with every (i,i) set;
        mixing (geometric) and init_one as src/pair_lj_cut.cpp:545-602."""
        nt1 = ntypes + 1
        eps = np.zeros((nt1, nt1)); sig = np.zeros((nt1, nt1)); cut = np.zeros((nt1, nt1))
        for (i, j), (e, s, c) in coeffs.items():
            i, j = min(i, j), max(i, j)
            eps[i, j], sig[i, j], cut[i, j] = e, s, c
        tabs = {k: np.zeros((nt1, nt1)) for k in ("cutsq", "lj1", "lj2", "lj3", "lj4", "offset")}
        for i in range(1, nt1):
            for j in range(i, nt1):
                if (i, j) not in coeffs and (j, i) not in coeffs:
                    eps[i, j] = np.sqrt(eps[i, i] * eps[j, j])
                    sig[i, j] = np.sqrt(sig[i, i] * sig[j, j])
                    cut[i, j] = np.sqrt(cut[i, i] * cut[j, j])
                e, s, c = eps[i, j], sig[i, j], cut[i, j]
                vals = {"cutsq": c * c, "lj1": 48.0 * e * s ** 12.0, "lj2": 24.0 * e * s ** 6.0,
                        "lj3": 4.0 * e * s ** 12.0, "lj4": 4.0 * e * s ** 6.0,
                        "offset": 4.0 * e * ((s / c) ** 12.0 - (s / c) ** 6.0) if shift else 0.0}
                for k, val in vals.items():
                    tabs[k][i, j] = tabs[k][j, i] = val
        self.pair_lj(ntypes, tabs["cutsq"], tabs["lj1"], tabs["lj2"], tabs["lj3"], tabs["lj4"], tabs["offset"])
        return tabs

    def pair_eam(self, ntypes, nrho, drho, nr, dr, rhomax, cutforcesq, frho_spline, rhor_spline, z2r_spline,
                 type2frho, type2rhor, type2z2r):
        fr, rr, zr = _f64(frho_spline), _f64(rhor_spline), _f64(z2r_spline)
        self._ck(self._L.mdx_pair_eam_init(
            self.h, ntypes, nrho, drho, nr, dr, rhomax, cutforcesq, fr.shape[0], _ptr(fr), rr.shape[0], _ptr(rr),
            zr.shape[0], _ptr(zr), _ptr(_i32(type2frho), _ip), _ptr(_i32(type2rhor), _ip),
            _ptr(_i32(type2z2r), _ip)))

    def precision(self, p):
        self._ck(self._L.mdx_set_precision(self.h, p))

    def position_format(self, fmt):
        """POS_Q32 (default) or POS_FP64 neighbour positions in the mixed gathers."""
        self._ck(self._L.mdx_set_position_format(self.h, fmt))

    def halo_sendlist(self, swap, lo, hi, shift):
        """Atoms with lo <= x < hi as ghost-shell send list `swap`, carrying the
        image shift `shift`; returns their count."""
        n = ctypes.c_int(0)
        lo = np.asarray(lo, dtype=np.float64); hi = np.asarray(hi, dtype=np.float64)
        sh = np.asarray(shift, dtype=np.float64)
        self._ck(self._L.mdx_halo_sendlist(self.h, swap, _ptr(lo), _ptr(hi), _ptr(sh), ctypes.byref(n)))
        return n.value

    def halo_sendlists(self, regions):
        """Send lists 0 .. len(regions)-1 from [(lo, hi, shift), ...] in one pass; returns the counts."""
        lo = np.ascontiguousarray([r[0] for r in regions], dtype=np.float64)
        hi = np.ascontiguousarray([r[1] for r in regions], dtype=np.float64)
        sh = np.ascontiguousarray([r[2] for r in regions], dtype=np.float64)
        counts = np.zeros(len(regions), dtype=np.int32)
        self._ck(self._L.mdx_halo_sendlists(self.h, len(regions), _ptr(lo), _ptr(hi), _ptr(sh), _ptr(counts, _ip)))
        return counts.tolist()

    def halo_pack(self, swap, dev_ptr):
        """x + shift of send list `swap` (-1: all, in swap order) into device
        memory at dev_ptr (3 doubles per atom)."""
        self._ck(self._L.mdx_halo_pack(self.h, swap, ctypes.c_void_p(dev_ptr)))

    def halo_unpack(self, first, count, dev_ptr):
        """Positions at device address dev_ptr into atoms first .. first+count-1."""
        self._ck(self._L.mdx_halo_unpack(self.h, first, count, ctypes.c_void_p(dev_ptr)))

    def neigh_modify(self, skin, every=1, delay=10, check=True, one=2000):
        self._ck(self._L.mdx_neigh_modify(self.h, skin, every, delay, 1 if check else 0, one))

    def set_launch(self, block=0, tpa=0):
        self._ck(self._L.mdx_set_launch(self.h, block, tpa))

    def get_launch(self):
        v = np.zeros(2, dtype=np.int32)
        self._ck(self._L.mdx_get_launch(self.h, _ptr(v[0:], _ip), _ptr(v[1:], _ip)))
        return dict(block=int(v[0]), tpa=int(v[1]))

    def autotune(self, reps=3):
        """Pick threads per atom and block of the pair kernels by timing (mdx_autotune)."""
        self._ck(self._L.mdx_autotune(self.h, reps))
        return self.get_launch()

    def fused(self, on=True):
        self._ck(self._L.mdx_set_fused(self.h, 1 if on else 0))

    def time_step_kernel(self, reps, dt, ftm2v=1.0):
        """Mean duration (ms) of the step's force kernel, launched `reps` times
        back to back between two HIP events (mdx_time_step_kernel)."""
        ms = ctypes.c_double(0.0)
        self._ck(self._L.mdx_time_step_kernel(self.h, reps, dt, ftm2v, ctypes.byref(ms)))
        return ms.value

    def copy_bandwidth(self, nbytes=1 << 30, iters=20):
        g = ctypes.c_double(0.0)
        self._ck(self._L.mdx_copy_bandwidth(self.h, nbytes, iters, ctypes.byref(g)))
        return g.value

    def profiling(self, on=True):
        self._ck(self._L.mdx_set_profiling(self.h, 1 if on else 0))

    # --- work ---
    def neigh_build(self):
        self._ck(self._L.mdx_neigh_build(self.h))

    def neigh_pairs(self):
        cnt = ctypes.c_long(0)
        self._ck(self._L.mdx_neigh_pairs(self.h, None, 0, ctypes.byref(cnt)))
        buf = np.zeros((cnt.value, 2), dtype=np.int32)
        self._ck(self._L.mdx_neigh_pairs(self.h, _ptr(buf, _ip), cnt.value, ctypes.byref(cnt)))
        return buf

    def compute(self, eflag=1, vflag=1, peratom=False):
        n = self.n
        f = np.zeros((n, 3))
        e = ctypes.c_double(0.0)
        w = np.zeros(6)
        ea = np.zeros(n) if peratom else None
        va = np.zeros((n, 6)) if peratom else None
        self._ck(self._L.mdx_compute(self.h, eflag, vflag, _ptr(f), ctypes.byref(e), _ptr(w), _ptr(ea), _ptr(va)))
        return {"f": f, "eng": e.value, "virial": w, "eatom": ea, "vatom": va}

    def compute_noalloc(self, n, x_dev, type_dev, f_dev, eflag=1, vflag=1):
        """mdx_compute_noalloc on device arrays given as integer addresses
        (x [n][3] and f [n][3] doubles, type int32[n], device memory of the
        context's GPU): f += the pair forces; returns (eng_vdwl, virial)."""
        e = ctypes.c_double(0.0)
        w = np.zeros(6)
        self._ck(self._L.mdx_compute_noalloc(self.h, n, ctypes.c_void_p(x_dev), ctypes.c_void_p(type_dev), eflag,
                                             vflag, ctypes.c_void_p(f_dev), ctypes.byref(e), _ptr(w)))
        return e.value, w

    def eam_rho_fp(self):
        r = np.zeros(self.n); fp = np.zeros(self.n)
        self._ck(self._L.mdx_eam_rho_fp(self.h, _ptr(r), _ptr(fp)))
        return r, fp

    def setup(self):
        self._thermo_rows = []
        self._ck(self._L.mdx_setup(self.h, self._cb, None))
        return list(self._thermo_rows)

    def run(self, nsteps, dt, ftm2v=1.0, thermo_every=0):
        self._thermo_rows = []
        self._ck(self._L.mdx_run(self.h, nsteps, dt, ftm2v, thermo_every, self._cb, None))
        return list(self._thermo_rows)

    def stats(self):
        s = Stats()
        self._ck(self._L.mdx_get_stats(self.h, ctypes.byref(s)))
        return {k: (list(getattr(s, k)) if k in ("nbins", "kernel_ms", "kernel_launches") else getattr(s, k))
                for k, _ in Stats._fields_}


LATTICE = {"sc": 0, "bcc": 1, "fcc": 2}
UNITS = {"lj": 0, "metal": 1}
# update.cpp:118-180
UNIT_CONSTANTS = {"lj": dict(boltz=1.0, mvv2e=1.0, ftm2v=1.0, nktv2p=1.0),
                  "metal": dict(boltz=8.617343e-5, mvv2e=1.0364269e-4, ftm2v=1.0 / 1.0364269e-4,
                                nktv2p=1.6021765e6)}


class EamInfo(ctypes.Structure):
    _fields_ = [("nrho", ctypes.c_int), ("nr", ctypes.c_int), ("nfrho", ctypes.c_int), ("nrhor", ctypes.c_int),
                ("nz2r", ctypes.c_int), ("drho", ctypes.c_double), ("dr", ctypes.c_double),
                ("rhomax", ctypes.c_double), ("cutmax", ctypes.c_double), ("mass", ctypes.c_double)]


def _host_lib():
    L = lib()
    if not getattr(L, "_host_ready", False):
        L.mdx_host_lattice.argtypes = [_i, _d, _i, _ip, _ip, _ip, _dp, _dp, _dp, ctypes.c_long,
                                       ctypes.POINTER(ctypes.c_long)]
        L.mdx_host_lattice.restype = _i
        L.mdx_host_velocity.argtypes = [ctypes.c_long, _dp, _ip, _dp, _d, _i, _i, _i, _i, _d, _d, _dp]
        L.mdx_host_velocity.restype = _i
        L.mdx_host_eam_funcfl.argtypes = [ctypes.c_char_p, ctypes.POINTER(EamInfo), _dp, _dp, _dp]
        L.mdx_host_eam_funcfl.restype = _i
        L.mdx_host_write_synthetic_funcfl.argtypes = [ctypes.c_char_p, _i, _d, _i, _i, _d, _d, _d, _dp]
        L.mdx_host_write_synthetic_funcfl.restype = _i
        L.mdx_host_eam_setfl.argtypes = [ctypes.c_char_p, _i, _i, ctypes.c_void_p, ctypes.POINTER(EamInfo), _dp, _dp,
                                         _dp, _ip, _ip, _ip, _dp]
        L.mdx_host_eam_setfl.restype = _i
        L.mdx_host_ranpark_uniform.argtypes = [_ip]
        L.mdx_host_ranpark_uniform.restype = _d
        L._host_ready = True
    return L


def _host_ck(rc):
    if rc != 0:
        raise MdxError(rc, lib().mdx_last_error().decode())


def lattice(style, arg, units, lo, hi, periodic=(1, 1, 1)):
    """create_atoms on a lattice: returns (boxlo, boxhi, x[n][3]) in atom-ID order."""
    L = _host_lib()
    lo, hi, per = _i32(lo), _i32(hi), _i32(periodic)
    blo, bhi = np.zeros(3), np.zeros(3)
    n = ctypes.c_long(0)
    _host_ck(L.mdx_host_lattice(LATTICE[style], arg, UNITS[units], _ptr(lo, _ip), _ptr(hi, _ip), _ptr(per, _ip),
                                _ptr(blo), _ptr(bhi), None, 0, ctypes.byref(n)))
    x = np.zeros((n.value, 3))
    _host_ck(L.mdx_host_lattice(LATTICE[style], arg, UNITS[units], _ptr(lo, _ip), _ptr(hi, _ip), _ptr(per, _ip),
                                _ptr(blo), _ptr(bhi), _ptr(x), n.value, ctypes.byref(n)))
    return blo, bhi, x


def velocity_create(x, types, mass, temp, seed, dist="uniform", loop="all", mom=True, units="lj"):
    L = _host_lib()
    x, t, mass = _f64(x), _i32(types), _f64(mass)
    v = np.zeros_like(x)
    u = UNIT_CONSTANTS[units]
    _host_ck(L.mdx_host_velocity(len(t), _ptr(x), _ptr(t, _ip), _ptr(mass), temp, seed,
                                 1 if dist == "gaussian" else 0, 1 if loop == "geom" else 0, 1 if mom else 0,
                                 u["boltz"], u["mvv2e"], _ptr(v)))
    return v


def eam_funcfl(path):
    """PairEAM tables of one funcfl file: dict with the sizes and the
    frho/rhor/z2r spline arrays ([k][n+1][7])."""
    L = _host_lib()
    info = EamInfo()
    _host_ck(L.mdx_host_eam_funcfl(path.encode(), ctypes.byref(info), None, None, None))
    fr = np.zeros((info.nfrho, info.nrho + 1, 7))
    rr = np.zeros((info.nrhor, info.nr + 1, 7))
    zr = np.zeros((info.nz2r, info.nr + 1, 7))
    _host_ck(L.mdx_host_eam_funcfl(path.encode(), ctypes.byref(info), _ptr(fr), _ptr(rr), _ptr(zr)))
    return dict(nrho=info.nrho, nr=info.nr, drho=info.drho, dr=info.dr, rhomax=info.rhomax, cutmax=info.cutmax,
                mass=info.mass, frho_spline=fr, rhor_spline=rr, z2r_spline=zr)


def eam_setfl(path, elements, fs=False):
    """pair_style eam/alloy (fs=False) or eam/fs + pair_coeff * * path elements...:
    sizes, spline arrays, type maps and per-type masses (mdx_host_eam_setfl)."""
    L = _host_lib()
    nt = len(elements)
    names = (ctypes.c_char_p * nt)(*[e.encode() for e in elements])
    info = EamInfo()
    args = (path.encode(), 1 if fs else 0, nt, ctypes.cast(names, ctypes.c_void_p), ctypes.byref(info))
    _host_ck(L.mdx_host_eam_setfl(*args, None, None, None, None, None, None, None))
    fr = np.zeros((info.nfrho, info.nrho + 1, 7))
    rr = np.zeros((info.nrhor, info.nr + 1, 7))
    zr = np.zeros((info.nz2r, info.nr + 1, 7))
    nt1 = nt + 1
    t2f = np.zeros(nt1, dtype=np.int32)
    t2r = np.zeros(nt1 * nt1, dtype=np.int32)
    t2z = np.zeros(nt1 * nt1, dtype=np.int32)
    mass = np.zeros(nt1)
    _host_ck(L.mdx_host_eam_setfl(*args, _ptr(fr), _ptr(rr), _ptr(zr), _ptr(t2f, _ip), _ptr(t2r, _ip),
                                  _ptr(t2z, _ip), _ptr(mass)))
    return dict(nrho=info.nrho, nr=info.nr, drho=info.drho, dr=info.dr, rhomax=info.rhomax, cutmax=info.cutmax,
                frho_spline=fr, rhor_spline=rr, z2r_spline=zr, type2frho=t2f, type2rhor=t2r, type2z2r=t2z,
                mass=mass)


def eam_context_setfl(ctx, path, elements, fs=False):
    """pair_style eam/alloy|eam/fs + pair_coeff * * path elements on a context; returns the tables."""
    t = eam_setfl(path, elements, fs)
    ctx.pair_eam(len(elements), t["nrho"], t["drho"], t["nr"], t["dr"], t["rhomax"], t["cutmax"] ** 2,
                 t["frho_spline"], t["rhor_spline"], t["z2r_spline"], t["type2frho"], t["type2rhor"], t["type2z2r"])
    return t


def write_synthetic_funcfl(path, seed=4951, a=3.615, nrho=500, nr=500, dr=0.01, cut=4.95, mass=63.55):
    L = _host_lib()
    params = np.zeros(6)
    _host_ck(L.mdx_host_write_synthetic_funcfl(path.encode(), seed, a, nrho, nr, dr, cut, mass, _ptr(params)))
    return dict(zip(("Z0", "beta", "chi", "E0", "re", "rho_e"), params.tolist()))


def eam_context_tables(ctx, path, ntypes=1):
    """pair_style eam + pair_coeff * * path on a context; returns the tables."""
    t = eam_funcfl(path)
    nt1 = ntypes + 1
    ctx.pair_eam(ntypes, t["nrho"], t["drho"], t["nr"], t["dr"], t["rhomax"], t["cutmax"] ** 2, t["frho_spline"],
                 t["rhor_spline"], t["z2r_spline"], np.zeros(nt1, dtype=np.int32),
                 np.zeros(nt1 * nt1, dtype=np.int32), np.zeros(nt1 * nt1, dtype=np.int32))
    return t


def octant_shell_regions(L, cut):
    """Send regions of a periodic cubic box [0, L)^3 for the ghost shell of its
    octant [0, L/2)^3 (one GPU's share when the box is split 2x2x2): for each
    of the 26 neighbouring directions, (lo, hi, shift) such that the atoms
    with lo <= x < hi, shifted by `shift`, are the ghosts within `cut` of that
    face, edge or corner (CommBrick::borders' selection, done in one pass per
    direction).  Requires L > 4 cut."""
    h = L / 2.0
    if not L > 4.0 * cut:
        raise ValueError("octant shell regions need L > 4 cut")
    regions = []
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for d in (-1, 0, 1):
                if (a, b, d) == (0, 0, 0):
                    continue
                lo, hi, sh = [], [], []
                for side in (a, b, d):
                    if side < 0:
                        lo.append(L - cut); hi.append(L); sh.append(-L)
                    elif side == 0:
                        lo.append(0.0); hi.append(h); sh.append(0.0)
                    else:
                        lo.append(h); hi.append(h + cut); sh.append(0.0)
                regions.append((lo, hi, sh))
    return regions


def thermo_columns(raw, natoms, prd, units="lj"):
    """Thermo row (Temp, E_pair, E_mol, TotEng, Press) from the raw sums the
    library reports, with thermo.cpp's normalisation (lj: per atom) and
    compute_temp/compute_pressure's formulas (dof = 3N-3)."""
    if units == "lj":
        boltz, mvv2e, nktv2p, norm = 1.0, 1.0, 1.0, True
    else:
        boltz, mvv2e, nktv2p, norm = 8.617343e-5, 1.0364269e-4, 1.6021765e6, False
    mv2, pe = raw[0], raw[1]
    dof = 3.0 * natoms - 3.0
    t = mv2 * mvv2e / (dof * boltz)
    ke = t * 0.5 * dof * boltz
    epair = pe / natoms if norm else pe
    if norm:
        ke /= natoms
    vol = prd[0] * prd[1] * prd[2]
    press = (dof * boltz * t + raw[2] + raw[3] + raw[4]) / 3.0 / vol * nktv2p
    return [t, epair, 0.0, epair + ke, press]
