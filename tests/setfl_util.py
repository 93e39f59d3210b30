"""Writers of DYNAMO setfl files (eam/alloy and eam/fs formats, as
PairEAMAlloy/PairEAMFS::read_file parse them) for the tests: a funcfl
potential re-expressed on its own knots, and a synthetic two-element
Cu-like potential."""
import numpy as np

HARTREE_BOHR = 27.2 * 0.529  # funcfl Z -> z2r factor (pair_eam.cpp:661)


def _emit(fh, vals):
    vals = list(vals)
    for k in range(0, len(vals), 5):
        fh.write(" ".join("%.17g" % v for v in vals[k: k + 5]) + "\n")


def write_setfl(path, names, masses, nrho, drho, nr, dr, cut, frho, rhor, z2r, fs=False):
    """frho[i] (nrho values); rhor[i] (nr values) or, for fs, rhor[i][j];
    z2r[(i, j)] for i >= j (nr values each)."""
    n = len(names)
    with open(path, "w") as fh:
        fh.write("synthetic setfl for tests\nline 2\nline 3\n")
        fh.write("%d %s\n" % (n, " ".join(names)))
        fh.write("%d %.17g %d %.17g %.17g\n" % (nrho, drho, nr, dr, cut))
        for i in range(n):
            fh.write("%d %.17g 3.615 fcc\n" % (29 + i, masses[i]))
            _emit(fh, frho[i])
            if fs:
                for j in range(n):
                    _emit(fh, rhor[i][j])
            else:
                _emit(fh, rhor[i])
        for i in range(n):
            for j in range(i + 1):
                _emit(fh, z2r[(i, j)])


def binary_cu_like(path, fs=False, nrho=300, nr=300, cut=4.95):
    """Two Cu-like elements A, B with different parameters (and, for fs,
    element-pair-dependent density amplitudes)."""
    par = {"A": dict(Z0=10.8, beta=1.5, chi=6.0, E0=3.0, re=2.556, m=63.55),
           "B": dict(Z0=9.6, beta=1.42, chi=5.6, E0=2.5, re=2.49, m=58.69)}
    names = ["A", "B"]
    dr = cut / (nr - 1)

    def dens(p, r, amp=1.0):
        return amp * np.exp(-p["chi"] * (r / p["re"] - 1.0))

    rho_e = {}
    for e in names:
        p = par[e]
        a = p["re"] * np.sqrt(2.0)
        shells = [(12, 0.5), (6, 1.0), (24, 1.5), (12, 2.0)]
        rho_e[e] = sum(k * dens(p, a * np.sqrt(q)) for k, q in shells if a * np.sqrt(q) < cut)
    drho = 2.0 * max(rho_e.values()) / (nrho - 1)
    rgrid = np.arange(nr) * dr
    rhogrid = np.arange(nrho) * drho
    frho = [-par[e]["E0"] * np.sqrt(rhogrid / rho_e[e]) for e in names]
    amp = {("A", "A"): 1.0, ("A", "B"): 0.8, ("B", "A"): 1.3, ("B", "B"): 1.0}
    if fs:
        rhor = [[dens(par[a], rgrid, amp[(a, b)]) for b in names] for a in names]
    else:
        rhor = [dens(par[e], rgrid) for e in names]
    Z = {e: par[e]["Z0"] * np.exp(-par[e]["beta"] * rgrid) for e in names}
    z2r = {(i, j): HARTREE_BOHR * Z[names[i]] * Z[names[j]] for i in range(2) for j in range(i + 1)}
    write_setfl(path, names, [par[e]["m"] for e in names], nrho, drho, nr, dr, cut, frho, rhor, z2r, fs=fs)
    return names


def spline_eval(sp, m, p):
    """value and derivative coefficients of one 7-coefficient spline row"""
    c = sp[m]
    return ((c[3] * p + c[4]) * p + c[5]) * p + c[6], (c[0] * p + c[1]) * p + c[2]
