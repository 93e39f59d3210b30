// host_setup.cpp - the reference's set-up commands on the host (include/mdx_host.h).
//
// Restates create_atoms on a lattice, velocity create (RanPark + Jenkins
// hash seeding) and the funcfl EAM reader/spliner so that callers without
// LAMMPS start the device path from the reference's own seeded state.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mdx.h"
#include "../../include/mdx_host.h"

namespace mdx {
void set_error(const std::string &msg);
}

namespace {

// Park-Miller minimal standard with Schrage's factorisation (random_park.cpp:42-49)
class ParkMiller {
 public:
  explicit ParkMiller(int seed) : state_(seed) {}
  double uniform() {
    const int k = state_ / 127773;
    state_ = 16807 * (state_ - k * 127773) - 2836 * k;
    if (state_ < 0) state_ += 2147483647;
    return (1.0 / 2147483647) * state_;
  }
  // Marsaglia polar pair, second value cached (random_park.cpp:55-76)
  double gaussian() {
    if (have_spare_) {
      have_spare_ = false;
      return spare_;
    }
    double v1, v2, rsq;
    do {
      v1 = 2.0 * uniform() - 1.0;
      v2 = 2.0 * uniform() - 1.0;
      rsq = v1 * v1 + v2 * v2;
    } while (!(rsq < 1.0 && rsq != 0.0));
    const double fac = std::sqrt(-2.0 * std::log(rsq) / rsq);
    spare_ = v1 * fac;
    have_spare_ = true;
    return v2 * fac;
  }
  // seed from (ibase, coordinates): one-at-a-time hash over the int's and
  // the 3 doubles' bytes, taken as signed char, then 5 warm-up draws
  // (random_park.cpp:94-130)
  void reseed(int ibase, const double *coord) {
    uint32_t h = 0;
    auto mix = [&h](const void *p, size_t nbytes) {
      const signed char *b = static_cast<const signed char *>(p);
      for (size_t q = 0; q < nbytes; q++) {
        h += static_cast<uint32_t>(static_cast<int>(b[q]));
        h += h << 10;
        h ^= h >> 6;
      }
    };
    mix(&ibase, sizeof(int));
    mix(coord, 3 * sizeof(double));
    h += h << 3;
    h ^= h >> 11;
    h += h << 15;
    state_ = static_cast<int>(h & 0x7ffffff);
    if (!state_) state_ = 1;
    for (int w = 0; w < 5; w++) uniform();
    have_spare_ = false;
  }
  int state() const { return state_; }

 private:
  int state_;
  bool have_spare_ = false;
  double spare_ = 0.0;
};

int fail(const char *msg) {
  mdx::set_error(msg);
  return MDX_ERR_ARG;
}

}  // namespace

extern "C" {

double mdx_host_ranpark_uniform(int *seed) {
  ParkMiller r(*seed);
  const double u = r.uniform();
  *seed = r.state();
  return u;
}

int mdx_host_lattice(int style, double arg, int units, const int lo[3], const int hi[3], const int periodic[3],
                     double *boxlo, double *boxhi, double *x, long cap, long *n) {
  static const double sc[1][3] = {{0, 0, 0}};
  static const double bcc[2][3] = {{0, 0, 0}, {0.5, 0.5, 0.5}};
  static const double fcc[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
  const double(*basis)[3];
  int nbasis;
  switch (style) {
    case MDX_LATTICE_SC: basis = sc; nbasis = 1; break;
    case MDX_LATTICE_BCC: basis = bcc; nbasis = 2; break;
    case MDX_LATTICE_FCC: basis = fcc; nbasis = 4; break;
    default: return fail("unknown lattice style");
  }
  if (!(arg > 0.0) || !lo || !hi || !periodic || !boxlo || !boxhi || !n) return fail("invalid lattice arguments");
  // lj units: the argument is rho*, scale = (nbasis / (V_cell rho*))^(1/3) (lattice.cpp:256-261)
  double scale = arg;
  if (units == MDX_UNITS_LJ) {
    const double cell_volume = 1.0;
    scale = std::pow(nbasis / cell_volume / arg, 1.0 / 3);
  }
  double sublo[3], subhi[3];
  for (int d = 0; d < 3; d++) {
    if (hi[d] <= lo[d]) return fail("empty lattice region");
    boxlo[d] = scale * static_cast<double>(lo[d]);
    boxhi[d] = scale * static_cast<double>(hi[d]);
    sublo[d] = boxlo[d];
    subhi[d] = boxhi[d];
    // periodic insertion box is shrunk so the hi face is not populated twice
    // (create_atoms.cpp:284-313)
    if (periodic[d]) {
      const double eps = (boxhi[d] - boxlo[d]) * 1.0e-6;
      sublo[d] -= eps;
      subhi[d] -= 2.0 * eps;
    }
  }
  long count = 0;
  for (int k = lo[2] - 1; k <= hi[2]; k++)
    for (int j = lo[1] - 1; j <= hi[1]; j++)
      for (int i = lo[0] - 1; i <= hi[0]; i++)
        for (int m = 0; m < nbasis; m++) {
          const double p0 = (i + basis[m][0]) * scale;
          const double p1 = (j + basis[m][1]) * scale;
          const double p2 = (k + basis[m][2]) * scale;
          if (p0 < sublo[0] || p0 >= subhi[0] || p1 < sublo[1] || p1 >= subhi[1] || p2 < sublo[2] ||
              p2 >= subhi[2])
            continue;
          if (x && count < cap) {
            x[3 * count] = p0;
            x[3 * count + 1] = p1;
            x[3 * count + 2] = p2;
          }
          count++;
        }
  *n = count;
  if (x && count > cap) return fail("lattice atom count exceeds the output capacity");
  return MDX_OK;
}

int mdx_host_velocity(long n, const double *x, const int *type, const double *mass, double t_desired, int seed,
                      int gaussian, int loop_geom, int mom, double boltz, double mvv2e, double *v) {
  if (seed <= 0) return fail("Illegal velocity create command");
  if (n <= 0 || !x || !type || !mass || !v) return fail("invalid velocity arguments");
  auto draw3 = [gaussian](ParkMiller &r, double *out) {
    for (int d = 0; d < 3; d++) out[d] = gaussian ? r.gaussian() : r.uniform();
  };
  if (!loop_geom) {
    // loop all: one stream over IDs 1..N; atom with ID k is array entry k-1
    ParkMiller r(seed);
    for (long i = 0; i < n; i++) {
      double u[3];
      draw3(r, u);
      const double s = 1.0 / std::sqrt(mass[type[i]]);
      for (int d = 0; d < 3; d++) v[3 * i + d] = u[d] * s;
    }
  } else {
    ParkMiller r(1);
    for (long i = 0; i < n; i++) {
      r.reseed(seed, &x[3 * i]);
      double u[3];
      draw3(r, u);
      const double s = 1.0 / std::sqrt(mass[type[i]]);
      for (int d = 0; d < 3; d++) v[3 * i + d] = u[d] * s;
    }
  }
  if (mom) {
    // Velocity::zero_momentum: subtract the centre-of-mass velocity
    double mtot = 0.0, p[3] = {0.0, 0.0, 0.0};
    for (long i = 0; i < n; i++) mtot += mass[type[i]];
    for (long i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) p[d] += v[3 * i + d] * mass[type[i]];
    if (mtot > 0.0)
      for (int d = 0; d < 3; d++) p[d] /= mtot;
    for (long i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) v[3 * i + d] -= p[d];
  }
  // Velocity::rescale to t_desired with ComputeTemp (dof = 3N - 3)
  double t = 0.0;
  for (long i = 0; i < n; i++) {
    const double *q = &v[3 * i];
    t += (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) * mass[type[i]];
  }
  const double dof = 3.0 * n - 3.0;
  t *= dof > 0.0 ? mvv2e / (dof * boltz) : 0.0;
  if (t == 0.0) return fail("Attempting to rescale a 0.0 temperature");
  const double factor = std::sqrt(t_desired / t);
  for (long i = 0; i < 3 * n; i++) v[i] *= factor;
  return MDX_OK;
}

}  // extern "C"

namespace {

// 7-coefficient cubic Hermite per knot (pair_eam.cpp:723-751); f is 1-based
void spline7(int n, double delta, const std::vector<double> &f, double *sp) {
  auto S = [sp](int m, int k) -> double & { return sp[m * 7 + k]; };
  for (int m = 1; m <= n; m++) S(m, 6) = f[m];
  S(1, 5) = S(2, 6) - S(1, 6);
  S(2, 5) = 0.5 * (S(3, 6) - S(1, 6));
  S(n - 1, 5) = 0.5 * (S(n, 6) - S(n - 2, 6));
  S(n, 5) = S(n, 6) - S(n - 1, 6);
  for (int m = 3; m <= n - 2; m++)
    S(m, 5) = ((S(m - 2, 6) - S(m + 2, 6)) + 8.0 * (S(m + 1, 6) - S(m - 1, 6))) / 12.0;
  for (int m = 1; m <= n - 1; m++) {
    S(m, 4) = 3.0 * (S(m + 1, 6) - S(m, 6)) - 2.0 * S(m, 5) - S(m + 1, 5);
    S(m, 3) = S(m, 5) + S(m + 1, 5) - 2.0 * (S(m + 1, 6) - S(m, 6));
  }
  S(n, 4) = 0.0;
  S(n, 3) = 0.0;
  for (int m = 1; m <= n; m++) {
    S(m, 2) = S(m, 5) / delta;
    S(m, 1) = 2.0 * S(m, 4) / delta;
    S(m, 0) = 3.0 * S(m, 3) / delta;
  }
}

// 4-point Lagrange interpolation of a 1-based file table at r (pair_eam.cpp:532-553)
double lagrange4(double r, double delta, int nfile, const std::vector<double> &tab) {
  const double sixth = 1.0 / 6.0;
  double p = r / delta + 1.0;
  int k = static_cast<int>(p);
  k = std::min(k, nfile - 2);
  k = std::max(k, 2);
  p -= k;
  p = std::min(p, 2.0);
  const double c1 = -sixth * p * (p - 1.0) * (p - 2.0);
  const double c2 = 0.5 * (p * p - 1.0) * (p - 2.0);
  const double c3 = -0.5 * p * (p + 1.0) * (p - 2.0);
  const double c4 = sixth * p * (p * p - 1.0);
  return c1 * tab[k - 1] + c2 * tab[k] + c3 * tab[k + 1] + c4 * tab[k + 2];
}

}  // namespace

extern "C" int mdx_host_eam_funcfl(const char *path, mdx_eam_info *info, double *frho_spline, double *rhor_spline,
                                   double *z2r_spline) {
  if (!path || !info) return fail("invalid funcfl arguments");
  std::ifstream in(path);
  if (!in) return fail("Cannot open EAM potential file");
  std::string line;
  int fnrho, fnr, tmp;
  double fdrho, fdr, fcut, fmass;
  if (!std::getline(in, line) || !std::getline(in, line)) return fail("funcfl header");
  {
    std::istringstream s(line);
    if (!(s >> tmp >> fmass)) return fail("funcfl line 2");
  }
  if (!std::getline(in, line)) return fail("funcfl header");
  {
    std::istringstream s(line);
    if (!(s >> fnrho >> fdrho >> fnr >> fdr >> fcut)) return fail("funcfl line 3");
  }
  if (fnrho < 5 || fnr < 5 || !(fdrho > 0) || !(fdr > 0)) return fail("funcfl table sizes");
  std::vector<double> F(fnrho + 1), Z(fnr + 1), R(fnr + 1);
  for (int k = 1; k <= fnrho; k++)
    if (!(in >> F[k])) return fail("funcfl F table truncated");
  for (int k = 1; k <= fnr; k++)
    if (!(in >> Z[k])) return fail("funcfl Z table truncated");
  for (int k = 1; k <= fnr; k++)
    if (!(in >> R[k])) return fail("funcfl rho table truncated");
  // file2array for a single file: common grid = the file grid minus one point
  const double rmax = (fnr - 1) * fdr, rhomax = (fnrho - 1) * fdrho;
  const int nr = static_cast<int>(rmax / fdr + 0.5), nrho = static_cast<int>(rhomax / fdrho + 0.5);
  info->nrho = nrho;
  info->nr = nr;
  info->nfrho = 2;  // the file + the zero table of non-EAM types
  info->nrhor = 1;
  info->nz2r = 1;
  info->drho = fdrho;
  info->dr = fdr;
  info->rhomax = rhomax;
  info->cutmax = fcut;
  info->mass = fmass;
  if (!frho_spline || !rhor_spline || !z2r_spline) return MDX_OK;
  std::vector<double> fr(nrho + 1, 0.0), rr(nr + 1, 0.0), zz(nr + 1, 0.0), zero(nrho + 1, 0.0);
  for (int m = 1; m <= nrho; m++) fr[m] = lagrange4((m - 1) * fdrho, fdrho, fnrho, F);
  for (int m = 1; m <= nr; m++) {
    const double r = (m - 1) * fdr;
    rr[m] = lagrange4(r, fdr, fnr, R);
    const double zi = lagrange4(r, fdr, fnr, Z);
    zz[m] = 27.2 * 0.529 * zi * zi;  // Hartree*Bohr; z2r = r*phi (pair_eam.cpp:661)
  }
  std::memset(frho_spline, 0, sizeof(double) * 2 * (nrho + 1) * 7);
  spline7(nrho, fdrho, fr, frho_spline);
  spline7(nrho, fdrho, zero, frho_spline + (size_t)(nrho + 1) * 7);
  std::memset(rhor_spline, 0, sizeof(double) * (nr + 1) * 7);
  std::memset(z2r_spline, 0, sizeof(double) * (nr + 1) * 7);
  spline7(nr, fdr, rr, rhor_spline);
  spline7(nr, fdr, zz, z2r_spline);
  return MDX_OK;
}

namespace {
// PairEAM::grab semantics: n values taken from whole lines
bool grab(std::istream &in, int n, double *out) {
  std::string line;
  int i = 0;
  while (i < n) {
    if (!std::getline(in, line)) return false;
    std::istringstream s(line);
    double v;
    while (i < n && (s >> v)) out[i++] = v;
  }
  return true;
}
}  // namespace

extern "C" int mdx_host_eam_setfl(const char *path, int fs, int ntypes, const char *const *type_elements,
                                  mdx_eam_info *info, double *frho_spline, double *rhor_spline, double *z2r_spline,
                                  int *type2frho, int *type2rhor, int *type2z2r, double *mass) {
  if (!path || !info || ntypes < 1 || ntypes > 8 || !type_elements) return fail("invalid setfl arguments");
  std::ifstream in(path);
  if (!in) return fail("Cannot open EAM potential file");
  std::string line;
  for (int k = 0; k < 3; k++)
    if (!std::getline(in, line)) return fail("setfl header");
  if (!std::getline(in, line)) return fail("setfl header");
  std::vector<std::string> names;
  int nel = 0;
  {
    std::istringstream s(line);
    if (!(s >> nel)) return fail("Incorrect element names in EAM potential file");
    std::string w;
    while (s >> w) names.push_back(w);
    if (nel < 1 || nel > 8 || (int)names.size() != nel) return fail("Incorrect element names in EAM potential file");
  }
  int nrho, nr;
  double drho, dr, cut;
  if (!std::getline(in, line)) return fail("setfl line 5");
  {
    std::istringstream s(line);
    if (!(s >> nrho >> drho >> nr >> dr >> cut)) return fail("setfl line 5");
  }
  if (nrho < 5 || nr < 5 || !(drho > 0) || !(dr > 0) || !(cut > 0)) return fail("setfl table sizes");
  const int per_rho = fs ? nel : 1;
  std::vector<double> el_mass(nel), F((size_t)nel * nrho), R((size_t)nel * per_rho * nr),
      Z((size_t)nel * (nel + 1) / 2 * nr);
  for (int i = 0; i < nel; i++) {
    if (!std::getline(in, line)) return fail("setfl element line");
    std::istringstream s(line);
    int z;
    if (!(s >> z >> el_mass[i])) return fail("setfl element line");
    if (!grab(in, nrho, &F[(size_t)i * nrho])) return fail("setfl table truncated");
    for (int j = 0; j < per_rho; j++)
      if (!grab(in, nr, &R[((size_t)i * per_rho + j) * nr])) return fail("setfl table truncated");
  }
  for (size_t k = 0; k < (size_t)nel * (nel + 1) / 2; k++)
    if (!grab(in, nr, &Z[k * nr])) return fail("setfl table truncated");
  // pair_coeff * * file E1 .. En: every type on an element (pair_eam_alloy.cpp:52-106)
  std::vector<int> map(ntypes + 1, 0);
  for (int t = 1; t <= ntypes; t++) {
    const char *want = type_elements[t - 1];
    int e = -1;
    for (int k = 0; k < nel && want; k++)
      if (names[k] == want) e = k;
    if (e < 0) return fail("No matching element in EAM potential file (NULL types need pair hybrid, not supported)");
    map[t] = e;
  }
  // file2array (pair_eam_alloy.cpp:220-330, pair_eam_fs.cpp): tables as read
  info->nrho = nrho;
  info->nr = nr;
  info->nfrho = nel + 1;
  info->nrhor = nel * per_rho;
  info->nz2r = nel * (nel + 1) / 2;
  info->drho = drho;
  info->dr = dr;
  info->rhomax = (nrho - 1) * drho;
  info->cutmax = cut;
  info->mass = el_mass[map[1]];
  if (!frho_spline || !rhor_spline || !z2r_spline) return MDX_OK;
  std::memset(frho_spline, 0, sizeof(double) * info->nfrho * (nrho + 1) * 7);
  std::memset(rhor_spline, 0, sizeof(double) * info->nrhor * (nr + 1) * 7);
  std::memset(z2r_spline, 0, sizeof(double) * info->nz2r * (nr + 1) * 7);
  std::vector<double> tab;
  for (int i = 0; i <= nel; i++) {  // + the zero table of non-EAM types
    tab.assign(nrho + 1, 0.0);
    if (i < nel)
      for (int m = 1; m <= nrho; m++) tab[m] = F[(size_t)i * nrho + m - 1];
    spline7(nrho, drho, tab, frho_spline + (size_t)i * (nrho + 1) * 7);
  }
  for (int k = 0; k < info->nrhor; k++) {
    tab.assign(nr + 1, 0.0);
    for (int m = 1; m <= nr; m++) tab[m] = R[(size_t)k * nr + m - 1];
    spline7(nr, dr, tab, rhor_spline + (size_t)k * (nr + 1) * 7);
  }
  for (int k = 0; k < info->nz2r; k++) {
    tab.assign(nr + 1, 0.0);
    for (int m = 1; m <= nr; m++) tab[m] = Z[(size_t)k * nr + m - 1];
    spline7(nr, dr, tab, z2r_spline + (size_t)k * (nr + 1) * 7);
  }
  const int nt1 = ntypes + 1;
  for (int i = 0; i < nt1; i++) {
    if (type2frho) type2frho[i] = i ? map[i] : 0;
    if (mass) mass[i] = i ? el_mass[map[i]] : 0.0;
    for (int j = 0; j < nt1; j++) {
      int rh = 0, zz = 0;
      if (i && j) {
        rh = fs ? map[i] * nel + map[j] : map[i];
        int irow = std::max(map[i], map[j]), icol = std::min(map[i], map[j]);
        zz = irow * (irow + 1) / 2 + icol;
      }
      if (type2rhor) type2rhor[i * nt1 + j] = rh;
      if (type2z2r) type2z2r[i * nt1 + j] = zz;
    }
  }
  return MDX_OK;
}

extern "C" int mdx_host_write_synthetic_funcfl(const char *path, int seed, double a, int nrho, int nr, double dr,
                                               double cut, double mass, double *params_out) {
  if (!path || seed <= 0 || !(a > 0) || nrho < 5 || nr < 5 || !(dr > 0) || !(cut > 0)) return fail("invalid arguments");
  ParkMiller r(seed);
  const double Z0 = 10.0 + 2.0 * r.uniform();
  const double beta = 1.4 + 0.2 * r.uniform();
  const double chi = 5.5 + 1.0 * r.uniform();
  const double E0 = 2.5 + 1.0 * r.uniform();
  const double re = a / std::sqrt(2.0), fe = 1.0;
  auto dens = [&](double rr) { return fe * std::exp(-chi * (rr / re - 1.0)); };
  // perfect-fcc density: shells at a/sqrt2 (12), a (6), a*sqrt(3/2) (24), a*sqrt2 (12), ...
  double rho_e = 0.0;
  const int shell_n[] = {12, 6, 24, 12, 24, 8, 48, 6};
  const double shell_r2[] = {0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0};
  for (int s = 0; s < 8; s++) {
    const double rs = a * std::sqrt(shell_r2[s]);
    if (rs < cut) rho_e += shell_n[s] * dens(rs);
  }
  const double drho = 2.0 * rho_e / (nrho - 1);
  FILE *fp = std::fopen(path, "w");
  if (!fp) return fail("cannot write funcfl file");
  std::fprintf(fp, "synthetic funcfl: Z0=%.10g beta=%.10g chi=%.10g E0=%.10g re=%.10g rho_e=%.10g seed=%d\n", Z0,
               beta, chi, E0, re, rho_e, seed);
  std::fprintf(fp, "29 %.6f %.6f FCC\n", mass, a);
  std::fprintf(fp, "%d %.16e %d %.16e %.16e\n", nrho, drho, nr, dr, cut);
  auto emit = [fp](int count, auto fn) {
    for (int k = 0; k < count; k++) std::fprintf(fp, "%.16e%s", fn(k), (k % 5 == 4 || k == count - 1) ? "\n" : " ");
  };
  emit(nrho, [&](int k) { return -E0 * std::sqrt((k * drho) / rho_e); });
  emit(nr, [&](int k) { return Z0 * std::exp(-beta * (k * dr)); });
  emit(nr, [&](int k) { return dens(k * dr); });
  std::fclose(fp);
  if (params_out) {
    params_out[0] = Z0;
    params_out[1] = beta;
    params_out[2] = chi;
    params_out[3] = E0;
    params_out[4] = re;
    params_out[5] = rho_e;
  }
  return MDX_OK;
}
