// Sanitizer driver for the CPU code: the FP64 oracle (oracle/mdref.c,
// oracle/fitref.c) and the host set-up code of the library
// (lammps_amd/csrc/host_setup.cpp), built with
// -fsanitize=address,undefined by tests/test_sanitizers.py and run as one
// process.  Each block drives a path the other tests drive through Python
// (lattice + velocity create, lj/cut and eam runs with every list style,
// funcfl/setfl readers, the synthetic funcfl writer, the batched
// force-matching oracle), so an out-of-bounds access or undefined
// behaviour on those paths ends the process with the sanitizer's report.
// Exit status 0 and the line "host_checks ok" mean every block ran clean.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/mdx.h"
#include "../../include/mdx_host.h"
#include "../../oracle/fitref.h"
#include "../../oracle/mdref.h"

// host_setup.cpp reports failures through the library's error slot
namespace mdx {
static std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
}  // namespace mdx
extern "C" const char *mdx_last_error(void) { return mdx::g_err.c_str(); }

#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
      std::exit(2);                                                        \
    }                                                                      \
  } while (0)

static void oracle_lj(int list_style) {
  mdref *s = mdref_new();
  const int lo[3] = {0, 0, 0}, hi[3] = {5, 5, 5}, per[3] = {1, 1, 1};
  CHECK(mdref_units(s, MDREF_UNITS_LJ) == 0);
  CHECK(mdref_lattice_box(s, 2, 0.8442, lo, hi, per, 2) == 0);
  std::vector<int> t(mdref_natoms(s));
  for (size_t i = 0; i < t.size(); i++) t[i] = 1 + (int)(i % 5 == 0);
  CHECK(mdref_set_type(s, (int)t.size(), t.data()) == 0);
  CHECK(mdref_mass(s, 1, 1.0) == 0 && mdref_mass(s, 2, 2.0) == 0);
  CHECK(mdref_velocity_create(s, 1.44, 87287, 1, 1, 1) == 0);
  CHECK(mdref_pair_lj(s, 2.5, 1) == 0);
  CHECK(mdref_pair_lj_coeff(s, 1, 1, 1.0, 1.0, -1.0) == 0);
  CHECK(mdref_pair_lj_coeff(s, 1, 2, 1.5, 0.8, 2.0) == 0);
  CHECK(mdref_pair_lj_coeff(s, 2, 2, 0.5, 0.88, 2.2) == 0);
  CHECK(mdref_neighbor(s, 0.3, 1, 0, 1, 2000) == 0);
  CHECK(mdref_list_style(s, list_style) == 0);
  CHECK(mdref_sortfreq(s, 7) == 0);
  CHECK(mdref_timestep(s, 0.005) == 0);
  CHECK(mdref_setup(s) == 0);
  double rows[6 * 64];
  int nrows = 0;
  CHECK(mdref_run(s, 30, rows, 64, &nrows) == 0);
  CHECK(mdref_compute(s, 2, 4) == 0);
  const int n = mdref_natoms(s);
  std::vector<double> x(3 * n), v(3 * n), f(3 * n), e(n), w(6 * n);
  std::vector<int> ty(n), im(3 * n);
  CHECK(mdref_get_by_tag(s, x.data(), v.data(), f.data(), ty.data(), im.data()) == 0);
  CHECK(mdref_get_peratom_by_tag(s, e.data(), w.data()) == 0);
  const long np = mdref_list_pairs(s, nullptr, 0);
  std::vector<int> pairs(2 * (size_t)np);
  CHECK(mdref_list_pairs(s, pairs.data(), np) == np);
  mdref_free(s);
}

static void oracle_eam(const char *path, bool cluster) {
  mdref *s = mdref_new();
  CHECK(mdref_units(s, MDREF_UNITS_METAL) == 0);
  if (!cluster) {
    const int lo[3] = {0, 0, 0}, hi[3] = {4, 4, 4}, per[3] = {1, 1, 1};
    CHECK(mdref_lattice_box(s, 2, 3.615, lo, hi, per, 1) == 0);
  } else {
    const double blo[3] = {-12, -12, -12}, bhi[3] = {12, 12, 12};
    const int per[3] = {0, 0, 0};
    CHECK(mdref_box(s, blo, bhi, per, 1) == 0);
    std::vector<double> x;
    for (int k = -2; k < 2; k++)
      for (int j = -2; j < 2; j++)
        for (int i = -2; i < 2; i++) {
          const double b[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
          for (auto &q : b) {
            x.push_back((i + q[0]) * 3.615);
            x.push_back((j + q[1]) * 3.615);
            x.push_back((k + q[2]) * 3.615);
          }
        }
    std::vector<int> t(x.size() / 3, 1);
    CHECK(mdref_add_atoms(s, (int)t.size(), x.data(), t.data()) == 0);
  }
  CHECK(mdref_pair_eam_file(s, path) == 0);
  CHECK(mdref_velocity_create(s, 1600.0, 376847, 0, 1, 1) == 0);
  CHECK(mdref_neighbor(s, 1.0, 1, 5, 1, 2000) == 0);
  CHECK(mdref_timestep(s, 0.005) == 0);
  CHECK(mdref_setup(s) == 0);
  double rows[6 * 64];
  int nrows = 0;
  CHECK(mdref_run(s, 20, rows, 64, &nrows) == 0);
  const int n = mdref_natoms(s);
  std::vector<double> rho(n), fp(n);
  CHECK(mdref_get_eam_by_tag(s, rho.data(), fp.data()) == 0);
  mdref_free(s);
}

static void oracle_fit(const mdx_eam_info &info, const std::vector<double> &fr, const std::vector<double> &rr,
                       const std::vector<double> &zr) {
  // two LJ replicas (periodic) and two eam clusters (non-periodic)
  const int natoms[2] = {108, 60};
  const double lo[6] = {0, 0, 0, -9, -9, -9}, hi[6] = {5.0, 5.0, 5.0, 9, 9, 9};
  const int per[6] = {1, 1, 1, 0, 0, 0};
  std::vector<double> x;
  std::vector<int> t;
  unsigned st = 12345;
  auto u = [&st]() {
    st = st * 1103515245u + 12345u;
    return (st >> 8) / 16777216.0;
  };
  for (int i = 0; i < 108; i++) {
    x.push_back(5.0 * ((i % 6) + 0.3 * u()) / 6.0);
    x.push_back(5.0 * (((i / 6) % 6) + 0.3 * u()) / 6.0);
    x.push_back(5.0 * ((i / 36) + 0.3 * u()) / 3.0);
    t.push_back(1 + (i % 4 == 0));
  }
  for (int i = 0; i < 60; i++) {
    x.push_back(2.5 * (i % 4) - 4.0 + 0.1 * u());
    x.push_back(2.5 * ((i / 4) % 3) - 3.0 + 0.1 * u());
    x.push_back(2.5 * (i / 12) - 5.0 + 0.1 * u());
    t.push_back(1);
  }
  fitref *s = fitref_new(2, natoms, lo, hi, per, 2, t.data(), x.data());
  CHECK(s != nullptr);
  const int nt = 9;
  std::vector<double> eps(nt, 0.0), sig(nt, 0.0), cut(nt, 0.0);
  const double e2[4] = {0, 1.0, 1.5, 0.5}, s2[4] = {0, 1.0, 0.8, 0.88};
  for (int a = 1; a <= 2; a++)
    for (int b = 1; b <= 2; b++) {
      const int k = (a == 1 && b == 1) ? 1 : (a == 2 && b == 2 ? 3 : 2);
      eps[a * 3 + b] = e2[k];
      sig[a * 3 + b] = s2[k];
      cut[a * 3 + b] = 2.5 * s2[k];
    }
  const size_t N = t.size();
  std::vector<double> f(3 * N), gf(3 * N), en(2), geps(nt), gsig(nt);
  CHECK(fitref_lj_forces(s, eps.data(), sig.data(), cut.data(), f.data(), en.data()) == 0);
  for (size_t k = 0; k < gf.size(); k++) gf[k] = 1e-3 * f[k];
  CHECK(fitref_lj_vjp(s, eps.data(), sig.data(), cut.data(), gf.data(), geps.data(), gsig.data()) == 0);
  // knot values from the spline tables' c6 column (knots 1..n)
  std::vector<double> kf(info.nrho), kr(info.nr), kz(info.nr);
  for (int m = 0; m < info.nrho; m++) kf[m] = fr[(size_t)(m + 1) * 7 + 6];
  for (int m = 0; m < info.nr; m++) {
    kr[m] = rr[(size_t)(m + 1) * 7 + 6];
    kz[m] = zr[(size_t)(m + 1) * 7 + 6];
  }
  CHECK(fitref_eam_set(s, info.nrho, info.drho, info.nr, info.dr, info.cutmax, kf.data(), kr.data(), kz.data()) == 0);
  std::vector<double> rho(N), fpv(N), gfr(info.nrho), grr(info.nr), gzr(info.nr);
  CHECK(fitref_eam_forces(s, f.data(), en.data(), rho.data(), fpv.data()) == 0);
  CHECK(fitref_eam_vjp(s, gf.data(), gfr.data(), grr.data(), gzr.data()) == 0);
  std::vector<double> sp((size_t)(info.nr + 1) * 7);
  fitref_spline(info.nr, info.dr, kr.data(), sp.data());
  fitref_free(s);
}

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: host_checks <scratch dir> <eam/alloy setfl file> <eam/fs setfl file>\n");
    return 2;
  }
  const std::string dir = argv[1];
  // host set-up: lattices, velocity create in every mode
  for (int style : {MDX_LATTICE_SC, MDX_LATTICE_BCC, MDX_LATTICE_FCC}) {
    const int lo[3] = {0, 0, 0}, hi[3] = {3, 4, 5}, per[3] = {1, 0, 1};
    double blo[3], bhi[3];
    long n = 0;
    CHECK(mdx_host_lattice(style, 0.8442, MDX_UNITS_LJ, lo, hi, per, blo, bhi, nullptr, 0, &n) != 0 || n >= 0);
    std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n);
    CHECK(mdx_host_lattice(style, 0.8442, MDX_UNITS_LJ, lo, hi, per, blo, bhi, x.data(), n, &n) == 0);
    std::vector<int> t((size_t)n, 1);
    const double mass[2] = {0.0, 1.0};
    for (int g = 0; g < 2; g++)
      for (int geom = 0; geom < 2; geom++)
        CHECK(mdx_host_velocity(n, x.data(), t.data(), mass, 1.44, 87287, g, geom, 1, 1.0, 1.0, v.data()) == 0);
  }
  int seed = 1;
  for (int k = 0; k < 10000; k++) (void)mdx_host_ranpark_uniform(&seed);
  CHECK(seed == 1043618065);
  // eam tables: synthetic funcfl writer and reader, setfl reader (alloy, fs)
  const std::string funcfl = dir + "/synthetic.eam";
  double params[6];
  CHECK(mdx_host_write_synthetic_funcfl(funcfl.c_str(), 4951, 3.615, 500, 500, 0.01, 4.95, 63.55, params) == 0);
  mdx_eam_info info;
  CHECK(mdx_host_eam_funcfl(funcfl.c_str(), &info, nullptr, nullptr, nullptr) == 0);
  std::vector<double> fr((size_t)info.nfrho * (info.nrho + 1) * 7), rr((size_t)info.nrhor * (info.nr + 1) * 7),
      zr((size_t)info.nz2r * (info.nr + 1) * 7);
  CHECK(mdx_host_eam_funcfl(funcfl.c_str(), &info, fr.data(), rr.data(), zr.data()) == 0);
  CHECK(mdx_host_eam_funcfl((dir + "/missing.eam").c_str(), &info, nullptr, nullptr, nullptr) != 0);
  for (int fs = 0; fs < 2; fs++) {
    const char *el[2] = {"B", "A"};
    const char *path = argv[2 + fs];
    mdx_eam_info si;
    CHECK(mdx_host_eam_setfl(path, fs, 2, el, &si, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr) == 0);
    std::vector<double> a((size_t)si.nfrho * (si.nrho + 1) * 7), b((size_t)si.nrhor * (si.nr + 1) * 7),
        c((size_t)si.nz2r * (si.nr + 1) * 7), m(3);
    std::vector<int> t2f(3), t2r(9), t2z(9);
    CHECK(mdx_host_eam_setfl(path, fs, 2, el, &si, a.data(), b.data(), c.data(), t2f.data(), t2r.data(),
                             t2z.data(), m.data()) == 0);
  }
  // oracle: every list style for lj/cut, funcfl eam periodic and cluster
  for (int ls : {MDREF_LIST_HALF_NEWTON, MDREF_LIST_HALF_NONEWTON, MDREF_LIST_FULL}) oracle_lj(ls);
  oracle_eam(funcfl.c_str(), false);
  oracle_eam(funcfl.c_str(), true);
  int st = 0;
  (void)mdref_ranpark_uniform_seq(1, 10000, &st);
  CHECK(st == 1043618065);
  const double coord[3] = {1.0, 2.0, 3.0};
  (void)mdref_ranpark_reset_coord(87287, coord);
  oracle_fit(info, fr, rr, zr);
  std::printf("host_checks ok\n");
  return 0;
}
