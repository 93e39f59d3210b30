// plugin.cpp - Pair / Neighbor plugin classes over libmdx (see plugin.h).
#include "plugin.h"

#include <cmath>
#include <cstdlib>

namespace lmpmdx {

void check(int rc) {
  if (rc != MDX_OK) throw Error(mdx_last_error());
}

void Units::set(const std::string &s) {
  style = s;
  if (s == "lj") {
    boltz = mvv2e = ftm2v = nktv2p = 1.0;
    dt = 0.005;
    skin = 0.3;
  } else if (s == "metal") {
    boltz = 8.617343e-5;
    mvv2e = 1.0364269e-4;
    ftm2v = 1.0 / 1.0364269e-4;
    nktv2p = 1.6021765e6;
    dt = 0.001;
    skin = 2.0;
  } else {
    throw Error("Illegal units command (lj or metal supported)");
  }
}

static double numeric(const std::string &s) {
  char *end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end) throw Error("Expected floating point parameter in input script: " + s);
  return v;
}

static int inumeric(const std::string &s) {
  char *end = nullptr;
  const long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str() || *end) throw Error("Expected integer parameter in input script: " + s);
  return (int)v;
}

// force->bounds: "*", "n", "n*", "*m", "n*m" -> [lo, hi] (src/force.cpp)
static void bounds(const std::string &s, int nmax, int &lo, int &hi) {
  const size_t star = s.find('*');
  if (star == std::string::npos) {
    lo = hi = inumeric(s);
  } else if (s.size() == 1) {
    lo = 1;
    hi = nmax;
  } else if (star == 0) {
    lo = 1;
    hi = inumeric(s.substr(1));
  } else if (star == s.size() - 1) {
    lo = inumeric(s.substr(0, star));
    hi = nmax;
  } else {
    lo = inumeric(s.substr(0, star));
    hi = inumeric(s.substr(star + 1));
  }
  if (lo < 1 || hi > nmax || lo > hi) throw Error("Numeric index is out of bounds");
}

void Pair::init(int n) {
  if (ntypes != n) allocate(n);
  init_style();
  const int nt1 = n + 1;
  cutsq.assign(nt1 * nt1, 0.0);
  cutforce = 0.0;
  for (int i = 1; i <= n; i++)
    for (int j = i; j <= n; j++) {
      const double c = init_one(i, j);
      cutsq[i * nt1 + j] = cutsq[j * nt1 + i] = c * c;
      cutforce = std::max(cutforce, c);
    }
}

void Pair::compute(mdx_ctx *ctx, int eflag, int vflag, double *f, double *eatom, double *vatom) {
  check(mdx_compute(ctx, eflag, vflag, f, &eng_vdwl, virial, eatom, vatom));
}

// ---------------------------------------------------------------- lj/cut
void PairLJCut::allocate(int n) {
  ntypes = n;
  nt1 = n + 1;
  for (auto *v : {&epsilon, &sigma, &cut, &lj1, &lj2, &lj3, &lj4, &offset}) v->assign(nt1 * nt1, 0.0);
  setflag.assign(nt1 * nt1, 0);
}

// pair_lj_cut.cpp:428-442
void PairLJCut::settings(const std::vector<std::string> &args) {
  if (args.size() != 1) throw Error("Illegal pair_style command");
  cut_global = numeric(args[0]);
}

// pair_lj_cut.cpp:448-476
void PairLJCut::coeff(const std::vector<std::string> &args) {
  if (args.size() < 4 || args.size() > 5) throw Error("Incorrect args for pair coefficients");
  if (!ntypes) throw Error("Pair coeff for a style without atom types");
  int ilo, ihi, jlo, jhi;
  bounds(args[0], ntypes, ilo, ihi);
  bounds(args[1], ntypes, jlo, jhi);
  const double eps = numeric(args[2]), sig = numeric(args[3]);
  const double c = args.size() == 5 ? numeric(args[4]) : cut_global;
  int count = 0;
  for (int i = ilo; i <= ihi; i++)
    for (int j = std::max(jlo, i); j <= jhi; j++) {
      at(epsilon, i, j) = eps;
      at(sigma, i, j) = sig;
      at(cut, i, j) = c;
      setflag[i * nt1 + j] = 1;
      count++;
    }
  if (count == 0) throw Error("Incorrect args for pair coefficients");
}

// pair_lj_cut.cpp:545-602 with geometric mixing (pair.cpp:639-664)
double PairLJCut::init_one(int i, int j) {
  if (!setflag[i * nt1 + i] || !setflag[j * nt1 + j]) throw Error("All pair coeffs are not set");
  if (setflag[i * nt1 + j] == 0) {
    at(epsilon, i, j) = std::sqrt(at(epsilon, i, i) * at(epsilon, j, j));
    at(sigma, i, j) = std::sqrt(at(sigma, i, i) * at(sigma, j, j));
    at(cut, i, j) = std::sqrt(at(cut, i, i) * at(cut, j, j));
  }
  const double e = at(epsilon, i, j), s = at(sigma, i, j), c = at(cut, i, j);
  at(lj1, i, j) = 48.0 * e * std::pow(s, 12.0);
  at(lj2, i, j) = 24.0 * e * std::pow(s, 6.0);
  at(lj3, i, j) = 4.0 * e * std::pow(s, 12.0);
  at(lj4, i, j) = 4.0 * e * std::pow(s, 6.0);
  if (offset_flag) {
    const double ratio = s / c;
    at(offset, i, j) = 4.0 * e * (std::pow(ratio, 12.0) - std::pow(ratio, 6.0));
  } else {
    at(offset, i, j) = 0.0;
  }
  for (auto *v : {&lj1, &lj2, &lj3, &lj4, &offset}) at(*v, j, i) = at(*v, i, j);
  return c;
}

void PairLJCut::upload(mdx_ctx *ctx) {
  check(mdx_pair_lj_init(ctx, ntypes, cutsq.data(), lj1.data(), lj2.data(), lj3.data(), lj4.data(), offset.data()));
}

// ---------------------------------------------------------------- eam
void PairEAM::settings(const std::vector<std::string> &args) {
  if (!args.empty()) throw Error("Illegal pair_style command");
}

// pair_coeff * * file (pair_eam.cpp:346-391); one funcfl file for all types
void PairEAM::coeff(const std::vector<std::string> &args) {
  if (args.size() != 3) throw Error("Incorrect args for pair coefficients");
  if (args[0] != "*" || args[1] != "*")
    throw Error("this build reads one funcfl file for all types: use pair_coeff * * file");
  file = args[2];
  check(mdx_host_eam_funcfl(file.c_str(), &info, nullptr, nullptr, nullptr));
  frho.assign((size_t)info.nfrho * (info.nrho + 1) * 7, 0.0);
  rhor.assign((size_t)info.nrhor * (info.nr + 1) * 7, 0.0);
  z2r.assign((size_t)info.nz2r * (info.nr + 1) * 7, 0.0);
  check(mdx_host_eam_funcfl(file.c_str(), &info, frho.data(), rhor.data(), z2r.data()));
}

void PairEAM::init_style() {
  if (file.empty()) throw Error("All pair coeffs are not set");
}

// pair_eam.cpp:411-427: one global cutoff
double PairEAM::init_one(int, int) { return info.cutmax; }

void PairEAM::upload(mdx_ctx *ctx) {
  const int nt1 = ntypes + 1;
  std::vector<int> t2f(nt1, 0), t2r(nt1 * nt1, 0), t2z(nt1 * nt1, 0);
  check(mdx_pair_eam_init(ctx, ntypes, info.nrho, info.drho, info.nr, info.dr, info.rhomax, info.cutmax * info.cutmax,
                          info.nfrho, frho.data(), info.nrhor, rhor.data(), info.nz2r, z2r.data(), t2f.data(),
                          t2r.data(), t2z.data()));
}

// ---------------------------------------------------------------- eam/alloy, eam/fs
void PairEAMSetfl::settings(const std::vector<std::string> &args) {
  if (!args.empty()) throw Error("Illegal pair_style command");
}

void PairEAMSetfl::coeff(const std::vector<std::string> &args) {
  if (!ntypes) throw Error("Pair coeff for a style without atom types");
  if ((int)args.size() != 3 + ntypes) throw Error("Incorrect args for pair coefficients");
  if (args[0] != "*" || args[1] != "*") throw Error("Incorrect args for pair coefficients");
  file = args[2];
  std::vector<const char *> names;
  for (int t = 0; t < ntypes; t++) names.push_back(args[3 + t].c_str());
  check(mdx_host_eam_setfl(file.c_str(), fs_ ? 1 : 0, ntypes, names.data(), &info, nullptr, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr));
  const int nt1 = ntypes + 1;
  frho.assign((size_t)info.nfrho * (info.nrho + 1) * 7, 0.0);
  rhor.assign((size_t)info.nrhor * (info.nr + 1) * 7, 0.0);
  z2r.assign((size_t)info.nz2r * (info.nr + 1) * 7, 0.0);
  t2f.assign(nt1, 0);
  t2r.assign(nt1 * nt1, 0);
  t2z.assign(nt1 * nt1, 0);
  mass.assign(nt1, 0.0);
  check(mdx_host_eam_setfl(file.c_str(), fs_ ? 1 : 0, ntypes, names.data(), &info, frho.data(), rhor.data(),
                           z2r.data(), t2f.data(), t2r.data(), t2z.data(), mass.data()));
}

void PairEAMSetfl::init_style() {
  if (file.empty()) throw Error("All pair coeffs are not set");
}

double PairEAMSetfl::init_one(int, int) { return info.cutmax; }

void PairEAMSetfl::upload(mdx_ctx *ctx) {
  check(mdx_pair_eam_init(ctx, ntypes, info.nrho, info.drho, info.nr, info.dr, info.rhomax, info.cutmax * info.cutmax,
                          info.nfrho, frho.data(), info.nrhor, rhor.data(), info.nz2r, z2r.data(), t2f.data(),
                          t2r.data(), t2z.data()));
}

Pair *create_pair(const std::string &style) {
  if (style == "lj/cut" || style == "lj/cut/hip") return new PairLJCut();
  if (style == "eam" || style == "eam/hip") return new PairEAM();
  if (style == "eam/alloy") return new PairEAMSetfl(false);
  if (style == "eam/fs") return new PairEAMSetfl(true);
  throw Error("Invalid pair style: " + style);
}

// ---------------------------------------------------------------- neighbor
void Neighbor::modify(const std::vector<std::string> &args) {
  for (size_t k = 0; k < args.size(); k++) {
    auto value = [&](void) -> const std::string & {
      if (k + 1 >= args.size()) throw Error("Illegal neigh_modify command");
      return args[++k];
    };
    if (args[k] == "every") every = inumeric(value());
    else if (args[k] == "delay") delay = inumeric(value());
    else if (args[k] == "check") {
      const std::string &v = value();
      if (v == "yes") check = 1;
      else if (v == "no") check = 0;
      else throw Error("Illegal neigh_modify command");
    } else if (args[k] == "one") one = inumeric(value());
    else if (args[k] == "page") value();  // paged storage is not used on the device
    else throw Error("Illegal neigh_modify command: " + args[k]);
  }
  if (every <= 0 || delay < 0 || one <= 0) throw Error("Illegal neigh_modify command");
}

}  // namespace lmpmdx
