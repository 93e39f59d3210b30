// plugin.h - the reference's Pair / Neighbor / Fix plugin surface on top of
// libmdx (include/mdx.h).  Class and method names, argument meanings and
// error behaviour follow LAMMPS: src/pair.h:21-222 (settings, coeff,
// init_style, init_one, compute(eflag, vflag), eng_vdwl, virial),
// src/pair_lj_cut.cpp, src/MANYBODY/pair_eam.cpp, src/neighbor.cpp
// (skin, every, delay, check, one) and src/fix_nve.cpp.  Errors are thrown
// as lmpmdx::Error carrying the reference's message text ("All pair coeffs
// are not set", "Neighbor list overflow, boost neigh_modify one", ...).
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mdx.h"
#include "../../include/mdx_host.h"

namespace lmpmdx {

struct Error : std::runtime_error {
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

// update.cpp:118-180
struct Units {
  std::string style = "lj";
  double boltz = 1.0, mvv2e = 1.0, ftm2v = 1.0, nktv2p = 1.0, dt = 0.005, skin = 0.3;
  void set(const std::string &s);
};

void check(int rc);  // throws Error(mdx_last_error()) on a non-zero status

// Pair plugin (src/pair.h)
class Pair {
 public:
  virtual ~Pair() = default;
  virtual void settings(const std::vector<std::string> &args) = 0;
  virtual void coeff(const std::vector<std::string> &args) = 0;
  virtual void init_style() {}
  virtual double init_one(int i, int j) = 0;
  // Pair::init (src/pair.cpp:173-234): init_style, init_one per pair, cutsq
  void init(int ntypes);
  // per-type storage for `ntypes` types (Pair::allocate on the first pair_coeff)
  void ensure_allocated(int n) {
    if (ntypes != n) allocate(n);
  }
  // hands the tables to the device context
  virtual void upload(mdx_ctx *ctx) = 0;
  // Pair::compute(eflag, vflag) on the context's positions; fills f [n][3]
  void compute(mdx_ctx *ctx, int eflag, int vflag, double *f, double *eatom = nullptr, double *vatom = nullptr);
  // pair_coeff * * file.eam sets the type masses (pair_eam.cpp:384)
  virtual bool sets_mass() const { return false; }
  virtual double mass_of(int /*type*/) const { return 0.0; }

  double eng_vdwl = 0.0, virial[6] = {0, 0, 0, 0, 0, 0};
  double cutforce = 0.0;
  int ntypes = 0;
  std::vector<double> cutsq;  // (ntypes+1)^2
 protected:
  virtual void allocate(int n) = 0;
};

class PairLJCut : public Pair {
 public:
  void settings(const std::vector<std::string> &args) override;
  void coeff(const std::vector<std::string> &args) override;
  double init_one(int i, int j) override;
  void upload(mdx_ctx *ctx) override;
  double cut_global = 0.0;
  int offset_flag = 0;  // pair_modify shift yes
 protected:
  void allocate(int n) override;
  std::vector<double> epsilon, sigma, cut, lj1, lj2, lj3, lj4, offset;
  std::vector<int> setflag;
  int nt1 = 0;
  double &at(std::vector<double> &v, int i, int j) { return v[i * nt1 + j]; }
};

class PairEAM : public Pair {
 public:
  void settings(const std::vector<std::string> &args) override;
  void coeff(const std::vector<std::string> &args) override;
  void init_style() override;
  double init_one(int i, int j) override;
  void upload(mdx_ctx *ctx) override;
  bool sets_mass() const override { return true; }
  double mass_of(int) const override { return info.mass; }
 protected:
  void allocate(int n) override { ntypes = n; }
  std::string file;
  mdx_eam_info info{};
  std::vector<double> frho, rhor, z2r;
};

// pair_style eam/alloy | eam/fs: one DYNAMO setfl file, pair_coeff * * file
// E1 .. En with one element name per atom type (pair_eam_alloy.cpp:52-106,
// pair_eam_fs.cpp:52-106); each type takes its element's mass
class PairEAMSetfl : public Pair {
 public:
  explicit PairEAMSetfl(bool fs) : fs_(fs) {}
  void settings(const std::vector<std::string> &args) override;
  void coeff(const std::vector<std::string> &args) override;
  void init_style() override;
  double init_one(int i, int j) override;
  void upload(mdx_ctx *ctx) override;
  bool sets_mass() const override { return true; }
  double mass_of(int type) const override { return type < (int)mass.size() ? mass[type] : 0.0; }
 protected:
  void allocate(int n) override { ntypes = n; }
  bool fs_;
  std::string file;
  mdx_eam_info info{};
  std::vector<double> frho, rhor, z2r, mass;
  std::vector<int> t2f, t2r, t2z;
};

Pair *create_pair(const std::string &style);  // lj/cut | eam | eam/alloy | eam/fs (force.cpp:128-147 style map)

// `neighbor <skin> bin` + `neigh_modify` (src/neighbor.cpp:1731-1862)
struct Neighbor {
  double skin = 0.3;
  int every = 1, delay = 10, check = 1, one = 2000;
  void modify(const std::vector<std::string> &args);
};

// `fix ID group nve` (src/fix_nve.cpp): integration runs inside mdx_run
struct FixNVE {
  std::string id;
};

}  // namespace lmpmdx
