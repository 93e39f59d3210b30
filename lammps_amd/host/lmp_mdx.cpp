// lmp_mdx.cpp - LAMMPS-style input-script runner on libmdx.
//
//   lmp_mdx -in in.script [-var name value ...] [-device N]
//
// Reads the subset of the reference's commands that set up and run the
// short-range path (src/input.cpp:532-600 dispatch): variable (index,
// equal), units, atom_style atomic, boundary, lattice, region block,
// create_box, create_atoms, mass, velocity create, pair_style
// lj/cut|eam|eam/alloy|eam/fs, pair_coeff, pair_modify shift, neighbor, neigh_modify,
// fix nve, timestep, thermo, compute pe/atom|stress/atom, dump custom,
// dump_modify, undump, run, plus `mdx` (precision, threads per atom, list
// layout, fusion).  Output follows the reference's log: thermo rows
// "Step Temp E_pair E_mol TotEng Press" (thermo.cpp: %8d, %12.8g), the
// loop time and the neighbour statistics (src/finish.cpp); dumps follow
// src/dump_custom.cpp.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "output.h"
#include "plugin.h"

namespace lmpmdx {

// --------------------------------------------------------------- equal-style arithmetic
class Expr {
 public:
  explicit Expr(const std::string &s) : s_(s) {}
  double eval() {
    const double v = sum();
    skip();
    if (p_ != s_.size()) throw Error("Invalid syntax in variable formula: " + s_);
    return v;
  }

 private:
  std::string s_;
  size_t p_ = 0;
  void skip() {
    while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) p_++;
  }
  double sum() {
    double v = prod();
    for (;;) {
      skip();
      if (p_ < s_.size() && (s_[p_] == '+' || s_[p_] == '-')) {
        const char op = s_[p_++];
        const double r = prod();
        v = op == '+' ? v + r : v - r;
      } else return v;
    }
  }
  double prod() {
    double v = power();
    for (;;) {
      skip();
      if (p_ < s_.size() && (s_[p_] == '*' || s_[p_] == '/')) {
        const char op = s_[p_++];
        const double r = power();
        v = op == '*' ? v * r : v / r;
      } else return v;
    }
  }
  double power() {
    const double b = unary();
    skip();
    if (p_ < s_.size() && s_[p_] == '^') {
      p_++;
      return std::pow(b, power());
    }
    return b;
  }
  double unary() {
    skip();
    if (p_ < s_.size() && s_[p_] == '-') {
      p_++;
      return -unary();
    }
    if (p_ < s_.size() && s_[p_] == '(') {
      p_++;
      const double v = sum();
      skip();
      if (p_ >= s_.size() || s_[p_] != ')') throw Error("Invalid syntax in variable formula: " + s_);
      p_++;
      return v;
    }
    const char *start = s_.c_str() + p_;
    char *end = nullptr;
    const double v = std::strtod(start, &end);
    if (end == start) throw Error("Invalid syntax in variable formula: " + s_);
    p_ += end - start;
    return v;
  }
};

static std::string fmt_g(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.15g", v);
  return buf;
}

// --------------------------------------------------------------- the world
class Lammps {
 public:
  explicit Lammps(int device) : device_(device) {}
  ~Lammps() {
    if (ctx_) mdx_destroy(ctx_);
  }
  void set_index_variable(const std::string &name, const std::string &value) { vars_[name] = value; }
  void file(const std::string &path);

 private:
  int device_;
  mdx_ctx *ctx_ = nullptr;
  std::map<std::string, std::string> vars_;
  Units units_;
  std::string lattice_style_;
  double lattice_arg_ = 0.0;
  int region_lo_[3] = {0, 0, 0}, region_hi_[3] = {0, 0, 0};
  bool region_ = false;
  int periodic_[3] = {1, 1, 1};
  std::string boundstr_ = "pp pp pp";
  double boxlo_[3] = {0, 0, 0}, boxhi_[3] = {0, 0, 0};
  bool box_ = false;
  int ntypes_ = 0;
  std::vector<double> x_, v_;
  std::vector<int> type_;
  std::vector<double> mass_;
  std::unique_ptr<Pair> pair_;
  Neighbor neighbor_;
  bool fix_nve_ = false;
  double dt_ = -1.0;
  int thermo_every_ = 0;
  int precision_ = MDX_MIXED;
  int tpa_ = 0, fused_ = 1;
  bool uploaded_ = false;
  long step_ = 0;
  std::map<std::string, ComputePerAtom> computes_;
  std::vector<std::unique_ptr<DumpCustom>> dumps_;

  std::string substitute(const std::string &line);
  void command(const std::vector<std::string> &w);
  void run(int nsteps);
  void upload();
  void write_dumps(bool setup);
  DumpCustom *find_dump(const std::string &id);
};

// $x and ${name} outside quotes (src/input.cpp:405-500)
std::string Lammps::substitute(const std::string &line) {
  std::string out;
  char quote = 0;
  for (size_t k = 0; k < line.size(); k++) {
    if (line[k] == quote) quote = 0;
    else if (!quote && (line[k] == '"' || line[k] == '\'')) quote = line[k];
    if (quote || line[k] != '$' || k + 1 >= line.size()) {
      out += line[k];
      continue;
    }
    std::string name;
    if (line[k + 1] == '{') {
      const size_t close = line.find('}', k + 2);
      if (close == std::string::npos) throw Error("Invalid variable name in input script");
      name = line.substr(k + 2, close - k - 2);
      k = close;
    } else {
      name = line.substr(k + 1, 1);
      k++;
    }
    auto it = vars_.find(name);
    if (it == vars_.end()) throw Error("Substitution for illegal variable " + name);
    out += it->second;
  }
  return out;
}

void Lammps::file(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw Error("Cannot open input script " + path);
  std::string line;
  while (std::getline(in, line)) {
    // a '#' outside single or double quotes starts a comment (src/input.cpp:318-335)
    char quote = 0;
    for (size_t k = 0; k < line.size(); k++) {
      if (line[k] == quote) quote = 0;
      else if (!quote && (line[k] == '"' || line[k] == '\'')) quote = line[k];
      else if (line[k] == '#' && !quote) {
        line = line.substr(0, k);
        break;
      }
    }
    line = substitute(line);
    // words split on whitespace; quoted text is one word with the quotes
    // removed (src/input.cpp:365-400)
    std::vector<std::string> w;
    std::string t;
    bool in_word = false;
    quote = 0;
    for (char ch : line) {
      if (quote) {
        if (ch == quote) quote = 0;
        else t += ch;
      } else if (ch == '"' || ch == '\'') {
        quote = ch;
        in_word = true;
      } else if (std::isspace((unsigned char)ch)) {
        if (in_word) w.push_back(t);
        t.clear();
        in_word = false;
      } else {
        t += ch;
        in_word = true;
      }
    }
    if (quote) throw Error("Unbalanced quotes in input line");
    if (in_word) w.push_back(t);
    if (w.empty()) continue;
    std::printf("%s\n", line.c_str());
    command(w);
  }
}

void Lammps::command(const std::vector<std::string> &w) {
  const std::string &c = w[0];
  std::vector<std::string> args(w.begin() + 1, w.end());
  if (c == "variable") {
    if (args.size() < 3) throw Error("Illegal variable command");
    if (args[1] == "index") {
      if (!vars_.count(args[0])) vars_[args[0]] = args[2];  // -var on the command line wins
    } else if (args[1] == "equal") {
      std::string e;
      for (size_t k = 2; k < args.size(); k++) e += args[k];
      vars_[args[0]] = fmt_g(Expr(e).eval());
    } else {
      throw Error("Unsupported variable style " + args[1]);
    }
  } else if (c == "units") {
    if (box_) throw Error("Units command after simulation box is defined");
    units_.set(args.at(0));
    neighbor_.skin = units_.skin;
  } else if (c == "atom_style") {
    if (args.at(0) != "atomic") throw Error("this build supports atom_style atomic");
  } else if (c == "boundary") {
    if (args.size() != 3) throw Error("Illegal boundary command");
    boundstr_.clear();
    for (int d = 0; d < 3; d++) {
      periodic_[d] = args[d] == "p" ? 1 : 0;
      // src/domain.cpp:1554-1567: two letters per dimension
      boundstr_ += (args[d].size() == 1 ? args[d] + args[d] : args[d]) + (d < 2 ? " " : "");
    }
  } else if (c == "lattice") {
    if (args.size() != 2) throw Error("Illegal lattice command");
    lattice_style_ = args[0];
    lattice_arg_ = std::strtod(args[1].c_str(), nullptr);
    const int nb = lattice_style_ == "fcc" ? 4 : lattice_style_ == "bcc" ? 2 : 1;
    const double a = units_.style == "lj" ? std::pow(nb / lattice_arg_, 1.0 / 3) : lattice_arg_;
    std::printf("Lattice spacing in x,y,z = %g %g %g\n", a, a, a);
  } else if (c == "region") {
    if (args.size() < 8 || args[1] != "block") throw Error("this build supports region <id> block xlo xhi ylo yhi zlo zhi");
    for (int d = 0; d < 3; d++) {
      region_lo_[d] = std::atoi(args[2 + 2 * d].c_str());
      region_hi_[d] = std::atoi(args[3 + 2 * d].c_str());
    }
    region_ = true;
  } else if (c == "create_box") {
    if (!region_ || lattice_style_.empty()) throw Error("Create_box region and lattice must be defined");
    ntypes_ = std::atoi(args.at(0).c_str());
    mass_.assign(ntypes_ + 1, 0.0);
    const int st = lattice_style_ == "fcc" ? MDX_LATTICE_FCC : lattice_style_ == "bcc" ? MDX_LATTICE_BCC : MDX_LATTICE_SC;
    long n = 0;
    check(mdx_host_lattice(st, lattice_arg_, units_.style == "lj" ? MDX_UNITS_LJ : MDX_UNITS_METAL, region_lo_,
                           region_hi_, periodic_, boxlo_, boxhi_, nullptr, 0, &n));
    box_ = true;
    std::printf("Created orthogonal box = (%g %g %g) to (%g %g %g)\n", boxlo_[0], boxlo_[1], boxlo_[2], boxhi_[0],
                boxhi_[1], boxhi_[2]);
  } else if (c == "create_atoms") {
    if (!box_) throw Error("Create_atoms command before simulation box is defined");
    const int t = std::atoi(args.at(0).c_str());
    const int st = lattice_style_ == "fcc" ? MDX_LATTICE_FCC : lattice_style_ == "bcc" ? MDX_LATTICE_BCC : MDX_LATTICE_SC;
    long n = 0;
    check(mdx_host_lattice(st, lattice_arg_, units_.style == "lj" ? MDX_UNITS_LJ : MDX_UNITS_METAL, region_lo_,
                           region_hi_, periodic_, boxlo_, boxhi_, nullptr, 0, &n));
    x_.assign(3 * n, 0.0);
    check(mdx_host_lattice(st, lattice_arg_, units_.style == "lj" ? MDX_UNITS_LJ : MDX_UNITS_METAL, region_lo_,
                           region_hi_, periodic_, boxlo_, boxhi_, x_.data(), n, &n));
    type_.assign(n, t);
    v_.assign(3 * n, 0.0);
    std::printf("Created %ld atoms\n", n);
  } else if (c == "mass") {
    const int t = std::atoi(args.at(0).c_str());
    if (t < 1 || t > ntypes_) throw Error("Invalid type for mass set");
    mass_[t] = std::strtod(args.at(1).c_str(), nullptr);
  } else if (c == "velocity") {
    if (args.size() < 4 || args[1] != "create") throw Error("this build supports velocity <group> create T seed ...");
    const double t = std::strtod(args[2].c_str(), nullptr);
    const int seed = std::atoi(args[3].c_str());
    int gaussian = 0, geom = 0, mom = 1;
    for (size_t k = 4; k + 1 < args.size(); k += 2) {
      if (args[k] == "dist") gaussian = args[k + 1] == "gaussian";
      else if (args[k] == "loop") geom = args[k + 1] == "geom";
      else if (args[k] == "mom") mom = args[k + 1] == "yes";
      else throw Error("Illegal velocity command: " + args[k]);
    }
    if (pair_ && pair_->sets_mass())
      for (int k = 1; k <= ntypes_; k++) mass_[k] = pair_->mass_of(k);
    check(mdx_host_velocity((long)type_.size(), x_.data(), type_.data(), mass_.data(), t, seed, gaussian, geom, mom,
                            units_.boltz, units_.mvv2e, v_.data()));
  } else if (c == "pair_style") {
    pair_.reset(create_pair(args.at(0)));
    pair_->settings(std::vector<std::string>(args.begin() + 1, args.end()));
    uploaded_ = false;
  } else if (c == "pair_coeff") {
    if (!pair_) throw Error("Pair_coeff command before pair_style is defined");
    if (!box_) throw Error("Pair_coeff command before simulation box is defined");
    pair_->ensure_allocated(ntypes_);
    pair_->coeff(args);
    if (pair_->sets_mass())
      for (int k = 1; k <= ntypes_; k++) mass_[k] = pair_->mass_of(k);
    uploaded_ = false;
  } else if (c == "pair_modify") {
    auto *lj = dynamic_cast<PairLJCut *>(pair_.get());
    if (args.size() == 2 && args[0] == "shift" && lj) lj->offset_flag = args[1] == "yes";
    else throw Error("this build supports pair_modify shift yes|no for lj/cut");
    uploaded_ = false;
  } else if (c == "neighbor") {
    if (args.size() != 2 || args[1] != "bin") throw Error("this build supports neighbor <skin> bin");
    neighbor_.skin = std::strtod(args[0].c_str(), nullptr);
    uploaded_ = false;
  } else if (c == "neigh_modify") {
    neighbor_.modify(args);
    uploaded_ = false;
  } else if (c == "fix") {
    if (args.size() < 3 || args[2] != "nve") throw Error("this build supports fix <id> all nve");
    fix_nve_ = true;
  } else if (c == "timestep") {
    dt_ = std::strtod(args.at(0).c_str(), nullptr);
  } else if (c == "thermo") {
    thermo_every_ = std::atoi(args.at(0).c_str());
  } else if (c == "compute") {
    if (args.empty()) throw Error("Illegal compute command");
    if (computes_.count(args[0])) throw Error("Reuse of compute ID");
    computes_[args[0]] = ComputePerAtom::create(args);
  } else if (c == "dump") {
    if (args.empty()) throw Error("Illegal dump command");
    if (find_dump(args[0])) throw Error("Reuse of dump ID");
    dumps_.emplace_back(new DumpCustom(args, computes_));
  } else if (c == "dump_modify") {
    if (args.size() < 2) throw Error("Illegal dump_modify command");
    DumpCustom *d = find_dump(args[0]);
    if (!d) throw Error("Could not find dump_modify ID");
    d->modify(std::vector<std::string>(args.begin() + 1, args.end()));
  } else if (c == "undump") {
    if (args.size() != 1) throw Error("Illegal undump command");
    size_t k = 0;
    while (k < dumps_.size() && dumps_[k]->id != args[0]) k++;
    if (k == dumps_.size()) throw Error("Could not find undump ID");
    dumps_.erase(dumps_.begin() + k);
  } else if (c == "mdx") {
    // mdx precision mixed|double ; mdx tpa N ; mdx fused yes|no
    for (size_t k = 0; k + 1 < args.size(); k += 2) {
      if (args[k] == "precision") precision_ = args[k + 1] == "double" ? MDX_FP64 : MDX_MIXED;
      else if (args[k] == "tpa") tpa_ = std::atoi(args[k + 1].c_str());
      else if (args[k] == "fused") fused_ = args[k + 1] == "yes";
      else throw Error("Illegal mdx command: " + args[k]);
    }
    uploaded_ = false;
  } else if (c == "run") {
    run(std::atoi(args.at(0).c_str()));
  } else {
    throw Error("Unknown command: " + c);
  }
}

void Lammps::upload() {
  if (!pair_) throw Error("Run command before pair_style is defined");
  if (!fix_nve_) throw Error("this build integrates with fix nve: add fix <id> all nve");
  for (int t = 1; t <= ntypes_; t++)
    if (!(mass_[t] > 0.0)) throw Error("Not all per-type masses are set");
  if (!ctx_) check(mdx_create(device_, &ctx_));
  check(mdx_set_box(ctx_, boxlo_, boxhi_, periodic_));
  std::vector<int> tags(type_.size());
  for (size_t k = 0; k < tags.size(); k++) tags[k] = (int)k + 1;
  check(mdx_set_atoms(ctx_, (int)type_.size(), ntypes_, x_.data(), v_.data(), type_.data(), tags.data(),
                      mass_.data()));
  pair_->init(ntypes_);
  pair_->upload(ctx_);
  check(mdx_set_precision(ctx_, precision_));
  check(mdx_set_launch(ctx_, 0, tpa_));
  check(mdx_set_fused(ctx_, fused_));
  check(mdx_neigh_modify(ctx_, neighbor_.skin, neighbor_.every, neighbor_.delay, neighbor_.check, neighbor_.one));
  uploaded_ = true;
}

struct ThermoState {
  const Units *u;
  double natoms, vol;
  long first, last;  // the run's first and last step
  int every;
};

static void thermo_row(void *user, long step, const double *raw) {
  const ThermoState *s = static_cast<const ThermoState *>(user);
  // a run split at dump steps reports the end of every piece: print only
  // the rows of Output (first, last and multiples of thermo N)
  if (!(step == s->first || step == s->last || (s->every > 0 && step % s->every == 0))) return;
  const double dof = 3.0 * s->natoms - 3.0;
  const double t = raw[0] * s->u->mvv2e / (dof * s->u->boltz);
  const bool norm = s->u->style == "lj";
  const double epair = norm ? raw[1] / s->natoms : raw[1];
  double ke = t * 0.5 * dof * s->u->boltz;
  if (norm) ke /= s->natoms;
  const double press = (dof * s->u->boltz * t + raw[2] + raw[3] + raw[4]) / 3.0 / s->vol * s->u->nktv2p;
  std::printf("%8ld %12.8g %12.8g %12.8g %12.8g %12.8g \n", step, t, epair, 0.0, epair + ke, press);
}

DumpCustom *Lammps::find_dump(const std::string &id) {
  for (auto &d : dumps_)
    if (d->id == id) return d.get();
  return nullptr;
}

// Output::setup / Output::write (src/output.cpp:171-214,291-320): a dump is
// written on steps that are multiples of its N and not yet written, and at
// setup also when `dump_modify first yes` asks for it.
void Lammps::write_dumps(bool setup) {
  std::vector<DumpCustom *> due;
  bool peratom = false;
  for (auto &d : dumps_) {
    const bool w = (step_ % d->every == 0 && d->last != step_) || (setup && d->last < 0 && d->first_flag);
    if (w) {
      due.push_back(d.get());
      peratom = peratom || d->needs_peratom();
    }
  }
  if (due.empty()) return;
  const int n = (int)type_.size();
  std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), f(3 * (size_t)n), eatom, vatom;
  std::vector<int> image(3 * (size_t)n);
  check(mdx_get_atoms(ctx_, x.data(), v.data(), f.data(), image.data()));
  if (peratom) {
    // the per-atom energy and virial of this step's positions (eflag_atom,
    // vflag_atom of Integrate::ev_set, src/integrate.cpp:99-112)
    eatom.resize(n);
    vatom.resize(6 * (size_t)n);
    pair_->compute(ctx_, 2, 4, nullptr, eatom.data(), vatom.data());
  }
  Snapshot snap;
  snap.step = step_;
  snap.n = n;
  snap.type = type_.data();
  snap.mass = mass_.data();
  snap.x = x.data();
  snap.v = v.data();
  snap.f = f.data();
  snap.image = image.data();
  snap.eatom = peratom ? eatom.data() : nullptr;
  snap.vatom = peratom ? vatom.data() : nullptr;
  for (int d = 0; d < 3; d++) {
    snap.boxlo[d] = boxlo_[d];
    snap.boxhi[d] = boxhi_[d];
  }
  snap.boundstr = boundstr_;
  snap.mvv2e = units_.mvv2e;
  snap.nktv2p = units_.nktv2p;
  for (DumpCustom *d : due) d->write(snap);
}

void Lammps::run(int nsteps) {
  if (!uploaded_) upload();
  const double dt = dt_ > 0 ? dt_ : units_.dt;
  ThermoState ts{&units_, (double)type_.size(),
                 (boxhi_[0] - boxlo_[0]) * (boxhi_[1] - boxlo_[1]) * (boxhi_[2] - boxlo_[2]), step_,
                 step_ + nsteps, thermo_every_};
  std::printf("Step Temp E_pair E_mol TotEng Press \n");
  // Verlet::setup for every run command, then the timed loop
  check(mdx_setup(ctx_, thermo_row, &ts));
  write_dumps(true);
  const auto t0 = std::chrono::steady_clock::now();
  // the device runs uninterrupted between dump steps; each piece ends with
  // x, v, f of its last step in place (mdx_run's final step is unfused)
  const long end = step_ + nsteps;
  while (step_ < end) {
    long stop = end;
    for (auto &d : dumps_) stop = std::min(stop, (step_ / d->every + 1) * d->every);
    check(mdx_run(ctx_, (int)(stop - step_), dt, units_.ftm2v, thermo_every_, thermo_row, &ts));
    step_ = stop;
    write_dumps(false);
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // host copy of the state, so a later re-upload continues from here
  if (!type_.empty()) check(mdx_get_atoms(ctx_, x_.data(), v_.data(), nullptr, nullptr));
  mdx_stats st;
  check(mdx_get_stats(ctx_, &st));
  std::printf("Loop time of %g on 1 procs for %d steps with %d atoms\n\n", secs, nsteps, (int)type_.size());
  const long half = st.total_neighbors / 2;
  std::printf("Total # of neighbors = %ld\n", half);
  std::printf("Ave neighs/atom = %g\n", type_.empty() ? 0.0 : (double)half / type_.size());
  std::printf("Neighbor list builds = %ld\n", st.ncalls);
  std::printf("Dangerous builds = %ld\n", st.ndanger);
  std::fflush(stdout);
}

}  // namespace lmpmdx

int main(int argc, char **argv) {
  std::string input;
  int device = 0;
  std::vector<std::pair<std::string, std::string>> vars;
  for (int k = 1; k < argc; k++) {
    if (!std::strcmp(argv[k], "-in") && k + 1 < argc) input = argv[++k];
    else if (!std::strcmp(argv[k], "-device") && k + 1 < argc) device = std::atoi(argv[++k]);
    else if (!std::strcmp(argv[k], "-var") && k + 2 < argc) {
      vars.emplace_back(argv[k + 1], argv[k + 2]);
      k += 2;
    } else {
      std::fprintf(stderr, "usage: lmp_mdx -in script [-var name value] [-device N]\n");
      return 2;
    }
  }
  if (input.empty()) {
    std::fprintf(stderr, "usage: lmp_mdx -in script [-var name value] [-device N]\n");
    return 2;
  }
  try {
    lmpmdx::Lammps lmp(device);
    for (auto &v : vars) lmp.set_index_variable(v.first, v.second);
    lmp.file(input);
  } catch (const lmpmdx::Error &e) {
    std::fprintf(stdout, "ERROR: %s\n", e.what());
    return 1;
  }
  return 0;
}
