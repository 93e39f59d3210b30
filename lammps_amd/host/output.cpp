// output.cpp - per-atom computes and the custom dump of the script runner
// (src/compute_pe_atom.cpp, src/compute_stress_atom.cpp, src/dump.cpp,
// src/dump_custom.cpp).
#include "output.h"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <sstream>

namespace lmpmdx {

ComputePerAtom ComputePerAtom::create(const std::vector<std::string> &args) {
  // args: ID group style [extra]
  if (args.size() < 3) throw Error("Illegal compute command");
  if (args[1] != "all") throw Error("this build supports per-atom computes on group all");
  ComputePerAtom c;
  if (args[2] == "pe/atom") {
    // src/compute_pe_atom.cpp:40-70: no keywords = every contribution
    c.style = PE;
    if (args.size() > 3) {
      c.pair = false;
      for (size_t k = 3; k < args.size(); k++) {
        const std::string &a = args[k];
        if (a == "pair") c.pair = true;
        else if (a != "bond" && a != "angle" && a != "dihedral" && a != "improper" && a != "kspace")
          throw Error("Illegal compute pe/atom command");
      }
    }
    return c;
  }
  if (args[2] == "stress/atom") {
    // src/compute_stress_atom.cpp:39-110
    c.style = STRESS;
    if (args.size() < 4) throw Error("Illegal compute stress/atom command");
    if (args[3] != "NULL") throw Error("Could not find compute stress/atom temperature ID");
    if (args.size() == 4) {
      c.ke = c.pair = true;
    } else {
      c.ke = c.pair = false;
      for (size_t k = 4; k < args.size(); k++) {
        const std::string &a = args[k];
        if (a == "ke") c.ke = true;
        else if (a == "pair" || a == "virial") c.pair = true;
        else if (a != "bond" && a != "angle" && a != "dihedral" && a != "improper" && a != "kspace" && a != "fix")
          throw Error("Illegal compute stress/atom command");
      }
    }
    return c;
  }
  throw Error("this build supports compute pe/atom and stress/atom, not " + args[2]);
}

// ------------------------------------------------------------------ dump custom

DumpCustom::DumpCustom(const std::vector<std::string> &args, const std::map<std::string, ComputePerAtom> &computes) {
  // args: ID group custom N file field ...
  if (args.size() < 5) throw Error("Illegal dump command");
  if (args[2] != "custom") throw Error("this build supports dump style custom");
  if (args.size() == 5) throw Error("No dump custom arguments specified");
  if (args[1] != "all") throw Error("this build supports dumps of group all");
  id = args[0];
  every = std::atoi(args[3].c_str());
  if (every <= 0) throw Error("Illegal dump command");
  file_ = args[4];
  multifile_ = file_.find('*') != std::string::npos;
  computes_.reserve(args.size());  // Field::compute points into this vector
  struct Name {
    const char *s;
    Kind k;
    VType t;
  };
  static const Name names[] = {
      {"id", ID, INT},       {"proc", PROC, INT},   {"procp1", PROCP1, INT}, {"type", TYPE, INT},
      {"element", ELEMENT, STRING}, {"mass", MASS, DOUBLE},
      {"x", X, DOUBLE},      {"y", Y, DOUBLE},      {"z", Z, DOUBLE},        {"xs", XS, DOUBLE},
      {"ys", YS, DOUBLE},    {"zs", ZS, DOUBLE},    {"xu", XU, DOUBLE},      {"yu", YU, DOUBLE},
      {"zu", ZU, DOUBLE},    {"xsu", XSU, DOUBLE},  {"ysu", YSU, DOUBLE},    {"zsu", ZSU, DOUBLE},
      {"ix", IX, INT},       {"iy", IY, INT},       {"iz", IZ, INT},         {"vx", VX, DOUBLE},
      {"vy", VY, DOUBLE},    {"vz", VZ, DOUBLE},    {"fx", FX, DOUBLE},      {"fy", FY, DOUBLE},
      {"fz", FZ, DOUBLE}};
  for (size_t k = 5; k < args.size(); k++) {
    const std::string &a = args[k];
    columns_ += a + " ";
    bool found = false;
    for (const Name &nm : names)
      if (a == nm.s) {
        fields_.push_back({nm.k, nm.t, 0, nullptr});
        found = true;
        break;
      }
    if (found) continue;
    if (a.compare(0, 2, "c_") != 0) {
      // properties this build's atom style does not hold (mol, q, mu, radius, ...)
      static const char *absent[] = {"mol", "q", "mux", "muy", "muz", "mu", "radius", "diameter", "omegax",
                                     "omegay", "omegaz", "angmomx", "angmomy", "angmomz", "tqx", "tqy", "tqz"};
      for (const char *p : absent)
        if (a == p) throw Error("Dumping an atom property that isn't allocated");
      throw Error("Invalid attribute " + a + " in dump custom command");
    }
    // c_ID or c_ID[k] (src/dump_custom.cpp:1196-1228)
    std::string name = a.substr(2);
    int argindex = 0;
    const size_t br = name.find('[');
    if (br != std::string::npos) {
      if (name.back() != ']') throw Error("Invalid attribute " + a + " in dump custom command");
      argindex = std::atoi(name.c_str() + br + 1);
      name = name.substr(0, br);
    }
    auto it = computes.find(name);
    if (it == computes.end()) throw Error("Could not find dump custom compute ID");
    const ComputePerAtom &c = it->second;
    if (argindex == 0 && c.size_peratom_cols() > 0)
      throw Error("Dump custom compute does not calculate per-atom vector");
    if (argindex > 0 && c.size_peratom_cols() == 0)
      throw Error("Dump custom compute does not calculate per-atom array");
    if (argindex < 0 || argindex > c.size_peratom_cols())
      throw Error("Dump custom compute vector is accessed out-of-range");
    computes_.push_back(c);
    fields_.push_back({c.style == ComputePerAtom::PE ? C_PE : C_STRESS, DOUBLE, argindex - 1, &computes_.back()});
    if (c.pair) needs_peratom_ = true;
  }
  init_formats();
}

DumpCustom::~DumpCustom() {
  if (fp_) std::fclose(fp_);
}

// One printf conversion per token, of the kind the column's value needs:
// d/i for integers, e/f/g for doubles, s for element names.  The reference
// hands the token to fprintf unchecked; a mismatched token is refused here.
static void check_format(const std::string &tok, char kind) {
  size_t p = 0, convs = 0;
  char conv = 0;
  while ((p = tok.find('%', p)) != std::string::npos) {
    if (p + 1 < tok.size() && tok[p + 1] == '%') {
      p += 2;
      continue;
    }
    size_t q = p + 1;
    while (q < tok.size() && std::strchr("-+ #0", tok[q])) q++;
    while (q < tok.size() && std::isdigit((unsigned char)tok[q])) q++;
    if (q < tok.size() && tok[q] == '.') {
      q++;
      while (q < tok.size() && std::isdigit((unsigned char)tok[q])) q++;
    }
    if (kind == 'g' && q < tok.size() && tok[q] == 'l') q++;
    if (q >= tok.size()) throw Error("Invalid dump_modify format token " + tok);
    conv = tok[q];
    convs++;
    p = q + 1;
  }
  const char *ok = kind == 'd' ? "di" : kind == 'g' ? "eEfFgG" : "s";
  if (convs != 1 || !std::strchr(ok, conv)) throw Error("Invalid dump_modify format token " + tok);
}

void DumpCustom::init_formats() {
  // src/dump_custom.cpp:126-137 (defaults) and 200-232 (tokenise, one space appended)
  std::vector<std::string> toks;
  if (format_user_.empty()) {
    for (const Field &fl : fields_) toks.push_back(fl.vtype == INT ? "%d" : fl.vtype == STRING ? "%s" : "%g");
  } else {
    std::istringstream ss(format_user_);
    std::string t;
    while (ss >> t) toks.push_back(t);
    if (toks.size() < fields_.size()) throw Error("Dump_modify format string is too short");
  }
  vformat_.clear();
  for (size_t i = 0; i < fields_.size(); i++) {
    check_format(toks[i], fields_[i].vtype == INT ? 'd' : fields_[i].vtype == STRING ? 's' : 'g');
    vformat_.push_back(toks[i] + " ");
  }
}

void DumpCustom::modify(const std::vector<std::string> &args) {
  // args: keyword value ... (src/dump.cpp:760-880, src/dump_custom.cpp:1455-1475)
  size_t k = 0;
  while (k < args.size()) {
    const std::string &key = args[k];
    if (key == "element") {
      if (k + 1 >= args.size()) throw Error("Illegal dump_modify command");
      typenames_.assign(1, "");
      for (k++; k < args.size() && !(args[k] == "format" || args[k] == "sort" || args[k] == "pad" ||
                                      args[k] == "flush" || args[k] == "first");
           k++)
        typenames_.push_back(args[k]);
      continue;
    }
    if (k + 1 >= args.size()) throw Error("Illegal dump_modify command");
    const std::string &val = args[k + 1];
    if (key == "format") {
      format_user_ = val == "none" ? "" : val;
      init_formats();
    } else if (key == "sort") {
      if (val == "off") sortcol_ = -1;
      else if (val == "id") {
        sortcol_ = 0;
        descend_ = false;
      } else {
        const int c = std::atoi(val.c_str());
        if (c == 0) throw Error("Illegal dump_modify command");
        sortcol_ = std::abs(c);
        descend_ = c < 0;
        if (sortcol_ > (int)fields_.size()) throw Error("Dump sort column is invalid");
      }
    } else if (key == "pad") {
      pad_ = std::atoi(val.c_str());
      if (pad_ < 0) throw Error("Illegal dump_modify command");
    } else if (key == "flush") {
      flush_ = val == "yes";
    } else if (key == "first") {
      first_flag = val == "yes";
    } else {
      throw Error("Illegal dump_modify command");
    }
    k += 2;
  }
}

void DumpCustom::open_for(long step) {
  // src/dump.cpp:430-470: one file, or one per snapshot with '*' -> step
  if (!multifile_ && fp_) return;
  if (fp_) std::fclose(fp_);
  std::string name = file_;
  if (multifile_) {
    const size_t star = name.find('*');
    char num[32];
    std::snprintf(num, sizeof(num), "%0*ld", pad_, step);
    name = name.substr(0, star) + num + name.substr(star + 1);
  }
  fp_ = std::fopen(name.c_str(), "w");
  if (!fp_) throw Error("Cannot open dump file " + name);
}

double DumpCustom::value(const Field &fl, const Snapshot &s, int i, const double *stress) const {
  const double prd[3] = {s.boxhi[0] - s.boxlo[0], s.boxhi[1] - s.boxlo[1], s.boxhi[2] - s.boxlo[2]};
  const double *xi = s.x + 3 * i;
  const int *im = s.image + 3 * i;
  switch (fl.kind) {
    case ID: return i + 1;
    case PROC: return 0;
    case PROCP1: return 1;
    case TYPE: case ELEMENT: return s.type[i];
    case MASS: return s.mass[s.type[i]];
    case X: case Y: case Z: return xi[fl.kind - X];
    // src/dump_custom.cpp:1966-2200 (scaled, unwrapped, scaled unwrapped)
    case XS: case YS: case ZS: {
      const int d = fl.kind - XS;
      return (xi[d] - s.boxlo[d]) * (1.0 / prd[d]);
    }
    case XU: case YU: case ZU: {
      const int d = fl.kind - XU;
      return xi[d] + im[d] * prd[d];
    }
    case XSU: case YSU: case ZSU: {
      const int d = fl.kind - XSU;
      return (xi[d] - s.boxlo[d]) * (1.0 / prd[d]) + im[d];
    }
    case IX: case IY: case IZ: return im[fl.kind - IX];
    case VX: case VY: case VZ: return s.v[3 * i + fl.kind - VX];
    case FX: case FY: case FZ: return s.f[3 * i + fl.kind - FX];
    case C_PE: return fl.compute->pair ? s.eatom[i] : 0.0;
    case C_STRESS: return stress[fl.col];
  }
  return 0.0;
}

// stress/atom of atom i (src/compute_stress_atom.cpp:140-360): the pair
// virial plus mvv2e m v_a v_b, times -nktv2p; component order xx yy zz xy xz yz
static void stress_of(const ComputePerAtom &c, const Snapshot &s, int i, double out[6]) {
  for (int k = 0; k < 6; k++) out[k] = c.pair ? s.vatom[6 * i + k] : 0.0;
  if (c.ke) {
    const double onemass = s.mvv2e * s.mass[s.type[i]];
    const double *v = s.v + 3 * i;
    out[0] += onemass * v[0] * v[0];
    out[1] += onemass * v[1] * v[1];
    out[2] += onemass * v[2] * v[2];
    out[3] += onemass * v[0] * v[1];
    out[4] += onemass * v[0] * v[2];
    out[5] += onemass * v[1] * v[2];
  }
  const double nktv2p = -s.nktv2p;
  for (int k = 0; k < 6; k++) out[k] *= nktv2p;
}

void DumpCustom::write(const Snapshot &s) {
  open_for(s.step);
  const int nf = (int)fields_.size();
  // pack every row (Dump::write -> pack), then order (Dump::sort)
  std::vector<double> buf((size_t)s.n * nf);
  for (int i = 0; i < s.n; i++) {
    double *row = buf.data() + (size_t)i * nf;
    const ComputePerAtom *last_c = nullptr;
    double st[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nf; j++) {
      const Field &fl = fields_[j];
      if (fl.kind == C_STRESS && fl.compute != last_c) {
        stress_of(*fl.compute, s, i, st);
        last_c = fl.compute;
      }
      row[j] = value(fl, s, i, st);
    }
  }
  std::vector<int> order(s.n);
  std::iota(order.begin(), order.end(), 0);  // tag order
  if (sortcol_ > 0) {
    const int c = sortcol_ - 1;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      const double va = buf[(size_t)a * nf + c], vb = buf[(size_t)b * nf + c];
      return descend_ ? va > vb : va < vb;
    });
  }
  // header (src/dump_custom.cpp:351-362)
  std::fprintf(fp_, "ITEM: TIMESTEP\n%ld\n", s.step);
  std::fprintf(fp_, "ITEM: NUMBER OF ATOMS\n%ld\n", (long)s.n);
  std::fprintf(fp_, "ITEM: BOX BOUNDS %s\n", s.boundstr.c_str());
  for (int d = 0; d < 3; d++) std::fprintf(fp_, "%g %g\n", s.boxlo[d], s.boxhi[d]);
  std::fprintf(fp_, "ITEM: ATOMS %s\n", columns_.c_str());
  // rows (src/dump_custom.cpp:984-1000); element names default to "C"
  for (int i : order) {
    const double *row = buf.data() + (size_t)i * nf;
    for (int j = 0; j < nf; j++) {
      const Field &fl = fields_[j];
      if (fl.vtype == INT) {
        std::fprintf(fp_, vformat_[j].c_str(), static_cast<int>(row[j]));
      } else if (fl.vtype == STRING) {
        const int t = static_cast<int>(row[j]);
        const char *nm = t < (int)typenames_.size() ? typenames_[t].c_str() : "C";
        std::fprintf(fp_, vformat_[j].c_str(), nm);
      } else {
        std::fprintf(fp_, vformat_[j].c_str(), row[j]);
      }
    }
    std::fputc('\n', fp_);
  }
  if (flush_) std::fflush(fp_);
  if (multifile_) {
    std::fclose(fp_);
    fp_ = nullptr;
  }
  last = s.step;
}

}  // namespace lmpmdx
