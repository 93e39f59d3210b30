// output.h - per-atom computes (pe/atom, stress/atom) and the custom dump
// of the script runner.  Command syntax, column names, file layout and
// error messages follow src/compute_pe_atom.cpp, src/compute_stress_atom.cpp,
// src/dump.cpp and src/dump_custom.cpp.
#pragma once
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "plugin.h"

namespace lmpmdx {

// compute ID group pe/atom [pair bond ...]     (src/compute_pe_atom.cpp:28-70)
// compute ID group stress/atom NULL [keywords]  (src/compute_stress_atom.cpp:39-110)
// This build has pair forces only, so the contributions that can be non-zero
// are the pair term and, for stress/atom, the kinetic term.
struct ComputePerAtom {
  enum Style { PE, STRESS };
  Style style = PE;
  bool pair = true, ke = false;
  // 0 for a per-atom vector (pe/atom), 6 for the stress/atom array
  int size_peratom_cols() const { return style == PE ? 0 : 6; }
  // args: the command's words after "compute" (ID group style ...)
  static ComputePerAtom create(const std::vector<std::string> &args);
};

// The atoms at one step, in tag order (tag k+1 at row k).
struct Snapshot {
  long step = 0;
  int n = 0;
  const int *type = nullptr;
  const double *mass = nullptr;  // per type, [ntypes+1]
  const double *x = nullptr, *v = nullptr, *f = nullptr;
  const int *image = nullptr;    // 3 ints per atom
  const double *eatom = nullptr, *vatom = nullptr;  // pair per-atom energy / virial
  double boxlo[3] = {0, 0, 0}, boxhi[3] = {0, 0, 0};
  std::string boundstr;  // "pp pp pp" (src/domain.cpp:1554-1567)
  double mvv2e = 1.0, nktv2p = 1.0;
};

// dump ID group custom N file field ... (src/dump_custom.cpp:55-150)
class DumpCustom {
 public:
  // args: the command's words after "dump"
  DumpCustom(const std::vector<std::string> &args, const std::map<std::string, ComputePerAtom> &computes);
  ~DumpCustom();
  DumpCustom(const DumpCustom &) = delete;
  DumpCustom &operator=(const DumpCustom &) = delete;
  // dump_modify ID keyword value ... (src/dump.cpp:760-880): format, sort,
  // pad, flush, first, element
  void modify(const std::vector<std::string> &args);
  // the per-atom pair energy and virial must be tallied for this dump
  bool needs_peratom() const { return needs_peratom_; }
  // write one snapshot (src/dump.cpp:300-400, src/dump_custom.cpp:351-362,984-1000)
  void write(const Snapshot &s);

  std::string id;
  int every = 0;
  long last = -1;         // last step written (Output::last_dump)
  bool first_flag = false;

 private:
  enum Kind { ID, PROC, PROCP1, TYPE, ELEMENT, MASS, X, Y, Z, XS, YS, ZS, XU, YU, ZU, XSU, YSU, ZSU,
              IX, IY, IZ, VX, VY, VZ, FX, FY, FZ, C_PE, C_STRESS };
  enum VType { INT, DOUBLE, STRING };
  struct Field {
    Kind kind;
    VType vtype;
    int col;                // stress/atom column 0-5
    const ComputePerAtom *compute;
  };
  std::vector<Field> fields_;
  std::vector<ComputePerAtom> computes_;  // copies, in field order of first use
  std::string file_, columns_, format_user_;
  std::vector<std::string> vformat_;
  std::vector<std::string> typenames_;
  bool multifile_ = false, flush_ = true, needs_peratom_ = false;
  int pad_ = 0;
  int sortcol_ = -1;  // -1 off, 0 id, k > 0 column k
  bool descend_ = false;
  FILE *fp_ = nullptr;

  void init_formats();
  void open_for(long step);
  double value(const Field &fl, const Snapshot &s, int i, const double *stress) const;
};

}  // namespace lmpmdx
