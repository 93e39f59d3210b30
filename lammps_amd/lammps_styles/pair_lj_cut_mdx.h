/* pair_style lj/cut/mdx: PairLJCut with its compute on libmdx (include/mdx.h).
   Goes into the reference as src/MDX/pair_lj_cut_mdx.h; follows the shape of
   src/GPU/pair_lj_cut_gpu.h (a CPU style subclass that hands the tables and
   coordinates to an accelerator library).  One process owns the whole box:
   libmdx builds the list on the device and handles periodic images itself. */
#ifdef PAIR_CLASS

PairStyle(lj/cut/mdx,PairLJCutMDX)

#else

#ifndef LMP_PAIR_LJ_CUT_MDX_H
#define LMP_PAIR_LJ_CUT_MDX_H

#include "pair_lj_cut.h"

struct mdx_ctx;

namespace LAMMPS_NS {

class PairLJCutMDX : public PairLJCut {
 public:
  PairLJCutMDX(class LAMMPS *);
  ~PairLJCutMDX();
  void compute(int, int);
  void init_style();
  double init_one(int, int);
  mdx_ctx *context() { return ctx; }  // the device state run_style verlet/mdx drives

 private:
  mdx_ctx *ctx;
  void check(int rc);
};

}  // namespace LAMMPS_NS

#endif
#endif
