/* pair_style eam/mdx: PairEAM (funcfl) with its compute on libmdx
   (include/mdx.h).  Goes into the reference as src/MDX/pair_eam_mdx.h; follows
   the shape of src/GPU/pair_eam_gpu.h.  libmdx keeps rho and fp on the device,
   so the forward and reverse fp communication of PairEAM is not needed. */
#ifdef PAIR_CLASS

PairStyle(eam/mdx,PairEAMMDX)

#else

#ifndef LMP_PAIR_EAM_MDX_H
#define LMP_PAIR_EAM_MDX_H

#include "pair_eam.h"

struct mdx_ctx;

namespace LAMMPS_NS {

class PairEAMMDX : public PairEAM {
 public:
  PairEAMMDX(class LAMMPS *);
  ~PairEAMMDX();
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
