/* run_style verlet/mdx: the whole run on the device (libmdx mdx_setup /
   mdx_run): Verlet's step order with fix nve, the neighbour decisions and
   builds, and the force kernels, with x, v, f and the lists resident in HBM.
   Goes into the reference as src/MDX/verlet_mdx.h, an Integrate style as
   src/verlet.h.  Requires pair_style lj/cut/mdx or eam/mdx (whose device
   context it drives) and integrates every atom as fix nve does. */
#ifdef INTEGRATE_CLASS

IntegrateStyle(verlet/mdx,VerletMDX)

#else

#ifndef LMP_VERLET_MDX_H
#define LMP_VERLET_MDX_H

#include "verlet.h"

struct mdx_ctx;

namespace LAMMPS_NS {

class VerletMDX : public Verlet {
 public:
  VerletMDX(class LAMMPS *, int, char **);
  void setup();
  void run(int);

 private:
  mdx_ctx *ctx;
  double raw[8];  // last thermo sums from the device: sum m v^2, eng_vdwl, virial[6]
  void check(int rc);
  void to_host();
  static void thermo_row(void *self, long step, const double *row);
};

}  // namespace LAMMPS_NS

#endif
#endif
