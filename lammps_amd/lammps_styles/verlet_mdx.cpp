/* run_style verlet/mdx (see verlet_mdx.h).  Replaces Verlet::setup and
   Verlet::run with fix nve (src/verlet.cpp:88-142,207-309,
   src/fix_nve.cpp:65-142) by mdx_setup / mdx_run.  The device runs between
   output steps; at each one the state comes back to the atom arrays, the
   pair's energy and virial are set from the device's sums, and Output::write
   prints thermo and dumps as usual. */
#include "verlet_mdx.h"

#include <cstring>
#include <vector>

#include "atom.h"
#include "error.h"
#include "force.h"
#include "output.h"
#include "pair.h"
#include "pair_eam_mdx.h"
#include "pair_lj_cut_mdx.h"
#include "update.h"
#include "mdx.h"

using namespace LAMMPS_NS;

VerletMDX::VerletMDX(LAMMPS *lmp, int narg, char **arg) : Verlet(lmp, narg, arg), ctx(NULL) {
  std::memset(raw, 0, sizeof(raw));
}

void VerletMDX::check(int rc) {
  if (rc) error->one(FLERR, mdx_last_error());
}

void VerletMDX::thermo_row(void *self, long, const double *row) {
  std::memcpy(static_cast<VerletMDX *>(self)->raw, row, sizeof(double) * 8);
}

void VerletMDX::setup() {
  // the reference's setup first: the pair style uploads atoms and tables and
  // computes the step-0 forces, then output->setup prints the first row
  Verlet::setup();
  PairLJCutMDX *lj = dynamic_cast<PairLJCutMDX *>(force->pair);
  PairEAMMDX *eam = dynamic_cast<PairEAMMDX *>(force->pair);
  if (!lj && !eam) error->all(FLERR, "run_style verlet/mdx requires pair_style lj/cut/mdx or eam/mdx");
  ctx = lj ? lj->context() : eam->context();
  // velocities too, then Verlet::setup on the device (same positions)
  check(mdx_set_atoms(ctx, atom->nlocal, atom->ntypes, atom->x[0], atom->v[0], atom->type,
                      (const int *)atom->tag, atom->mass));
  check(mdx_setup(ctx, &VerletMDX::thermo_row, this));
}

// device state -> atom arrays (x, v, f, image) and the pair's tallies
void VerletMDX::to_host() {
  const int n = atom->nlocal;
  std::vector<int> im(3 * (size_t)n);
  check(mdx_get_atoms(ctx, atom->x[0], atom->v[0], atom->f[0], im.data()));
  for (int i = 0; i < n; i++)
    atom->image[i] = ((imageint)(im[3 * i + 2] + IMGMAX) << IMG2BITS) |
                     ((imageint)(im[3 * i + 1] + IMGMAX) << IMGBITS) | (imageint)(im[3 * i] + IMGMAX);
  force->pair->eng_vdwl = raw[1];
  for (int k = 0; k < 6; k++) force->pair->virial[k] = raw[2 + k];
  // the last device step tallied energy and virial (mdx_run's final step),
  // which compute pe / pressure check (Integrate::ev_set's bookkeeping)
  update->eflag_global = update->vflag_global = update->ntimestep;
}

void VerletMDX::run(int n) {
  const bigint end = update->ntimestep + n;
  while (update->ntimestep < end) {
    // run on the device up to the next output step (Output::next), as
    // Verlet::run calls output->write there
    bigint stop = output->next < end ? output->next : end;
    if (stop <= update->ntimestep) stop = end;
    const int chunk = (int)(stop - update->ntimestep);
    check(mdx_run(ctx, chunk, update->dt, force->ftm2v, 0, &VerletMDX::thermo_row, this));
    update->ntimestep = stop;
    to_host();
    if (update->ntimestep == output->next) output->write(update->ntimestep);
  }
}
