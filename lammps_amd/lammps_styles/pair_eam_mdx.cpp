/* pair_style eam/mdx (see pair_eam_mdx.h).  Replaces eam_gpu_init /
   eam_gpu_compute / eam_gpu_compute_force (src/GPU/pair_eam_gpu.cpp:38-140)
   with the libmdx C ABI. */
#include "pair_eam_mdx.h"

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "force.h"
#include "neighbor.h"
#include "mdx.h"

using namespace LAMMPS_NS;

PairEAMMDX::PairEAMMDX(LAMMPS *lmp) : PairEAM(lmp), ctx(NULL) {
  if (comm->nprocs != 1) error->all(FLERR, "eam/mdx runs on one process owning the whole box");
  check(mdx_create(0, &ctx));
}

PairEAMMDX::~PairEAMMDX() { mdx_destroy(ctx); }

void PairEAMMDX::check(int rc) {
  if (rc) error->one(FLERR, mdx_last_error());
}

// file2array and array2spline run in PairEAM::init_style; no neighbour request
void PairEAMMDX::init_style() {
  PairEAM::init_style();
  int periodic[3] = {domain->xperiodic, domain->yperiodic, domain->zperiodic};
  check(mdx_set_box(ctx, domain->boxlo, domain->boxhi, periodic));
  check(mdx_neigh_modify(ctx, neighbor->skin, neighbor->every, neighbor->delay, neighbor->dist_check,
                         neighbor->oneatom));
}

// the spline arrays as array2spline left them, once cutforcesq is known
double PairEAMMDX::init_one(int i, int j) {
  double cut = PairEAM::init_one(i, j);
  if (i == atom->ntypes && j == atom->ntypes)
    check(mdx_pair_eam_init(ctx, atom->ntypes, nrho, drho, nr, dr, rhomax, cutforcesq, nfrho, frho_spline[0][0],
                            nrhor, rhor_spline[0][0], nz2r, z2r_spline[0][0], type2frho, type2rhor[0],
                            type2z2r[0]));
  return cut;
}

void PairEAMMDX::compute(int eflag, int vflag) {
  if (eflag || vflag) ev_setup(eflag, vflag);
  else evflag = vflag_fdotr = 0;
  const int n = atom->nlocal;
  if (neighbor->ago == 0) {
    check(mdx_set_atoms(ctx, n, atom->ntypes, atom->x[0], NULL, atom->type, (const int *)atom->tag, atom->mass));
    check(mdx_neigh_build(ctx));
  } else {
    check(mdx_set_positions(ctx, atom->x[0]));
  }
  double e = 0.0, v[6];
  check(mdx_compute(ctx, eflag, vflag, atom->f[0], &e, v, eflag_atom ? eatom : NULL, vflag_atom ? vatom[0] : NULL));
  if (eflag_global) eng_vdwl += e;
  if (vflag_global)
    for (int k = 0; k < 6; k++) virial[k] += v[k];
  vflag_fdotr = 0;
}
