/* pair_style lj/cut/mdx (see pair_lj_cut_mdx.h).  Replaces the calls of
   src/GPU/pair_lj_cut_gpu.cpp:42-65 (ljl_gpu_init / ljl_gpu_compute) with
   the libmdx C ABI. */
#include "pair_lj_cut_mdx.h"

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "force.h"
#include "neighbor.h"
#include "mdx.h"

using namespace LAMMPS_NS;

PairLJCutMDX::PairLJCutMDX(LAMMPS *lmp) : PairLJCut(lmp), ctx(NULL) {
  if (comm->nprocs != 1) error->all(FLERR, "lj/cut/mdx runs on one process owning the whole box");
  check(mdx_create(0, &ctx));
}

PairLJCutMDX::~PairLJCutMDX() { mdx_destroy(ctx); }

void PairLJCutMDX::check(int rc) {
  if (rc) error->one(FLERR, mdx_last_error());
}

// no neighbour request: the list is built on the device
void PairLJCutMDX::init_style() {
  int periodic[3] = {domain->xperiodic, domain->yperiodic, domain->zperiodic};
  check(mdx_set_box(ctx, domain->boxlo, domain->boxhi, periodic));
  check(mdx_neigh_modify(ctx, neighbor->skin, neighbor->every, neighbor->delay, neighbor->dist_check,
                         neighbor->oneatom));
}

// The (ntypes+1)^2 tables go to the device once init_one has filled them
// all: (ntypes, ntypes) is the last pair Pair::init visits.  Pair::init
// stores cutsq[i][j] only after init_one returns (src/pair.cpp:220-223), so
// this pair's entry is stored here first, as PairLJCutGPU::init_style
// fills cutsq before its upload (src/GPU/pair_lj_cut_gpu.cpp:136-158).
double PairLJCutMDX::init_one(int i, int j) {
  double cut = PairLJCut::init_one(i, j);
  cutsq[i][j] = cutsq[j][i] = cut * cut;
  if (i == atom->ntypes && j == atom->ntypes)
    check(mdx_pair_lj_init(ctx, atom->ntypes, cutsq[0], lj1[0], lj2[0], lj3[0], lj4[0], offset[0]));
  return cut;
}

void PairLJCutMDX::compute(int eflag, int vflag) {
  if (eflag || vflag) ev_setup(eflag, vflag);
  else evflag = vflag_fdotr = 0;
  const int n = atom->nlocal;
  if (neighbor->ago == 0) {  // atoms may have been sorted or exchanged
    check(mdx_set_atoms(ctx, n, atom->ntypes, atom->x[0], NULL, atom->type, (const int *)atom->tag, atom->mass));
    check(mdx_neigh_build(ctx));
  } else {
    check(mdx_set_positions(ctx, atom->x[0]));
  }
  double e = 0.0, v[6];
  // forces are added into atom->f, as every Pair::compute does
  check(mdx_compute(ctx, eflag, vflag, atom->f[0], &e, v, eflag_atom ? eatom : NULL, vflag_atom ? vatom[0] : NULL));
  if (eflag_global) eng_vdwl += e;
  if (vflag_global)
    for (int k = 0; k < 6; k++) virial[k] += v[k];
  vflag_fdotr = 0;  // the device returned the pairwise virial
}
