/* mdref.h - CPU FP64 oracle for the short-range MD path (TEST INFRASTRUCTURE).
 *
 * This header declares a plain-C restatement of the LAMMPS (8 Aug 2014)
 * algorithms on the binned-neighbour / lj/cut / eam / fix nve path.  It is
 * the checker that the HIP implementation in lammps_amd/ is compared with.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product never links or calls it.
 *
 * Every function in mdref.c cites the reference file:line it restates.
 */
#ifndef MDREF_H
#define MDREF_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdref mdref;

enum { MDREF_UNITS_LJ = 0, MDREF_UNITS_METAL = 1 };
enum { MDREF_LIST_HALF_NEWTON = 0, MDREF_LIST_HALF_NONEWTON = 1, MDREF_LIST_FULL = 2 };
enum { MDREF_PAIR_NONE = 0, MDREF_PAIR_LJ = 1, MDREF_PAIR_EAM = 2 };

mdref *mdref_new(void);
void mdref_free(mdref *s);
const char *mdref_error(mdref *s);

/* units lj|metal (update.cpp:118-180); sets dt and skin defaults */
int mdref_units(mdref *s, int style);
/* orthogonal box (create_box), periodic flags per dim (boundary p/f) */
int mdref_box(mdref *s, const double lo[3], const double hi[3], const int periodic[3], int ntypes);
/* lattice style: 0=sc 1=bcc 2=fcc; arg = rho* in lj units or a in metal;
   region block [ilo,ihi) in lattice units; create_atoms type 1 box.
   Creates the box from the region too (create_box). */
int mdref_lattice_box(mdref *s, int style, double arg, const int lo[3], const int hi[3],
                      const int periodic[3], int ntypes);
/* append atoms (x: n*3, type 1-based); tags continue from natoms */
int mdref_add_atoms(mdref *s, int n, const double *x, const int *type);
int mdref_set_type(mdref *s, int n, const int *type);   /* by local index */
int mdref_mass(mdref *s, int type, double m);
/* velocity all create T seed dist(0 uniform/1 gaussian) loop(0 all/1 geom) mom(0/1) */
int mdref_velocity_create(mdref *s, double t, int seed, int dist, int loop_geom, int mom);
int mdref_set_velocities(mdref *s, int n, const double *v);   /* by local index */

int mdref_pair_lj(mdref *s, double cut_global, int shift);
int mdref_pair_lj_coeff(mdref *s, int i, int j, double eps, double sigma, double cut /* <0: global */);
int mdref_pair_eam_file(mdref *s, const char *path);   /* pair_coeff * * file (funcfl, 1 type) */
/* pair_style eam/alloy (fs = 0) or eam/fs (fs = 1): pair_coeff * * file E1 .. En,
   type_elements[t-1] = element name of type t */
int mdref_pair_eam_setfl(mdref *s, const char *path, int fs, const char *const *type_elements);
/* EAM table accessors after init (pair_eam.h public members) */
int mdref_eam_info(mdref *s, int *nrho, double *drho, int *nr, double *dr, double *rhomax, double *cutmax,
                   int *nfrho, int *nrhor, int *nz2r);
int mdref_eam_splines(mdref *s, double *frho_spline, double *rhor_spline, double *z2r_spline,
                      int *type2frho, int *type2rhor, int *type2z2r);
double mdref_mass_of(mdref *s, int type);

int mdref_neighbor(mdref *s, double skin, int every, int delay, int check, int one);
int mdref_newton(mdref *s, int newton);          /* newton on/off (pair) */
int mdref_list_style(mdref *s, int style);       /* half newton / half no-newton / full */
int mdref_sortfreq(mdref *s, int freq);
int mdref_timestep(mdref *s, double dt);
int mdref_thermo_every(mdref *s, int n);

/* Verlet::setup: init, pbc, ghosts, sort, build, force at step 0 */
int mdref_setup(mdref *s);
/* run n steps; each thermo row written as {step,temp,epair,emol,etotal,press} */
int mdref_run(mdref *s, int nsteps, double *thermo, int max_rows, int *nrows);
/* one force evaluation on the current positions (rebuilds ghosts + lists first);
   eflag/vflag as Integrate::ev_set (1 global, 2 per-atom; vflag 1/2 global, 4 per-atom) */
int mdref_compute(mdref *s, int eflag, int vflag);

/* Threads of the threaded CPU-baseline build (libmdref_omp.so, -fopenmp
   -DMDREF_OMP): sets the count when n > 0 and returns it; the serial build
   returns 1.  The threaded build runs the force passes of steps without
   energy/virial tallies and the list build on threads, with per-thread
   force arrays summed after each pass (the reference's USER-OMP scheme). */
int mdref_threads(int n);

int mdref_natoms(mdref *s);
int mdref_nghost(mdref *s);
/* per-atom output ordered by tag (tag 1..N -> row 0..N-1) */
int mdref_get_by_tag(mdref *s, double *x, double *v, double *f, int *type, int *image);
int mdref_get_peratom_by_tag(mdref *s, double *eatom, double *vatom);
int mdref_get_eam_by_tag(mdref *s, double *rho, double *fp);
int mdref_energy_virial(mdref *s, double *eng, double *virial6);
int mdref_thermo_now(mdref *s, double *row6);
/* neighbour statistics: total neighbours in the current list, builds, dangerous builds */
int mdref_neigh_stats(mdref *s, long *total, int *builds, int *dangerous, int *maxneigh);
/* current list as tag pairs (tag_i, tag_j) in list order; returns count (or needed size) */
long mdref_list_pairs(mdref *s, int *pairs, long maxpairs);
long mdref_ntimestep(mdref *s);

/* raw RanPark access (random_park.cpp) for pinning */
double mdref_ranpark_uniform_seq(int seed, int n, int *state_out);
int mdref_ranpark_reset_coord(int ibase, const double coord[3]);

#ifdef __cplusplus
}
#endif
#endif
