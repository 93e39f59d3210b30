/* mdx_host.h - host-side set-up commands of the reference, for callers that
 * drive libmdx without LAMMPS around it (the bench, the script runner).
 *
 * These are CPU routines (no device work).  Each restates one reference
 * command so that the device path starts from the same seeded state:
 *   create_atoms on a lattice (src/lattice.cpp:80-300, src/create_atoms.cpp:269-326,659-770)
 *   velocity create          (src/velocity.cpp:136-340,666-716; src/random_park.cpp:42-130)
 *   pair_coeff * * file.eam  (src/MANYBODY/pair_eam.cpp:346-481 read_file,
 *                             488-694 file2array, 698-751 array2spline)
 */
#ifndef MDX_HOST_H
#define MDX_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

enum { MDX_LATTICE_SC = 0, MDX_LATTICE_BCC = 1, MDX_LATTICE_FCC = 2 };
enum { MDX_UNITS_LJ = 0, MDX_UNITS_METAL = 1 };

/* `lattice <style> <arg>` + `region box block lo hi ...` + `create_box` +
   `create_atoms 1 box`: fills boxlo/boxhi and up to cap atoms into x [n][3]
   in the reference's creation (= atom ID) order; *n gets the count (also
   when it exceeds cap, then MDX_ERR_ARG is returned).  arg is rho* in lj
   units and the lattice constant in metal units. */
int mdx_host_lattice(int style, double arg, int units, const int lo[3], const int hi[3], const int periodic[3],
                     double *boxlo, double *boxhi, double *x, long cap, long *n);

/* `velocity all create T seed dist uniform|gaussian loop all|geom mom yes|no`
   for atoms with IDs 1..n in array order; writes v [n][3]. */
int mdx_host_velocity(long n, const double *x, const int *type, const double *mass, double t_desired, int seed,
                      int gaussian, int loop_geom, int mom, double boltz, double mvv2e, double *v);

/* Park-Miller minimal standard generator state (RanPark::uniform). */
double mdx_host_ranpark_uniform(int *seed);

/* EAM tables of one funcfl file as PairEAM leaves them after init_style:
   call with NULL arrays to get the sizes, then with arrays of
   nfrho*(nrho+1)*7, nrhor*(nr+1)*7, nz2r*(nr+1)*7 doubles.
   The type maps are those of `pair_coeff * * file` with every type on the file. */
typedef struct {
  int nrho, nr, nfrho, nrhor, nz2r;
  double drho, dr, rhomax, cutmax, mass;
} mdx_eam_info;
int mdx_host_eam_funcfl(const char *path, mdx_eam_info *info, double *frho_spline, double *rhor_spline,
                        double *z2r_spline);

/* eam/alloy (fs = 0) or eam/fs (fs = 1) DYNAMO setfl file, with atom type t
   on the element named type_elements[t-1] (pair_coeff * * file E1 .. En),
   as PairEAMAlloy/PairEAMFS read_file + coeff + file2array leave the
   tables (src/MANYBODY/pair_eam_alloy.cpp:52-330, pair_eam_fs.cpp:52-330):
   nfrho = nelements + 1 (a zero table last), nrhor = nelements (alloy) or
   nelements^2 (fs, [i*n + j] = file block i, function j), nz2r =
   n(n+1)/2 (lower triangle, file order).  Call with NULL spline arrays for
   the sizes.  type2frho[ntypes+1], type2rhor/type2z2r[(ntypes+1)^2] and
   mass[ntypes+1] (element masses per type) may be NULL. */
int mdx_host_eam_setfl(const char *path, int fs, int ntypes, const char *const *type_elements, mdx_eam_info *info,
                       double *frho_spline, double *rhor_spline, double *z2r_spline, int *type2frho, int *type2rhor,
                       int *type2z2r, double *mass);

/* Writes a funcfl file tabulating F(rho) = -E0 sqrt(rho/rho_e),
   Z(r) = Z0 exp(-beta r) and rho(r) = fe exp(-chi (r/re - 1)) on nrho/nr
   points, cutoff `cut`; rho_e is the perfect-fcc density at lattice
   constant a (neighbour shells within cut).  Parameters (Z0, beta, chi, E0)
   are drawn with RanPark from `seed` in that order from the given ranges,
   re = a/sqrt(2), fe = 1, drho = 2 rho_e/(nrho-1); params_out gets
   {Z0, beta, chi, E0, re, rho_e} (may be NULL). */
int mdx_host_write_synthetic_funcfl(const char *path, int seed, double a, int nrho, int nr, double dr, double cut,
                                    double mass, double *params_out);

#ifdef __cplusplus
}
#endif
#endif
