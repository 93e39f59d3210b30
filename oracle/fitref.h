/* fitref.h - TEST INFRASTRUCTURE ONLY (CPU oracle; never part of the product).
 *
 * FP64 restatement of force evaluation on a batch of independent
 * configurations ("replicas", each an orthogonal box, periodic or not per
 * dimension) and of the gradients of a force-matching loss with respect to
 * the potential parameters:
 *   - lj/cut forces and energies (pair_lj_cut.cpp:68-141, init_one
 *     :545-602), and dL/d(epsilon_ab, sigma_ab) given G_i = dL/dF_i;
 *   - eam forces and energies from knot values (pair_eam.cpp:125-306 with
 *     the tables array2spline/interpolate builds, :698-751), and dL/d(knot
 *     values) of frho, rhor and z2r given G_i.
 * The reference has no gradient code; the adjoint is derived here from the
 * reference's force expressions and pinned by central finite differences
 * of this file's own FP64 forces (tests/test_fit_oracle.py).
 */
#ifndef FITREF_H
#define FITREF_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct fitref fitref;

/* nrep replicas; replica r holds natoms[r] atoms, in order, of the arrays
   type [N] (1-based) and x [N][3]; lo/hi [nrep][3], periodic [nrep][3]. */
fitref *fitref_new(int nrep, const int *natoms, const double *lo, const double *hi, const int *periodic,
                   int ntypes, const int *type, const double *x);
void fitref_free(fitref *s);
const char *fitref_error(fitref *s);

/* lj/cut: eps, sigma, cut are (ntypes+1)^2 row-major tables (symmetric).
   f [N][3]; energy [nrep] (may be NULL). */
int fitref_lj_forces(fitref *s, const double *eps, const double *sigma, const double *cut, double *f,
                     double *energy);
/* geps, gsigma (ntypes+1)^2: dL/d(eps_ab), dL/d(sigma_ab) accumulated into
   the [a][b] entry with a <= b (a symmetric parameter gets both orders). */
int fitref_lj_vjp(fitref *s, const double *eps, const double *sigma, const double *cut, const double *gf,
                  double *geps, double *gsigma);

/* eam, one element for every type: knot values frho[nrho], rhor[nr],
   z2r[nr] (the arrays interpolate() turns into splines, pair_eam.cpp:723),
   spacing drho, dr, force cutoff cut; rhomax = (nrho-1) drho. */
int fitref_eam_set(fitref *s, int nrho, double drho, int nr, double dr, double cut, const double *frho,
                   const double *rhor, const double *z2r);
/* f [N][3]; energy [nrep], rho [N], fp [N] may be NULL. */
int fitref_eam_forces(fitref *s, double *f, double *energy, double *rho, double *fp);
/* dL/d(knot values) for G = gf [N][3]. */
int fitref_eam_vjp(fitref *s, const double *gf, double *g_frho, double *g_rhor, double *g_z2r);

/* the 7-coefficient spline of n knots (pair_eam.cpp:723-751), (n+1)*7 */
void fitref_spline(int n, double delta, const double *knots /* [n] */, double *spline);

#ifdef __cplusplus
}
#endif
#endif
