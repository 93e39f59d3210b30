/* fitref.c - TEST INFRASTRUCTURE ONLY: FP64 CPU oracle for batched force
 * evaluation and force-matching gradients (see fitref.h).  Built with
 * -ffp-contract=off so every product rounds on its own, as the reference's
 * x86-64 build does.
 *
 * Neighbours: a cell list per replica with cells no smaller than the largest
 * cutoff; every (cell, periodic shift) of the 27 surrounding cells in
 * unwrapped cell coordinates is visited once, so every periodic image within
 * the cutoff is found (the reference's ghost atoms, comm_brick.cpp:696-864),
 * also in boxes only one or two cells wide.  The ghost coordinate is
 * x_j + shift*prd rounded once (atom_vec_atomic.cpp:290-297).
 */
#include "fitref.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

struct fitref {
  int nrep, ntypes, n;
  int *natoms, *off;
  double *lo, *hi, *prd;
  int *periodic;
  int *type;
  double *x; /* wrapped into the box in periodic dimensions */
  char err[256];
  /* eam */
  int eam_set, nrho, nr;
  double drho, dr, rdrho, rdr, cutsq, rhomax;
  double *frho_sp, *rhor_sp, *z2r_sp;
};

typedef struct {
  int *first, *num, *j, *shift; /* CSR over the replica's atoms; shift = 9(sx+1)+3(sy+1)+(sz+1) */
  int cap;
} plist;

static int fail(fitref *s, const char *m) {
  snprintf(s->err, sizeof(s->err), "%s", m);
  return -1;
}
const char *fitref_error(fitref *s) { return s->err; }

fitref *fitref_new(int nrep, const int *natoms, const double *lo, const double *hi, const int *periodic,
                   int ntypes, const int *type, const double *x) {
  if (nrep <= 0 || ntypes <= 0) return NULL;
  fitref *s = (fitref *)calloc(1, sizeof(fitref));
  s->nrep = nrep;
  s->ntypes = ntypes;
  s->natoms = (int *)malloc(sizeof(int) * nrep);
  s->off = (int *)malloc(sizeof(int) * (nrep + 1));
  s->lo = (double *)malloc(sizeof(double) * 3 * nrep);
  s->hi = (double *)malloc(sizeof(double) * 3 * nrep);
  s->prd = (double *)malloc(sizeof(double) * 3 * nrep);
  s->periodic = (int *)malloc(sizeof(int) * 3 * nrep);
  s->off[0] = 0;
  for (int r = 0; r < nrep; r++) {
    s->natoms[r] = natoms[r];
    s->off[r + 1] = s->off[r] + natoms[r];
    for (int d = 0; d < 3; d++) {
      s->lo[3 * r + d] = lo[3 * r + d];
      s->hi[3 * r + d] = hi[3 * r + d];
      s->prd[3 * r + d] = hi[3 * r + d] - lo[3 * r + d];
      s->periodic[3 * r + d] = periodic[3 * r + d];
    }
  }
  s->n = s->off[nrep];
  s->type = (int *)malloc(sizeof(int) * (s->n + 1));
  s->x = (double *)malloc(sizeof(double) * 3 * (s->n + 1));
  memcpy(s->type, type, sizeof(int) * s->n);
  /* Domain::pbc (domain.cpp:478) in periodic dimensions */
  for (int r = 0; r < nrep; r++)
    for (int i = s->off[r]; i < s->off[r + 1]; i++)
      for (int d = 0; d < 3; d++) {
        double v = x[3 * i + d];
        if (s->periodic[3 * r + d]) {
          if (v < s->lo[3 * r + d]) v += s->prd[3 * r + d];
          if (v >= s->hi[3 * r + d]) v -= s->prd[3 * r + d];
          v = MAX(v, s->lo[3 * r + d]);
        }
        s->x[3 * i + d] = v;
      }
  return s;
}

void fitref_free(fitref *s) {
  if (!s) return;
  free(s->natoms); free(s->off); free(s->lo); free(s->hi); free(s->prd); free(s->periodic);
  free(s->type); free(s->x);
  free(s->frho_sp); free(s->rhor_sp); free(s->z2r_sp);
  free(s);
}

static void plist_free(plist *p) {
  free(p->first); free(p->num); free(p->j); free(p->shift);
  memset(p, 0, sizeof(*p));
}

/* full list of replica r with rsq < cutsq_max (strict), atoms in order */
static int build_list(fitref *s, int r, double cutmax, plist *pl) {
  const int a0 = s->off[r], n = s->natoms[r];
  const double *lo = &s->lo[3 * r], *prd = &s->prd[3 * r];
  const int *per = &s->periodic[3 * r];
  int nc[3];
  for (int d = 0; d < 3; d++) {
    nc[d] = (int)floor(prd[d] / cutmax);
    if (nc[d] < 1) nc[d] = 1;
    if (nc[d] > 512) nc[d] = 512;
  }
  const int ncell = nc[0] * nc[1] * nc[2];
  int *head = (int *)malloc(sizeof(int) * ncell), *next = (int *)malloc(sizeof(int) * (n + 1));
  int *cell = (int *)malloc(sizeof(int) * 3 * (n + 1));
  for (int c = 0; c < ncell; c++) head[c] = -1;
  for (int k = n - 1; k >= 0; k--) {
    const double *xi = &s->x[3 * (a0 + k)];
    int c3[3];
    for (int d = 0; d < 3; d++) {
      if (!per[d] && (xi[d] < lo[d] || xi[d] > lo[d] + prd[d])) {
        free(head); free(next); free(cell);
        return fail(s, "atom outside a non-periodic box dimension");
      }
      int ci = (int)floor((xi[d] - lo[d]) / prd[d] * nc[d]);
      c3[d] = MIN(MAX(ci, 0), nc[d] - 1);
      cell[3 * k + d] = c3[d];
    }
    const int c = (c3[2] * nc[1] + c3[1]) * nc[0] + c3[0];
    next[k] = head[c];
    head[c] = k;
  }
  const double cutsq = cutmax * cutmax;
  pl->first = (int *)malloc(sizeof(int) * (n + 1));
  pl->num = (int *)malloc(sizeof(int) * (n + 1));
  pl->cap = 64 * n + 64;
  pl->j = (int *)malloc(sizeof(int) * pl->cap);
  pl->shift = (int *)malloc(sizeof(int) * pl->cap);
  int used = 0;
  for (int k = 0; k < n; k++) {
    const double *xi = &s->x[3 * (a0 + k)];
    pl->first[k] = used;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int cc[3], sh[3] = {0, 0, 0}, ok = 1;
          const int dd[3] = {dx, dy, dz};
          for (int d = 0; d < 3; d++) {
            int u = cell[3 * k + d] + dd[d];
            if (per[d]) {
              /* unwrapped cell u -> (cell, shift) */
              int q = (int)floor((double)u / nc[d]);
              sh[d] = q;
              u -= q * nc[d];
            } else if (u < 0 || u >= nc[d]) {
              ok = 0;
            }
            cc[d] = u;
          }
          if (!ok) continue;
          const int c = (cc[2] * nc[1] + cc[1]) * nc[0] + cc[0];
          for (int m = head[c]; m >= 0; m = next[m]) {
            if (m == k && sh[0] == 0 && sh[1] == 0 && sh[2] == 0) continue;
            const double *xj = &s->x[3 * (a0 + m)];
            double gx = xj[0] + sh[0] * prd[0];
            double gy = xj[1] + sh[1] * prd[1];
            double gz = xj[2] + sh[2] * prd[2];
            double dx_ = xi[0] - gx, dy_ = xi[1] - gy, dz_ = xi[2] - gz;
            double rsq = dx_ * dx_ + dy_ * dy_ + dz_ * dz_;
            if (rsq < cutsq) {
              if (used == pl->cap) {
                pl->cap *= 2;
                pl->j = (int *)realloc(pl->j, sizeof(int) * pl->cap);
                pl->shift = (int *)realloc(pl->shift, sizeof(int) * pl->cap);
              }
              pl->j[used] = m;
              pl->shift[used] = 9 * (sh[0] + 1) + 3 * (sh[1] + 1) + (sh[2] + 1);
              used++;
            }
          }
        }
    pl->num[k] = used - pl->first[k];
  }
  free(head); free(next); free(cell);
  return 0;
}

/* del = x_i - (x_j + shift*prd) */
static double pair_del(fitref *s, int r, int i, int j, int code, double *del) {
  const double *prd = &s->prd[3 * r];
  const int sh[3] = {code / 9 - 1, (code / 3) % 3 - 1, code % 3 - 1};
  for (int d = 0; d < 3; d++) {
    double g = s->x[3 * j + d] + sh[d] * prd[d];
    del[d] = s->x[3 * i + d] - g;
  }
  return del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
}

static double max_cut(fitref *s, const double *cut) {
  double c = 0.0;
  const int nt1 = s->ntypes + 1;
  for (int a = 1; a <= s->ntypes; a++)
    for (int b = 1; b <= s->ntypes; b++) c = MAX(c, cut[a * nt1 + b]);
  return c;
}

/* ---------------------------------------------------------------- lj/cut */

int fitref_lj_forces(fitref *s, const double *eps, const double *sigma, const double *cut, double *f,
                     double *energy) {
  const int nt1 = s->ntypes + 1;
  const double cmax = max_cut(s, cut);
  if (!(cmax > 0)) return fail(s, "no positive cutoff");
  for (int r = 0; r < s->nrep; r++) {
    plist pl;
    memset(&pl, 0, sizeof(pl));
    if (build_list(s, r, cmax, &pl)) return -1;
    const int a0 = s->off[r];
    double e = 0.0;
    for (int k = 0; k < s->natoms[r]; k++) {
      const int i = a0 + k, ti = s->type[i];
      double fx = 0.0, fy = 0.0, fz = 0.0;
      for (int q = pl.first[k]; q < pl.first[k] + pl.num[k]; q++) {
        const int j = a0 + pl.j[q], tj = s->type[j];
        double del[3];
        const double rsq = pair_del(s, r, i, j, pl.shift[q], del);
        const int t = ti * nt1 + tj;
        if (rsq < cut[t] * cut[t]) {
          /* pair_lj_cut.cpp:108-112 with init_one's lj1..lj4 (:561-564) */
          const double lj1 = 48.0 * eps[t] * pow(sigma[t], 12.0), lj2 = 24.0 * eps[t] * pow(sigma[t], 6.0);
          const double lj3 = 4.0 * eps[t] * pow(sigma[t], 12.0), lj4 = 4.0 * eps[t] * pow(sigma[t], 6.0);
          const double r2inv = 1.0 / rsq, r6inv = r2inv * r2inv * r2inv;
          const double forcelj = r6inv * (lj1 * r6inv - lj2);
          const double fpair = forcelj * r2inv;
          fx += del[0] * fpair;
          fy += del[1] * fpair;
          fz += del[2] * fpair;
          e += 0.5 * (r6inv * (lj3 * r6inv - lj4));
        }
      }
      f[3 * i] = fx;
      f[3 * i + 1] = fy;
      f[3 * i + 2] = fz;
    }
    if (energy) energy[r] = e;
    plist_free(&pl);
  }
  return 0;
}

/* dfpair/deps = fpair/eps; dfpair/dsigma = r2inv r6inv (576 eps s^11 r6inv - 144 eps s^5) */
int fitref_lj_vjp(fitref *s, const double *eps, const double *sigma, const double *cut, const double *gf,
                  double *geps, double *gsigma) {
  const int nt1 = s->ntypes + 1;
  const double cmax = max_cut(s, cut);
  if (!(cmax > 0)) return fail(s, "no positive cutoff");
  double *ge = (double *)calloc(nt1 * nt1, sizeof(double)), *gs = (double *)calloc(nt1 * nt1, sizeof(double));
  for (int r = 0; r < s->nrep; r++) {
    plist pl;
    memset(&pl, 0, sizeof(pl));
    if (build_list(s, r, cmax, &pl)) { free(ge); free(gs); return -1; }
    const int a0 = s->off[r];
    for (int k = 0; k < s->natoms[r]; k++) {
      const int i = a0 + k, ti = s->type[i];
      for (int q = pl.first[k]; q < pl.first[k] + pl.num[k]; q++) {
        const int j = a0 + pl.j[q], tj = s->type[j];
        double del[3];
        const double rsq = pair_del(s, r, i, j, pl.shift[q], del);
        const int t = ti * nt1 + tj;
        if (rsq < cut[t] * cut[t]) {
          const double a = gf[3 * i] * del[0] + gf[3 * i + 1] * del[1] + gf[3 * i + 2] * del[2];
          const double r2inv = 1.0 / rsq, r6inv = r2inv * r2inv * r2inv;
          const double s6 = pow(sigma[t], 6.0), s12 = pow(sigma[t], 12.0);
          const double s5 = pow(sigma[t], 5.0), s11 = pow(sigma[t], 11.0);
          ge[t] += a * (r2inv * r6inv * (48.0 * s12 * r6inv - 24.0 * s6));
          gs[t] += a * (r2inv * r6inv * (576.0 * eps[t] * s11 * r6inv - 144.0 * eps[t] * s5));
        }
      }
    }
    plist_free(&pl);
  }
  for (int a = 0; a < nt1 * nt1; a++) geps[a] = gsigma[a] = 0.0;
  for (int a = 1; a <= s->ntypes; a++)
    for (int b = 1; b <= s->ntypes; b++) {
      const int lo = MIN(a, b), hi = MAX(a, b);
      geps[lo * nt1 + hi] += ge[a * nt1 + b];
      gsigma[lo * nt1 + hi] += gs[a * nt1 + b];
    }
  free(ge); free(gs);
  return 0;
}

/* ---------------------------------------------------------------- eam */

/* pair_eam.cpp:723-751 */
void fitref_spline(int n, double delta, const double *knots, double *sp) {
#define SP(m, k) sp[(m)*7 + (k)]
  for (int k = 0; k < 7; k++) SP(0, k) = 0.0;
  for (int m = 1; m <= n; m++) SP(m, 6) = knots[m - 1];
  SP(1, 5) = SP(2, 6) - SP(1, 6);
  SP(2, 5) = 0.5 * (SP(3, 6) - SP(1, 6));
  SP(n - 1, 5) = 0.5 * (SP(n, 6) - SP(n - 2, 6));
  SP(n, 5) = SP(n, 6) - SP(n - 1, 6);
  for (int m = 3; m <= n - 2; m++)
    SP(m, 5) = ((SP(m - 2, 6) - SP(m + 2, 6)) + 8.0 * (SP(m + 1, 6) - SP(m - 1, 6))) / 12.0;
  for (int m = 1; m <= n - 1; m++) {
    SP(m, 4) = 3.0 * (SP(m + 1, 6) - SP(m, 6)) - 2.0 * SP(m, 5) - SP(m + 1, 5);
    SP(m, 3) = SP(m, 5) + SP(m + 1, 5) - 2.0 * (SP(m + 1, 6) - SP(m, 6));
  }
  SP(n, 4) = 0.0;
  SP(n, 3) = 0.0;
  for (int m = 1; m <= n; m++) {
    SP(m, 2) = SP(m, 5) / delta;
    SP(m, 1) = 2.0 * SP(m, 4) / delta;
    SP(m, 0) = 3.0 * SP(m, 3) / delta;
  }
#undef SP
}

int fitref_eam_set(fitref *s, int nrho, double drho, int nr, double dr, double cut, const double *frho,
                   const double *rhor, const double *z2r) {
  if (nrho < 5 || nr < 5 || !(drho > 0) || !(dr > 0) || !(cut > 0)) return fail(s, "invalid eam tables");
  s->nrho = nrho;
  s->nr = nr;
  s->drho = drho;
  s->dr = dr;
  s->rdrho = 1.0 / drho;
  s->rdr = 1.0 / dr;
  s->cutsq = cut * cut;
  s->rhomax = (nrho - 1) * drho;
  free(s->frho_sp); free(s->rhor_sp); free(s->z2r_sp);
  s->frho_sp = (double *)malloc(sizeof(double) * (nrho + 1) * 7);
  s->rhor_sp = (double *)malloc(sizeof(double) * (nr + 1) * 7);
  s->z2r_sp = (double *)malloc(sizeof(double) * (nr + 1) * 7);
  fitref_spline(nrho, drho, frho, s->frho_sp);
  fitref_spline(nr, dr, rhor, s->rhor_sp);
  fitref_spline(nr, dr, z2r, s->z2r_sp);
  s->eam_set = 1;
  return 0;
}

/* r-spline interval and fraction (pair_eam.cpp:170-174) */
static const double *r_coeff(fitref *s, const double *sp, double r, double *p_out, int *m_out) {
  double p = r * s->rdr + 1.0;
  int m = (int)p;
  m = MIN(m, s->nr - 1);
  p -= m;
  p = MIN(p, 1.0);
  *p_out = p;
  *m_out = m;
  return &sp[m * 7];
}

/* rho-spline interval and fraction (pair_eam.cpp:215-219); *clamped when p was cut to 1 */
static const double *rho_coeff(fitref *s, double rho, double *p_out, int *m_out, int *clamped) {
  double p = rho * s->rdrho + 1.0;
  int m = (int)p;
  m = MAX(1, MIN(m, s->nrho - 1));
  p -= m;
  *clamped = p > 1.0;
  p = MIN(p, 1.0);
  *p_out = p;
  *m_out = m;
  return &s->frho_sp[m * 7];
}

static int eam_pass(fitref *s, double *f, double *energy, double *rho_out, double *fp_out, const double *gf,
                    double *gc_frho, double *gc_rhor, double *gc_z2r) {
  if (!s->eam_set) return fail(s, "eam tables not set");
  const double cut = sqrt(s->cutsq);
  for (int r = 0; r < s->nrep; r++) {
    plist pl;
    memset(&pl, 0, sizeof(pl));
    if (build_list(s, r, cut, &pl)) return -1;
    const int a0 = s->off[r], n = s->natoms[r];
    double *rho = (double *)calloc(n + 1, sizeof(double)), *fp = (double *)calloc(n + 1, sizeof(double));
    double e = 0.0;
    /* rho pass (pair_eam.cpp:160-202, full list: own row only) */
    for (int k = 0; k < n; k++) {
      const int i = a0 + k;
      for (int q = pl.first[k]; q < pl.first[k] + pl.num[k]; q++) {
        double del[3];
        const double rsq = pair_del(s, r, i, a0 + pl.j[q], pl.shift[q], del);
        if (rsq < s->cutsq) {
          double p;
          int m;
          const double *c = r_coeff(s, s->rhor_sp, sqrt(rsq), &p, &m);
          rho[k] += ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
        }
      }
    }
    /* fp pass (pair_eam.cpp:213-228) */
    for (int k = 0; k < n; k++) {
      double p;
      int m, cl;
      const double *c = rho_coeff(s, rho[k], &p, &m, &cl);
      fp[k] = (c[0] * p + c[1]) * p + c[2];
      double phi = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
      if (rho[k] > s->rhomax) phi += fp[k] * (rho[k] - s->rhomax);
      e += phi;
    }
    if (!gf) {
      /* force pass (pair_eam.cpp:237-303) */
      for (int k = 0; k < n; k++) {
        const int i = a0 + k;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int q = pl.first[k]; q < pl.first[k] + pl.num[k]; q++) {
          const int jj = pl.j[q];
          double del[3];
          const double rsq = pair_del(s, r, i, a0 + jj, pl.shift[q], del);
          if (rsq < s->cutsq) {
            const double rr = sqrt(rsq);
            double p;
            int m;
            const double *c = r_coeff(s, s->rhor_sp, rr, &p, &m);
            const double rhoip = (c[0] * p + c[1]) * p + c[2];
            const double rhojp = rhoip;
            c = &s->z2r_sp[m * 7];
            const double z2p = (c[0] * p + c[1]) * p + c[2];
            const double z2 = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
            const double recip = 1.0 / rr;
            const double phi = z2 * recip;
            const double phip = z2p * recip - phi * recip;
            const double psip = fp[k] * rhojp + fp[jj] * rhoip + phip;
            const double fpair = -psip * recip;
            fx += del[0] * fpair;
            fy += del[1] * fpair;
            fz += del[2] * fpair;
            e += 0.5 * phi;
          }
        }
        f[3 * i] = fx;
        f[3 * i + 1] = fy;
        f[3 * i + 2] = fz;
        if (rho_out) rho_out[i] = rho[k];
        if (fp_out) fp_out[i] = fp[k];
      }
      if (energy) energy[r] = e;
    } else {
      /* adjoint.  With b = -(G_i . del)/r per row entry:
         dL/dz2'(r) += b/r, dL/dz2(r) += -b/r^2, dL/drho'(r) += b (fp_i + fp_j),
         dL/dfp_i += b rho'(r), dL/dfp_j += b rho'(r);
         then dL/dF'(rho_i) = u_i and dL/drho_i = u_i F''(rho_i), which
         reaches rho(r) through the rho pass. */
      double *u = (double *)calloc(n + 1, sizeof(double));
      for (int k = 0; k < n; k++) {
        const int i = a0 + k;
        for (int q = pl.first[k]; q < pl.first[k] + pl.num[k]; q++) {
          const int jj = pl.j[q];
          double del[3];
          const double rsq = pair_del(s, r, i, a0 + jj, pl.shift[q], del);
          if (rsq < s->cutsq) {
            const double rr = sqrt(rsq);
            double p;
            int m;
            const double *c = r_coeff(s, s->rhor_sp, rr, &p, &m);
            const double rhop = (c[0] * p + c[1]) * p + c[2];
            const double b = -(gf[3 * i] * del[0] + gf[3 * i + 1] * del[1] + gf[3 * i + 2] * del[2]) / rr;
            /* value weights (p^3, p^2, p, 1) and derivative weights (3p^2, 2p, 1, 0)/dr on c3..c6 */
            const double vw[4] = {p * p * p, p * p, p, 1.0};
            const double dw[4] = {3.0 * p * p * s->rdr, 2.0 * p * s->rdr, s->rdr, 0.0};
            const double gz2p = b / rr, gz2 = -b / (rr * rr), grhop = b * (fp[k] + fp[jj]);
            for (int t = 0; t < 4; t++) {
              gc_z2r[m * 4 + t] += gz2p * dw[t] + gz2 * vw[t];
              gc_rhor[m * 4 + t] += grhop * dw[t];
            }
            u[k] += b * rhop;
            u[jj] += b * rhop;
          }
        }
      }
      double *v = (double *)calloc(n + 1, sizeof(double));
      for (int k = 0; k < n; k++) {
        double p;
        int m, cl;
        const double *c = rho_coeff(s, rho[k], &p, &m, &cl);
        const double dw[4] = {3.0 * p * p * s->rdrho, 2.0 * p * s->rdrho, s->rdrho, 0.0};
        for (int t = 0; t < 4; t++) gc_frho[m * 4 + t] += u[k] * dw[t];
        const double fpp = cl ? 0.0 : (2.0 * c[0] * p + c[1]) * s->rdrho;
        v[k] = u[k] * fpp;
      }
      for (int k = 0; k < n; k++) {
        const int i = a0 + k;
        if (v[k] == 0.0) continue;
        for (int q = pl.first[k]; q < pl.first[k] + pl.num[k]; q++) {
          double del[3];
          const double rsq = pair_del(s, r, i, a0 + pl.j[q], pl.shift[q], del);
          if (rsq < s->cutsq) {
            double p;
            int m;
            r_coeff(s, s->rhor_sp, sqrt(rsq), &p, &m);
            const double vw[4] = {p * p * p, p * p, p, 1.0};
            for (int t = 0; t < 4; t++) gc_rhor[m * 4 + t] += v[k] * vw[t];
          }
        }
      }
      free(u);
      free(v);
    }
    free(rho);
    free(fp);
    plist_free(&pl);
  }
  return 0;
}

int fitref_eam_forces(fitref *s, double *f, double *energy, double *rho, double *fp) {
  return eam_pass(s, f, energy, rho, fp, NULL, NULL, NULL, NULL);
}

/* knot gradient = J^T (coefficient gradient), J = the linear map of fitref_spline
   restricted to c3..c6 (c0..c2 enter through the derivative weights) */
static void knot_grad(int n, double delta, const double *gc /* (n+1)*4 */, double *g) {
  double *e = (double *)calloc(n, sizeof(double)), *sp = (double *)malloc(sizeof(double) * (n + 1) * 7);
  for (int k = 0; k < n; k++) {
    e[k] = 1.0;
    fitref_spline(n, delta, e, sp);
    double acc = 0.0;
    for (int m = 1; m <= n; m++)
      for (int t = 0; t < 4; t++) acc += gc[m * 4 + t] * sp[m * 7 + 3 + t];
    g[k] = acc;
    e[k] = 0.0;
  }
  free(e);
  free(sp);
}

int fitref_eam_vjp(fitref *s, const double *gf, double *g_frho, double *g_rhor, double *g_z2r) {
  if (!s->eam_set) return fail(s, "eam tables not set");
  double *gf_c = (double *)calloc((s->nrho + 1) * 4, sizeof(double));
  double *gr_c = (double *)calloc((s->nr + 1) * 4, sizeof(double));
  double *gz_c = (double *)calloc((s->nr + 1) * 4, sizeof(double));
  int rc = eam_pass(s, NULL, NULL, NULL, NULL, gf, gf_c, gr_c, gz_c);
  if (!rc) {
    knot_grad(s->nrho, s->drho, gf_c, g_frho);
    knot_grad(s->nr, s->dr, gr_c, g_rhor);
    knot_grad(s->nr, s->dr, gz_c, g_z2r);
  }
  free(gf_c); free(gr_c); free(gz_c);
  return rc;
}
