/* mdx_batch.h - C ABI for batched force evaluation over many independent
 * configurations ("replicas": snapshots of one system, or clusters of
 * uneven size) and for the adjoint of a force-matching loss.
 *
 * The reference has no batched or differentiable entry point: its force
 * styles are PairLJCut::compute (src/pair_lj_cut.cpp:68-141) and
 * PairEAM::compute (src/MANYBODY/pair_eam.cpp:125-306) on one system, with
 * tables from PairLJCut::init_one (:545-602) and PairEAM::interpolate
 * (pair_eam.cpp:723-751).  These entry points evaluate exactly those forces
 * for every replica in one persistent launch per pass, and return the
 * gradient of L with respect to the potential parameters given
 * G = dL/dF (a vector-Jacobian product), recomputing pair distances from
 * the positions instead of storing per-pair state.
 *
 * Conventions as in mdx.h: int status (MDX_OK or MDX_ERR_*), mdx_last_error()
 * for the message, calls synchronous at return.  Per-atom arrays are
 * [N][3] doubles in the caller's order (replica 0's atoms, then replica 1's,
 * ...).  Pointers marked "host or device" may point to host memory or to
 * device memory on the batch's device (the library checks which).
 */
#ifndef MDX_BATCH_H
#define MDX_BATCH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdx_batch mdx_batch;

/* kernel kinds for tuning and timing */
enum {
  MDX_BK_BUILD = 0,     /* neighbour build over all replicas */
  MDX_BK_LJ = 1,        /* lj/cut forces + per-atom energy */
  MDX_BK_LJ_VJP = 2,    /* lj/cut dL/d(eps, sigma) */
  MDX_BK_EAM_RHO = 3,   /* eam density + embedding (fp) pass */
  MDX_BK_EAM_FORCE = 4, /* eam force pass */
  MDX_BK_EAM_ADJ_A = 5, /* eam adjoint: z2r/rhor' terms, dL/dfp, frho terms */
  MDX_BK_EAM_ADJ_B = 6, /* eam adjoint: rhor terms through rho */
  MDX_BK_NKIND = 8
};

/* nrep replicas; replica r holds natoms[r] atoms of `type` [N] (1-based,
   ntypes <= 8), box lo/hi [nrep][3], periodic [nrep][3] (0: boundary f,
   atoms must then lie inside [lo, hi]). */
int mdx_batch_create(int device, int nrep, const int *natoms, const double *lo, const double *hi,
                     const int *periodic, int ntypes, const int *type, mdx_batch **out);
int mdx_batch_destroy(mdx_batch *b);

/* Stream of the batch's work: external = 0 its own stream (the default),
   1 the caller's `stream` (a hipStream_t of the batch's device; NULL is that
   device's null stream), so the kernels queue behind the caller's work on
   that stream (e.g. PyTorch's current stream) with no device-wide
   synchronisation.  Calls still return when their work is done. */
int mdx_batch_set_stream(mdx_batch *b, int external, void *stream);

/* positions [N][3], host or device; wrapped into periodic boxes */
int mdx_batch_set_positions(mdx_batch *b, const double *x);
/* MDX_FP64 or MDX_MIXED (FP64 positions and distances, FP32 potential
   arithmetic, FP64 accumulation) for the force passes; the adjoint passes
   run FP64 arithmetic in both modes (their per-pair factors cancel near the
   potential minimum, where FP32 would cost ~5e-6 relative gradient error) */
int mdx_batch_set_precision(mdx_batch *b, int precision);
/* neighbour skin added to the largest pair cutoff (default 0) */
int mdx_batch_set_skin(mdx_batch *b, double skin);

/* lj/cut parameters, (ntypes+1)^2 row-major symmetric tables of epsilon,
   sigma and cutoff (pair_coeff i j eps sigma cut; no energy shift). */
int mdx_batch_lj_init(mdx_batch *b, const double *eps, const double *sigma, const double *cut);
/* f [N][3] host or device (may be NULL); energy [nrep] host (may be NULL) */
int mdx_batch_lj_forces(mdx_batch *b, double *f, double *energy);
/* gf = dL/dF [N][3] host or device; geps, gsigma (ntypes+1)^2 host:
   dL/d(eps_ab), dL/d(sigma_ab) summed into [a][b] with a <= b. */
int mdx_batch_lj_vjp(mdx_batch *b, const double *gf, double *geps, double *gsigma);

/* eam, one element for every type, from knot values: frho[nrho] at
   rho = (m-1) drho, rhor[nr] and z2r[nr] at r = (m-1) dr (the arrays
   PairEAM::array2spline turns into splines); force cutoff `cut`;
   rhomax = (nrho-1) drho (pair_eam.cpp:213-228 semantics beyond it). */
int mdx_batch_eam_init(mdx_batch *b, int nrho, double drho, int nr, double dr, double cut, const double *frho,
                       const double *rhor, const double *z2r);
/* f [N][3] host or device (may be NULL); energy [nrep] host; rho, fp [N]
   host (may be NULL) */
int mdx_batch_eam_forces(mdx_batch *b, double *f, double *energy, double *rho, double *fp);
/* gf [N][3] host or device; knot gradients g_frho[nrho], g_rhor[nr],
   g_z2r[nr] host.  Reuses the rho/fp of the last FP64 mdx_batch_eam_forces
   on the current positions and tables, and otherwise reruns the rho pass
   in FP64 first. */
int mdx_batch_eam_vjp(mdx_batch *b, const double *gf, double *g_frho, double *g_rhor, double *g_z2r);

/* Launch shape per kernel kind: block (64..1024, multiple of 64), threads
   per atom (1..64, power of two), grid (blocks; 0 = one block per
   atom tile).  Kernels loop over their share of atoms, so any grid is valid. */
int mdx_batch_set_launch(mdx_batch *b, int kind, int block, int tpa, int grid);
int mdx_batch_get_launch(mdx_batch *b, int kind, int *block, int *tpa, int *grid);
/* Autotuner: when on, the next call of each kernel kind times a set of
   (threads per atom, block, grid) candidates on the real inputs, keeps the
   fastest and finishes with it. */
int mdx_batch_autotune(mdx_batch *b, int on);
/* accumulated device time (ms) and launch count of a kernel kind */
int mdx_batch_kernel_time(mdx_batch *b, int kind, double *ms, long *launches);
/* neighbour statistics of the current list */
int mdx_batch_neigh_stats(mdx_batch *b, long *total, int *max_per_atom);

#ifdef __cplusplus
}
#endif
#endif
