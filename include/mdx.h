/* mdx.h - C ABI of the MI355X (gfx950) short-range MD hot path.
 *
 * The library replaces, behind the reference's own plugin surface, the work
 * LAMMPS does in Neighbor::build (binned lists), PairLJCut::compute,
 * PairEAM::compute (rho -> fp -> force) and FixNVE::initial/final_integrate.
 * Reference citations are relative to the LAMMPS "8 Aug 2014" source tree.
 *
 * Conventions
 *  - Every entry point returns an int status: MDX_OK (0) or a negative
 *    MDX_ERR_* code; mdx_last_error() returns the message of the last failure
 *    on the calling thread.  (The reference's GPU package reports failure the
 *    same way through an int return of *_gpu_init and a bool& success of
 *    *_gpu_compute: src/GPU/pair_lj_cut_gpu.cpp:42-65.)
 *  - Host arrays are borrowed for the duration of a call only.  Per-atom
 *    arrays are AoS [n][3] doubles laid out as LAMMPS's atom->x[0]
 *    (src/memory.h 2-d create), in the caller's atom order.
 *  - Per-type tables are (ntypes+1)^2 row-major doubles, indexed [i][j] with
 *    1-based types, exactly as Pair::init_one fills them.
 *  - A context owns all device memory; one context per host thread.
 *    Calls are synchronous at return.
 *  - Plain pointers and sizes only; no HIP or torch types cross the ABI.
 */
#ifndef MDX_H
#define MDX_H

#ifdef __cplusplus
extern "C" {
#endif

#define MDX_VERSION 1

enum {
  MDX_OK = 0,
  MDX_ERR_ARG = -1,      /* invalid argument or state */
  MDX_ERR_HIP = -2,      /* HIP runtime failure (no device, launch failure) */
  MDX_ERR_NOMEM = -3,    /* device allocation failed ("Insufficient memory on accelerator") */
  MDX_ERR_OVERFLOW = -4  /* an atom has more than neigh_modify `one` neighbours */
};

enum { MDX_FP64 = 0, MDX_MIXED = 1 };        /* pair arithmetic precision */
enum { MDX_PAIR_NONE = 0, MDX_PAIR_LJ = 1, MDX_PAIR_EAM = 2 };
/* neighbour positions read by the mixed-precision gather kernels: 32-bit
   fixed point (16 B, the default) or the FP64 positions (32 B) */
enum { MDX_POS_FP64 = 0, MDX_POS_Q32 = 1 };

typedef struct mdx_ctx mdx_ctx;

/* Thermo callback: step, then {temp-independent raw sums}:
   row[0] = sum_i m_i v_i^2 (mass*velocity^2, no unit factor)
   row[1] = pair energy (eng_vdwl), row[2..7] = pair virial xx yy zz xy xz yz */
typedef void (*mdx_thermo_fn)(void *user, long step, const double *row);

/* kernels timed when profiling is on (index into mdx_stats.kernel_ms) */
enum {
  MDX_K_BUILD = 0,       /* whole list build: wrap/bin, sort, permute, list */
  MDX_K_LJ = 1,          /* lj/cut force kernel */
  MDX_K_EAM_RHO = 2,     /* eam density + embedding derivative kernel */
  MDX_K_EAM_FORCE = 3,   /* eam force kernel */
  MDX_K_NVE_INITIAL = 4, /* fix nve initial integrate (+ rebuild trigger) */
  MDX_K_NVE_FINAL = 5,   /* fix nve final integrate */
  MDX_K_FUSED = 6,       /* force + final + next initial integrate, fused */
  MDX_K_HALO = 7,        /* ghost-shell pack / unpack (mdx_halo_pack, mdx_halo_unpack) */
  MDX_NKERNEL = 8
};

typedef struct {
  long ncalls;          /* neighbour builds during mdx_run (Neighbor::ncalls) */
  long ndanger;         /* dangerous builds (neighbor.cpp:1370) */
  long total_neighbors; /* full-list entries at the last build (= 2x half-list pairs) */
  int max_neighbors;    /* largest per-atom count at the last build */
  int capacity;         /* per-atom list capacity currently allocated */
  int nbins[3];         /* device bin grid */
  double kernel_ms[MDX_NKERNEL]; /* accumulated device time per kernel kind (ms) */
  long kernel_launches[MDX_NKERNEL];
} mdx_stats;

const char *mdx_last_error(void);
int mdx_version(void);

/* Creates a context on HIP device `device` (the accelerator set-up that
   `package gpu` performs in src/GPU/fix_gpu.cpp:77-150). */
int mdx_create(int device, mdx_ctx **out);
int mdx_destroy(mdx_ctx *ctx);

/* Orthogonal box and periodicity (Domain boxlo/boxhi/periodicity,
   src/domain.h).  Non-periodic dimensions behave as `boundary f`. */
int mdx_set_box(mdx_ctx *ctx, const double lo[3], const double hi[3], const int periodic[3]);

/* Atoms: n owned atoms with x, v ([n][3]), 1-based type, tag (atom ID), and
   per-type mass[ntypes+1] (src/atom.h x/v/type/tag/mass).  v may be NULL. */
int mdx_set_atoms(mdx_ctx *ctx, int n, int ntypes, const double *x, const double *v, const int *type,
                  const int *tag, const double *mass);
/* Per-step position update in caller order: the role of the H2D copy that
   ljl_gpu_compute performs on atom->x each step (src/GPU/pair_lj_cut_gpu.cpp:87-122). */
int mdx_set_positions(mdx_ctx *ctx, const double *x);
/* Copies back state in caller order; any pointer may be NULL.  image is
   3 ints per atom (periodic image counts, src/domain.cpp:478 pbc). */
int mdx_get_atoms(mdx_ctx *ctx, double *x, double *v, double *f, int *image);

/* lj/cut coefficients as PairLJCut::init_one leaves them
   (src/pair_lj_cut.cpp:545-602; replaces ljl_gpu_init, src/GPU/pair_lj_cut_gpu.cpp:42-48). */
int mdx_pair_lj_init(mdx_ctx *ctx, int ntypes, const double *cutsq, const double *lj1, const double *lj2,
                     const double *lj3, const double *lj4, const double *offset);

/* eam spline tables as PairEAM::array2spline leaves them
   (src/MANYBODY/pair_eam.cpp:698-751; members pair_eam.h:37-45):
   frho_spline[nfrho][nrho+1][7], rhor_spline[nrhor][nr+1][7], z2r_spline[nz2r][nr+1][7],
   type2frho[ntypes+1], type2rhor/type2z2r[(ntypes+1)^2].
   Replaces eam_gpu_init (src/GPU/pair_eam_gpu.cpp:38-47). */
int mdx_pair_eam_init(mdx_ctx *ctx, int ntypes, int nrho, double drho, int nr, double dr, double rhomax,
                      double cutforcesq, int nfrho, const double *frho_spline, int nrhor,
                      const double *rhor_spline, int nz2r, const double *z2r_spline, const int *type2frho,
                      const int *type2rhor, const int *type2z2r);

/* FP64, or mixed: FP64 positions and distances, FP32 potential arithmetic,
   FP64 force/energy/virial accumulation. */
int mdx_set_precision(mdx_ctx *ctx, int precision);

/* `neighbor <skin> bin` + `neigh_modify every N delay N check yes|no one N`
   (src/neighbor.cpp:1731-1862). */
int mdx_neigh_modify(mdx_ctx *ctx, double skin, int every, int delay, int check, int one);

/* Neighbor::build on the current positions (wrap into the box, bin-sort,
   list build; src/neighbor.cpp:1380-1461, src/neigh_half_bin.cpp:285-419). */
int mdx_neigh_build(mdx_ctx *ctx);
/* Pair set of the current list as (tag_i, tag_j) with tag_i < tag_j, each
   pair once (the half-list pair set), in no particular order.  *npairs gets
   the count; pairs may be NULL to query it. */
int mdx_neigh_pairs(mdx_ctx *ctx, int *pairs, long cap, long *npairs);

/* Pair::compute(eflag, vflag) on the current positions (src/pair.h:135;
   eflag/vflag encoding of src/integrate.cpp:99-112: eflag 1 global, 2 per-atom;
   vflag 1|2 global, 4 per-atom).  Builds the list first if none exists.
   Outputs are optional: f [n][3] (caller order) receives f += the pair
   forces, as atom->f accumulates over the force styles of a step
   (src/pair.h:33-35, src/pair_lj_cut.cpp:115-117); eng_vdwl, virial[6],
   eatom[n], vatom[n][6] are overwritten, as ev_setup zeroes them at every
   compute (src/pair.cpp:706-725).  Replaces ljl_gpu_compute /
   eam_gpu_compute + eam_gpu_compute_force + lmp_gpu_forces
   (src/GPU/pair_lj_cut_gpu.cpp:87-122, src/GPU/pair_eam_gpu.cpp:96-140,
   src/GPU/fix_gpu.cpp:236-257). */
int mdx_compute(mdx_ctx *ctx, int eflag, int vflag, double *f, double *eng_vdwl, double *virial,
                double *eatom, double *vatom);

/* Limits of mdx_compute_noalloc (a file-scope device array of
   MDX_NOALLOC_SCRATCH_BYTES per GPU holds its work arrays). */
#define MDX_NOALLOC_MAX_ATOMS 32768
#define MDX_NOALLOC_NEIGH_PER_ATOM 128 /* list entries per atom at the maximum atom count */
#define MDX_NOALLOC_MAX_BINS (1 << 19)
#define MDX_NOALLOC_SCRATCH_BYTES (36L << 20)

/* mdx_compute without device allocation.  No counterpart in the reference:
   for callers that cannot allocate device memory once set up (serving
   loops, harnesses that forbid hipMalloc after start-up).  One
   Pair::compute on n atoms whose positions x ([n][3] doubles) and types
   (int[n], 1-based) are DEVICE memory of the context's GPU, in an order the
   caller chooses; f ([n][3] doubles, device) receives f += the pair forces
   in that order, eng_vdwl and virial[6] (host, optional) the global sums
   when eflag/vflag ask for them (per-atom outputs are not offered).  The
   context supplies the box, pair style and tables, precision, neighbour
   skin and `one`, and the launch shape; its own atoms, if any, are neither
   read nor changed.  The list is built from scratch at every call (skin
   included, as mdx_compute's first build), in a file-scope device array
   of MDX_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends,
   so n <= MDX_NOALLOC_MAX_ATOMS, at most MDX_NOALLOC_MAX_BINS bins, and at
   most MDX_NOALLOC_MAX_ATOMS * MDX_NOALLOC_NEIGH_PER_ATOM / n (and `one`)
   neighbours per atom; beyond these the call returns MDX_ERR_ARG or
   MDX_ERR_OVERFLOW without computing.  The call allocates nothing, is
   synchronous at return, and every call on one GPU shares that GPU's
   array: calls that overlap on one GPU (from several threads or contexts)
   race. */
int mdx_compute_noalloc(mdx_ctx *ctx, int n, const double *x, const int *type, int eflag, int vflag, double *f,
                        double *eng_vdwl, double *virial);

/* EAM intermediate arrays after the last compute (rho, fp), caller order. */
int mdx_eam_rho_fp(mdx_ctx *ctx, double *rho, double *fp);

/* Verlet::setup (src/verlet.cpp:88-142): wrap, sort and build the list,
   compute forces with energy and virial, and report the current step to the
   callback.  Resets the build counters, as setup does (neighbor->ncalls = 0). */
int mdx_setup(mdx_ctx *ctx, mdx_thermo_fn cb, void *user);

/* Verlet::run with fix nve, fully device resident
   (src/verlet.cpp:88-142,207-309; src/fix_nve.cpp:65-142;
   src/neighbor.cpp:1294-1372).  Step numbering continues across calls.
   dt and ftm2v as Update/Force hold them (src/update.cpp:118-180).
   Calls mdx_setup first if it has not run since the state last changed.
   The callback (may be NULL) receives every step that is a multiple of
   thermo_every (0: none) and the last step. */
int mdx_run(mdx_ctx *ctx, int nsteps, double dt, double ftm2v, int thermo_every, mdx_thermo_fn cb,
            void *user);

int mdx_get_stats(mdx_ctx *ctx, mdx_stats *out);

/* mdx_run fuses fix nve into the force kernel by default (final integrate of
   step n and initial integrate of step n+1 run inside step n's force
   kernel; same arithmetic, bit-identical trajectory).  0 runs the three
   kernels separately, as FixNVE::initial_integrate / Pair::compute /
   FixNVE::final_integrate are separate calls in the reference. */
int mdx_set_fused(mdx_ctx *ctx, int on);

/* Times every list build, pair launch and integrator launch with HIP event
   pairs on the context's stream (no synchronisation is added; the events
   are resolved by mdx_get_stats).  Resets the accumulated times. */
int mdx_set_profiling(mdx_ctx *ctx, int on);

/* HBM calibration probe (no counterpart in the reference): copies nbytes
   device-to-device `iters` times with a streaming kernel (k_copy) on the
   context's stream and returns the achieved read+write rate in GB/s. */
int mdx_copy_bandwidth(mdx_ctx *ctx, long nbytes, int iters, double *gbs);

/* Kernel shape of the pair kernels, set by the autotuner (0 = library
   default): block = threads per workgroup (64..256, multiple of 64), tpa =
   threads per atom (1, 2, 4, 8, 16 or 32; the `threads_per_atom` keyword of
   the reference's `package gpu`, src/GPU/fix_gpu.cpp:112-150). */
int mdx_set_launch(mdx_ctx *ctx, int block, int tpa);
int mdx_get_launch(mdx_ctx *ctx, int *block, int *tpa);
/* Autotuner (the role of `package gpu ... tpa` chosen by hand in the
   reference, src/GPU/fix_gpu.cpp:112-150): times the force kernel of the
   current configuration for threads per atom {4, 8, 16} x block {128, 256},
   `reps` launches each; keeps the fastest and recomputes the forces with it. */
int mdx_autotune(mdx_ctx *ctx, int reps);
/* Timing probe (no counterpart in the reference): the force kernel of a
   non-thermo step of mdx_run (with fix nve fused: force + final integrate +
   next initial integrate, the same reads and writes), launched `reps` times
   back to back between two HIP events on the context's stream, into spare
   buffers (velocities and positions are left as they were); *ms is the
   mean duration of one launch.  The list must be built (mdx_setup). */
int mdx_time_step_kernel(mdx_ctx *ctx, int reps, double dt, double ftm2v, double *ms);
/* Position format of the mixed-precision gathers and of the list build's
   candidate test.
   MDX_POS_Q32 (the default): each neighbour is read as 32-bit fixed point
   per dimension (resolution prd * 2^-32 periodic, prd * 2^-30 otherwise)
   and its separation rebuilt in FP64; pairs within 2e-6 relative of the
   cutoff are re-tested on the FP64 positions, so the cutoff decision is the
   FP64 one.  The list build tests its candidates the same way in either
   precision (same pairs as an FP64 test).  FP64 precision's pair kernels
   always read FP64 positions.
   MDX_POS_FP64: neighbours and build candidates are read as FP64
   positions. */
int mdx_set_position_format(mdx_ctx *ctx, int format);

/* Ghost-shell exchange of a spatially decomposed run (one context per GPU
   sub-box; the packed buffers travel between GPUs by RCCL send/recv, which
   the caller issues).  Replaces CommBrick's per-step forward_comm of ghost
   positions (src/comm_brick.cpp:398-460) with the send lists that borders
   builds at each reneighbouring (src/comm_brick.cpp:696-864), and
   AtomVecAtomic::pack_comm / unpack_comm with the pbc shift
   (src/atom_vec_atomic.cpp:117-190).

   mdx_halo_sendlist: swap in [0, 27) selects the atoms of this context with
   lo <= x < hi (each dimension), in device order, and records `shift` (the
   image shift prd * pbc a ghost of them carries).  *count gets the number
   selected.  The list stays valid until the next list build of this
   context, which reorders the atoms; rebuild it after each one, as borders
   does (mdx_halo_pack returns MDX_ERR_ARG for a list older than the last
   build).
   mdx_halo_pack: writes the listed atoms' x + shift, 3 doubles each, into
   buf (device memory of this context's GPU).  swap = -1 packs every swap
   in one launch, swap 0's atoms first, then swap 1's, ... (each swap's rows
   start at the sum of the earlier swaps' counts).  Pack and unpack run on
   the context's stream and return when done; work that the caller queued
   on other streams for the same buffers must be complete before the call.
   mdx_halo_unpack: writes count received positions (3 doubles each, device
   memory) into this context's atoms first .. first+count-1 (caller order):
   the ghost atoms, which the caller appended after the owned ones.  Ghost
   atoms are ordinary atoms of the context to the list build and the force
   kernels (mdx_compute), and their forces are not used.  mdx_run
   integrates every atom of a context, so the step loop of a sharded run
   (owned-only integration, ghost refresh between the initial integrate and
   the force kernel) is not part of this library. */
int mdx_halo_sendlist(mdx_ctx *ctx, int swap, const double lo[3], const double hi[3], const double shift[3],
                      int *count);
/* Send lists 0 .. nswap-1 in one pass over the atoms (regions may overlap):
   lo, hi, shift are [nswap][3], counts[nswap] gets each list's size.  The
   lists equal what nswap calls of mdx_halo_sendlist would build; lists
   nswap .. 26 of earlier calls are dropped. */
int mdx_halo_sendlists(mdx_ctx *ctx, int nswap, const double *lo, const double *hi, const double *shift,
                       int *counts);
int mdx_halo_pack(mdx_ctx *ctx, int swap, double *buf);
int mdx_halo_unpack(mdx_ctx *ctx, int first, int count, const double *buf);

#ifdef __cplusplus
}
#endif
#endif
