// mdx_internal.h - device data layout and kernel launchers shared by the
// HIP sources of libmdx (gfx950 only).
//
// Device layout (one context, atoms kept in bin-sorted order, re-sorted at
// every neighbour build):
//   x      double4[npad]  x, y, z, w = (double)type   -- 32 B gathers
//   v      double4[npad]  vx, vy, vz, w unused
//   f      double4[npad]  fx, fy, fz, w = per-atom energy of the last compute
//   tag    int[npad]      atom ID;  orig int[npad]: caller index of each atom
//   image  int[npad]      periodic images, 10 bits per dimension (IMGMAX 512)
//   nbr    int4[maxneigh/4][npad]  full list in chunks of four entries,
//                                chunk-major: chunk c of atom i at c * npad +
//                                i, so one chunk of consecutive atoms is
//                                contiguous (a wavefront's index loads are
//                                whole lines, none kept for a later trip);
//                                bits 0-26 of an entry hold the neighbour
//                                index, bits 27-31 a periodic image code
//                                (sx+1) + 3(sy+1) + 9(sz+1)
//   vatom  double[6][npad] per-atom virial (SoA)
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mdx.h"

#define MDX_MAXT 8
#define MDX_NT1 (MDX_MAXT + 1)
#define MDX_IDX_BITS 27
#define MDX_IDX_MASK ((1 << MDX_IDX_BITS) - 1)
#define MDX_HALO_SWAPS 27  // ghost-shell send lists per context (26 neighbour regions of a sub-box, + 1)
#define MDX_NOSHIFT 13
#define MDX_IMGMAX 512
#define MDX_IMGMASK 1023
#define MDX_RED_SLOTS 8  // reduction slots: e, v0..v5, ke

namespace mdx {

// Bin geometry of the reference's Neighbor::setup_bins (neighbor.cpp:1538-1702),
// used only to decide which side of a periodic pair the reference would
// have evaluated the distance from (matters only at exact cutoff ties).
struct RefBins {
  double bboxlo[3], bboxhi[3], bininv[3];
  int nbin[3];
};

struct NeighArgs {
  double lo[3], prd[3], bininv[3];
  int nbin[3];
  int S[3];
  int periodic[3];
  int n, npad, maxneigh, ntypes1;
  double cutneighsq[MDX_NT1 * MDX_NT1];
  RefBins ref;
};

struct LJTab {  // (ntypes+1)^2 row-major, as Pair::init_one fills them
  double cutsq[MDX_NT1 * MDX_NT1], lj1[MDX_NT1 * MDX_NT1], lj2[MDX_NT1 * MDX_NT1],
      lj3[MDX_NT1 * MDX_NT1], lj4[MDX_NT1 * MDX_NT1], offset[MDX_NT1 * MDX_NT1];
};

struct EAMArgs {
  int nrho, nr, nfrho, nrhor, nz2r, ntypes1;
  double rdr, rdrho, rhomax, cutforcesq;
  double clo, chi;  // cutforcesq (1 -/+ the fixed-point re-test band), set at each launch
  int type2frho[MDX_NT1];
  int type2rhor[MDX_NT1 * MDX_NT1];
  int type2z2r[MDX_NT1 * MDX_NT1];
  const double *frho, *rhor, *z2r;  // [k][n+1][7]
  const float *frho_f, *rhor_f, *z2r_f;
  // rhor and z2r repacked per knot as two 4-vectors {c0,c1,c2,0},{c3,c4,c5,c6}
  // (derivative, value): element 2*(k*(nr+1)+m) and the next one
  const double4 *rhor_p, *z2r_p;
  const float4 *rhor_pf, *z2r_pf;
};

// ---- eam spline evaluation of the pair kernels (every variant of the
// density and force passes uses these, so a pair gives the same bits in
// each).  The spline
// index arithmetic is PairEAM::compute's (pair_eam.cpp:188-192); tables are
// the per-knot packed rhor/z2r of EAMArgs (MIXED: FP32 coefficients).
template <typename T>
__device__ __forceinline__ void eam_spline_index(T r, T rdr, int nr, int &m, T &p) {
  p = r * rdr + T(1);
  m = static_cast<int>(p);
  m = m < 1 ? 1 : (m > nr - 1 ? nr - 1 : m);  // m >= 1 for any r >= 0; the lower clamp only guards NaN
  p -= m;
  p = p < T(1) ? p : T(1);
}

// Mixed precision: r, the knot index m and the fraction p from the FP64
// rsq.  r is the FP32 reciprocal square root refined by one FP64 Newton
// step (relative error ~1e-14; sqrtf of the FP32-rounded rsq is ~1e-7), the
// index arithmetic of pair_eam.cpp:188-192 runs in FP64, and only r and the
// fraction are rounded to FP32 for the spline arithmetic.  With r and p in
// FP32 the spline position carries ~2e-7 relative error, which the steep
// pair terms turn into force errors of ~1e-5 of the RMS force on large
// systems; here the FP32 coefficients and arithmetic bound it (~2e-6).
// Returns r in FP32.
__device__ __forceinline__ float eam_index_mixed(double rsq, double rdr, int nr, int &m, float &p) {
  const double y = (double)__builtin_amdgcn_rsqf((float)rsq);
  const double r0 = rsq * y;
  const double r = fma(0.5 * y, fma(-r0, r0, rsq), r0);
  double pp = fma(r, rdr, 1.0);
  m = static_cast<int>(pp);
  m = m < 1 ? 1 : (m > nr - 1 ? nr - 1 : m);
  pp -= m;
  p = (float)(pp < 1.0 ? pp : 1.0);
  return (float)r;
}

// FP32 rho spline value from the knot's c3..c6 vector at fraction p
__device__ __forceinline__ float eam_rho_eval(const float4 &c, float p) { return ((c.x * p + c.y) * p + c.z) * p + c.w; }

// rho contribution of a neighbour at rsq through table `tab` (pair_eam.cpp:193-196)
template <bool MIXED>
__device__ __forceinline__ double eam_rho_term(const EAMArgs &a, int tab, double rsq) {
  int m;
  if (MIXED) {
    float p;
    (void)eam_index_mixed(rsq, a.rdr, a.nr, m, p);
    return (double)eam_rho_eval(a.rhor_pf[2 * ((size_t)tab * (a.nr + 1) + m) + 1], p);  // c3..c6
  }
  double p;
  eam_spline_index<double>(sqrt(rsq), a.rdr, a.nr, m, p);
  const double4 c = a.rhor_p[2 * ((size_t)tab * (a.nr + 1) + m) + 1];
  return ((c.x * p + c.y) * p + c.z) * p + c.w;
}

// FP32 force-pass arithmetic of one pair from its knot vectors: rho'_i
// (ci: c0..c2 of type2rhor[ti][tj]), rho'_j (cj), z2' (zd: c0..c2) and z2
// (zv: c3..c6) at r, fraction p
__device__ __forceinline__ void eam_pair_eval(const float4 &ci, const float4 &cj, const float4 &zd,
                                              const float4 &zv, float r, float p, double fpi, double fpj,
                                              double &fpair, double &phi) {
  const float rhoip = (ci.x * p + ci.y) * p + ci.z;
  const float rhojp = (cj.x * p + cj.y) * p + cj.z;
  const float z2p = (zd.x * p + zd.y) * p + zd.z;
  const float z2 = ((zv.x * p + zv.y) * p + zv.z) * p + zv.w;
  const float recip = __builtin_amdgcn_rcpf(r);  // hardware reciprocal, 1 ulp
  const float phif = z2 * recip;
  const float phip = z2p * recip - phif * recip;
  const float psip = (float)fpi * rhojp + (float)fpj * rhoip + phip;
  fpair = (double)(-psip * recip);
  phi = (double)phif;
}

// pair force factor fpair (f_i += del * fpair) and phi of the force pass
// (pair_eam.cpp:245-282); ri = type2rhor[ti][tj], rj = type2rhor[tj][ti],
// zi = type2z2r[ti][tj]
template <bool MIXED>
__device__ __forceinline__ void eam_pair_term(const EAMArgs &a, int ri, int rj, int zi, double rsq, double fpi,
                                              double fpj, double &fpair, double &phi) {
  const size_t kn = (size_t)(a.nr + 1);
  if (MIXED) {
    int m;
    float p;
    const float r = eam_index_mixed(rsq, a.rdr, a.nr, m, p);
    const float4 ci = a.rhor_pf[2 * (ri * kn + m)];  // c0..c2
    float4 cj = ci;  // (a conditional lvalue here would put ci in scratch memory)
    if (ri != rj) cj = a.rhor_pf[2 * (rj * kn + m)];
    eam_pair_eval(ci, cj, a.z2r_pf[2 * (zi * kn + m)], a.z2r_pf[2 * (zi * kn + m) + 1], r, p, fpi, fpj, fpair, phi);
    return;
  }
  const double r = sqrt(rsq);
  int m;
  double p;
  eam_spline_index<double>(r, a.rdr, a.nr, m, p);
  const double4 ci = a.rhor_p[2 * (ri * kn + m)];
  double4 cj = ci;
  if (ri != rj) cj = a.rhor_p[2 * (rj * kn + m)];
  const double4 zd = a.z2r_p[2 * (zi * kn + m)], zv = a.z2r_p[2 * (zi * kn + m) + 1];
  const double rhoip = (ci.x * p + ci.y) * p + ci.z;
  const double rhojp = (cj.x * p + cj.y) * p + cj.z;
  const double z2p = (zd.x * p + zd.y) * p + zd.z;
  const double z2 = ((zv.x * p + zv.y) * p + zv.z) * p + zv.w;
  const double recip = 1.0 / r;
  phi = z2 * recip;
  const double phip = z2p * recip - phi * recip;
  const double psip = fpi * rhojp + fpj * rhoip + phip;
  fpair = -psip * recip;
}

__host__ __device__ inline int imax2(int a, int b) { return a > b ? a : b; }
__host__ __device__ inline int imin2(int a, int b) { return a < b ? a : b; }

// Bin linearisation.  x-rows of bins stay contiguous (x fastest), so a
// stencil row is at most three contiguous atom ranges; the rows themselves
// are ordered in y-tiles of kBinTileY rows, z-major inside a tile:
// row(y, z) = y0*nbz + z*th + (y - y0), y0 = the tile's first row, th its
// height.  Kernels walking atoms in storage order then sweep a tile's
// z-planes, so the neighbour positions a wavefront window needs (within two
// rows and planes) stay resident in L2 instead of being re-fetched for every
// plane of a z-major order.
#ifndef MDX_BIN_TILE_Y
#define MDX_BIN_TILE_Y 8
#endif
constexpr int kBinTileY = MDX_BIN_TILE_Y;
__host__ __device__ inline int bin_row(const NeighArgs &a, int y, int z) {
  const int y0 = (y / kBinTileY) * kBinTileY;
  const int th = imin2(kBinTileY, a.nbin[1] - y0);
  return y0 * a.nbin[2] + z * th + (y - y0);
}
__host__ __device__ inline int bin_index(const NeighArgs &a, int x, int y, int z) {
  return bin_row(a, y, z) * a.nbin[0] + x;
}
__host__ __device__ inline void bin_coords(const NeighArgs &a, int b, int &x, int &y, int &z) {
  x = b % a.nbin[0];
  const int r = b / a.nbin[0];
  const int y0 = (r / (kBinTileY * a.nbin[2])) * kBinTileY;
  const int th = imin2(kBinTileY, a.nbin[1] - y0);
  const int rem = r - y0 * a.nbin[2];
  z = rem / th;
  y = y0 + rem % th;
}

// 32-bit fixed-point positions read by the mixed-precision gathers (16 B
// per neighbour instead of 32).  Per dimension q = round((x - lo) * scale)
// modulo 2^32, with scale = 2^32 / prd in a periodic dimension, so that the
// wrapping int32 difference q_i - q_j is the minimum-image separation, and
// scale = 2^30 / prd otherwise; step = 1 / scale = prd * 2^-32 (or 2^-30) is
// exact.  w holds the type.
struct QScale {
  double lo[3], scale[3], step[3];
};

__device__ __forceinline__ int qcoord(double x, double lo, double scale) {
  return (int)(unsigned)(long long)rint((x - lo) * scale);
}
__device__ __forceinline__ int4 quantize(const QScale &q, const double4 &p) {
  return make_int4(qcoord(p.x, q.lo[0], q.scale[0]), qcoord(p.y, q.lo[1], q.scale[1]),
                   qcoord(p.z, q.lo[2], q.scale[2]), static_cast<int>(p.w));
}
// separation q_i - q_j of one dimension in length units (FP64, exact product)
__device__ __forceinline__ double qdelta(int qi, int qj, double step) {
  return (double)(int)((unsigned)qi - (unsigned)qj) * step;
}

// Rebuild trigger of the displacement check (Neighbor::check_distance,
// src/neighbor.cpp:1322-1372): the first triggering lane of each wavefront
// stores the check's stamp into the pinned host word (the host's decide
// reads it) and into a device word (kernels queued before that decide read
// it), so a wavefront makes at most one write of each however many of its
// atoms moved.
__device__ __forceinline__ void mark_trigger(bool moved, int *trig, int *trig_dev, int stamp) {
  const unsigned long long m = __ballot(moved);
  if (moved && (int)(__ffsll((long long)m) - 1) == (int)(threadIdx.x & 63)) {
    *trig = stamp;
    *trig_dev = stamp;
  }
}

// Gate of a step's kernels queued before the host has checked what they
// depend on (mdx_run's look-ahead): they do nothing when
//   * the displacement trigger they were queued behind fired (the device
//     word holds the check's stamp), or
//   * the list build queued before them failed: its largest row exceeds
//     the list capacity, an atom left a non-periodic box or a bin overflowed
//     (the build's flags 0, 4 and 5).
// The host then rebuilds (or reports the error) and queues the step again.
// The answer is the same for every thread of the launch: no kernel of the
// step writes the build flags, and while the trigger word holds `stamp` no
// thread runs a body that could write it; otherwise the only value a body
// can write is a later check's stamp, which differs from `stamp` as well.
struct StepGate {
  const int *trig;    // device trigger word, or null
  int stamp;
  const int *bflags;  // build flags, or null
  int cap;            // list capacity the build ran with
};

__device__ __forceinline__ bool gate_closed(const StepGate &g) {
  if (g.trig && __hip_atomic_load(g.trig, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g.stamp) return true;
  if (g.bflags) {
    const int mx = __hip_atomic_load(g.bflags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int lost = __hip_atomic_load(g.bflags + 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int bin = __hip_atomic_load(g.bflags + 5, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (mx > g.cap || lost || bin) return true;
  }
  return false;
}

// Velocity-Verlet work fused into the force kernel (fix nve final integrate
// of this step, then initial integrate of the next, src/fix_nve.cpp:65-142)
// entry k of atom i in the chunk-major list (int offset)
__host__ __device__ __forceinline__ size_t list_index(int i, int k, int npad) {
  return (((size_t)(k >> 2) * npad + i) << 2) + (k & 3);
}

struct IntegArgs {
  double mass[MDX_NT1];
  double dtf, dtv, triggersq;
  int next;      // also do the next step's initial integrate (x -> xn)
  int check;     // test |xn - xhold|^2 > triggersq (next step is a decide step)
  int write_f;   // store f (last step of a run, energy steps)
  double4 *v;
  double4 *xn;
  int4 *xqn;  // fixed-point copy of xn (mixed gathers), or null
  QScale qs;
  const double4 *xhold;
  int *trig;      // pinned host word: set to stamp when an atom moved past the trigger
  int *trig_dev;  // device copy of the same word
  int stamp;
  StepGate gate;  // look-ahead launch: do nothing when the gate is closed (all null: always open)
  // energy/virial launches (any pair kernel): 1 = per-atom virial wanted,
  // tallied per pair (ev_tally); 0 = global virial only, from sum_i x_i f_i
  // plus the image-shift term of the pairs that cross a periodic boundary
  // (virial_fdotr_compute), about 6 FMAs per pair cheaper
  int pervir;
};

struct Launch {
  int block = 256;     // threads per workgroup of the pair kernels
  int tpa = 8;         // threads per atom in the pair kernels
};

}  // namespace mdx

struct mdx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // box
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, prd[3] = {0, 0, 0};
  int periodic[3] = {1, 1, 1};
  bool box_set = false;
  // atoms
  int n = 0, npad = 0, ntypes = 0, cap_atoms = 0;
  double mass[MDX_NT1] = {0};
  double4 *x = nullptr, *xb = nullptr, *v = nullptr, *vb = nullptr, *f = nullptr, *xhold = nullptr;
  // fixed-point positions for the mixed-precision gathers (xqb: the fused
  // path's next-step buffer); maintained by every position writer while qpos
  int4 *xq = nullptr, *xqb = nullptr;
  mdx::QScale qs{};
  int pos_format = MDX_POS_Q32;  // mixed precision gathers: MDX_POS_Q32 or MDX_POS_FP64
  bool qpos = false;             // xq is maintained (mixed precision, Q32)
  bool qbuild = false;           // xq is written at each build for the fixed-point candidate test
  bool xq_stale = true;          // xq must be rebuilt from x before the next force call
  int *tag = nullptr, *tagb = nullptr, *orig = nullptr, *origb = nullptr, *image = nullptr, *imageb = nullptr;
  int *inv = nullptr;  // caller index -> device index
  double *vatom = nullptr;
  double *dscratch = nullptr;    // device, 6*npad doubles (host<->device reordering)
  double *host_stage = nullptr;  // pinned, 6*npad doubles
  hipEvent_t stage_ev[8] = {};   // chunked staging copies: one event per chunk in flight
  int *host_istage = nullptr;    // pinned, 3*npad ints
  // bins / sort
  int nbins = 0, cap_bins = 0;
  mdx::NeighArgs na{};
  unsigned *bkey = nullptr, *bkey_s = nullptr;
  int *bidx = nullptr, *bidx_s = nullptr, *bin_count = nullptr, *bin_start = nullptr;
  void *scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  // neighbour settings and state
  double skin = 0.3;
  int every = 1, delay = 10, check = 1, one = 2000;
  double triggersq = 0.0225, cutneighmax = 0.0;
  // dflags: [0] max count, [4] lost atom, [8..9] 64-bit neighbour total
  int *nbr = nullptr, *numneigh = nullptr, *dflags = nullptr;
  // ghost-shell exchange (halo.hip): per swap, the device indices of the
  // atoms to send (valid until the next list build) and their image shift
  int *halo_list[MDX_HALO_SWAPS] = {};
  int halo_cap[MDX_HALO_SWAPS] = {}, halo_count[MDX_HALO_SWAPS] = {};
  double halo_shift[MDX_HALO_SWAPS][3] = {};
  // the list build a send list belongs to: every build permutes the atoms,
  // so a list made before the last build names the wrong ones
  long build_serial = 0;
  long halo_serial[MDX_HALO_SWAPS] = {};
  int *halo_flag = nullptr;  // [npad + 1] selection flags, then their exclusive scan
  int halo_flag_cap = 0;
  void *halo_scan_tmp = nullptr;
  size_t halo_scan_bytes = 0;
  int maxneigh = 0;
  long nbr_elems = 0;
  bool built = false;
  int ago = 0;
  long ncalls = 0, ndanger = 0;
  long ntimestep = 0;
  bool run_setup_done = false;
  bool fuse = true;  // fuse fix nve into the force kernel in mdx_run
  // pair
  int pair_style = MDX_PAIR_NONE;
  int precision = MDX_FP64;
  double cutforce = 0.0;
  double cutsq[MDX_NT1 * MDX_NT1] = {0};
  mdx::LJTab *lj_dev = nullptr;
  mdx::LJTab lj_host{};
  mdx::EAMArgs eam{};
  double *eam_tab = nullptr;  // device doubles: frho | rhor | z2r
  float *eam_tab_f = nullptr;
  double *eam_pk = nullptr;  // packed rhor | z2r (EAMArgs::rhor_p, z2r_p)
  float *eam_pk_f = nullptr;
  double *rho = nullptr, *fp = nullptr, *eemb = nullptr;
  // reductions
  double *partials = nullptr;
  long cap_partials = 0;
  int last_grid = 0;  // blocks of the last pair launch (partials to reduce)
  double *red_out = nullptr;    // device, MDX_RED_SLOTS
  double *red_host = nullptr;   // pinned, MDX_RED_SLOTS
  // Rebuild trigger of the displacement check, written by the kernels
  // straight into pinned host memory: a check stores its own stamp, so the
  // word needs no clearing and no device-to-host copy before it is read.
  int *trig_host = nullptr;
  int *trig_dev = nullptr;       // device copy of the trigger word
  int *bflags_host = nullptr;    // pinned copy of the last build's flags
  int *dflags_base = nullptr;    // two halves of 16 build flags (null: dflags alone, cleared per build)
  bool flags_dirty = false;      // a build started and did not pass its flags check
  bool bins_dirty = true;        // bin counters not known to be clear
  hipEvent_t build_ev = nullptr; // recorded after that copy
  int trig_stamp = 0;
  hipEvent_t step_ev = nullptr;  // recorded after each fused step's kernels (look-ahead decide)
  // results of the last compute
  double eng_vdwl = 0.0, virial[6] = {0, 0, 0, 0, 0, 0};
  // launch shape and timing (event pairs recorded on `stream`, resolved
  // without a per-launch sync when the statistics are read)
  mdx::Launch launch;
  bool profile = false;
  struct EvRec {
    hipEvent_t a, b;
    int kind;
  };
  std::vector<EvRec> evpool;
  size_t evused = 0;
  bool prof_open = false;  // a begin event was recorded for the current launch
  double kernel_ms[MDX_NKERNEL] = {0};
  long kernel_n[MDX_NKERNEL] = {0};
  long total_neighbors = 0;
  int max_neighbors = 0;
  int pair_ntypes = 0;  // ntypes of the last pair init (atoms may not be set)
  // storage owned by someone else (mdx_compute_noalloc's scratch): growth
  // paths report an error instead of freeing and allocating
  bool fixed_storage = false;
  long nbr_cap_elems = 0;  // list entries available when fixed_storage
};

namespace mdx {

// MI355X has 8 XCDs, each with its own L2; workgroup k is dispatched to XCD
// k % 8.  Kernels over bin-sorted atoms use this logical block index so
// that consecutive (spatially adjacent) tiles run on the same XCD and share
// its L2 for the neighbour gathers.  Grids are launched as a multiple of 8.
constexpr int kXCD = 8;
__device__ __forceinline__ int xcd_block() {
  const int per = gridDim.x / kXCD;
  return (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
}
// the stamp of the next displacement check (never 0, the word's initial value)
inline int next_trigger_stamp(mdx_ctx *c) {
  c->trig_stamp = c->trig_stamp == 0x7fffffff ? 1 : c->trig_stamp + 1;
  return c->trig_stamp;
}

inline int xcd_grid(long blocks) { return (int)(((blocks + kXCD - 1) / kXCD) * kXCD); }

void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
void halo_free(mdx_ctx *c);

// kernel timing with HIP events (no-ops unless profiling is on)
void prof_begin(mdx_ctx *c, int kind);
void prof_end(mdx_ctx *c);
int prof_resolve(mdx_ctx *c);

// Pair::init + Neighbor::init cutoffs, bin grid, first list and partials (api.cpp)
int prepare_ctx(mdx_ctx *c);
// no-allocation compute (noalloc.hip)
int noalloc_init(int device);
int compute_noalloc(mdx_ctx *c, int n, const double *x, const int *type, int eflag, int vflag, double *f,
                    double *eng_vdwl, double *virial);

// neighbour build (neigh.hip)
int neigh_setup_grid(mdx_ctx *c);
int neigh_build(mdx_ctx *c);
// neigh_build in two halves: the launches (sort, list, an asynchronous copy
// of the build flags to pinned memory), and the check of those flags, which
// waits only for that copy.  finish sets *regrow when the list capacity was
// too small (the caller rebuilds with neigh_build, which regrows).
int neigh_build_launch(mdx_ctx *c);
int neigh_build_finish(mdx_ctx *c, bool *regrow);
int alloc_list(mdx_ctx *c, int cap);
int neigh_check_displacement(mdx_ctx *c, int *flag);

// pair kernels (pair.hip); eflag/vflag as Pair::compute.  Writes f (device,
// sorted order), per-atom energy in f.w, per-atom virial in vatom, and the
// global energy/virial into c->red_host (after a sync) when ev is set.
int pair_compute(mdx_ctx *c, int ev, const IntegArgs *ig = nullptr);

// integrator (integrate.hip)
int nve_initial(mdx_ctx *c, double dtv, double dtf, int check);
int nve_final(mdx_ctx *c, double dtf);
int reduce_ke(mdx_ctx *c);
int finish_reduction(mdx_ctx *c, int nblocks, int nslots);
int scatter_positions(mdx_ctx *c, const double *host_x);
int refresh_xq(mdx_ctx *c);  // rebuild the fixed-point positions if stale
int copy_probe(mdx_ctx *c, long nbytes, int iters, double *gbs);
// device order -> caller order [n][3] on the host; add: host += device
int gather_vec3(mdx_ctx *c, const double4 *dev, double *host_out, bool add = false);

}  // namespace mdx
