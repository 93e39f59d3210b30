// neigh.hip - device neighbour-list build for gfx950.
//
// Replaces Neighbor::build -> bin_atoms -> half_bin_newton
// (src/neighbor.cpp:1380-1461,1868-1983; src/neigh_half_bin.cpp:285-419)
// and the periodic ghost shell of CommBrick::borders (src/comm_brick.cpp:696-864).
//
// Design: no ghost atoms.  At each build the owned atoms are wrapped into the
// box (Domain::pbc, src/domain.cpp:478), binned on the reference's own bin
// grid (binsize = cutneigh/2, src/neighbor.cpp:1580-1620), and stably
// sorted by bin (counting sort: arrival slots, then ranks by atom index)
// so that every x-row of bins is one contiguous atom range.  One lane per
// atom then scans the 5x5 rows of its stencil (each row at most 3
// contiguous segments: wrapped-low, main, wrapped-high; k_build_flat takes
// them a plane at a time as one candidate stream) and keeps j iff
// rsq <= cutneighsq, with the rsq the reference computes for a ghost atom:
// FP64 without contraction on the periodic image x_j + s*L
// (atom_vec_atomic.cpp:290-297), decided first on the fixed-point mirror in
// FP32 and re-tested that way within a narrow band of the cutoff.
// Candidates are listed in scan order.  The result is a FULL list in chunks
// of four entries, chunk-major (chunk c of atom i at c * npad + i); each
// entry holds the neighbour index (bits 0-26) and its periodic image code
// (bits 27-31) so pair kernels rebuild the same image between builds.
#include <hipcub/hipcub.hpp>

#include "mdx_internal.h"

namespace mdx {

namespace {

// closest-corner stencil extent, as Neighbor::setup_bins sx/sy/sz
int stencil_extent(double cut, double binsize, double bininv) {
  int s = static_cast<int>(cut * bininv);
  if (s * binsize < cut) s++;
  return s;
}

__device__ __forceinline__ int ref_coord2bin(double c, double lo, double hi, double bininv, int nbin) {
  // src/neighbor.cpp:1918-1945, one dimension
  int ib;
  if (c >= hi)
    ib = static_cast<int>((c - hi) * bininv) + nbin;
  else if (c >= lo) {
    ib = static_cast<int>((c - lo) * bininv);
    ib = min(ib, nbin - 1);
  } else
    ib = static_cast<int>((c - lo) * bininv) - 1;
  return ib;
}

// True when the reference's half_bin_newton would evaluate the pair from the
// owned atom at xi with the ghost image xg: its bin offset lies in the upper
// half stencil (src/neigh_stencil.cpp:143-158).
__device__ bool ref_owner_side(const RefBins &r, double xi0, double xi1, double xi2, double g0, double g1,
                               double g2) {
  int di = ref_coord2bin(g0, r.bboxlo[0], r.bboxhi[0], r.bininv[0], r.nbin[0]) -
           ref_coord2bin(xi0, r.bboxlo[0], r.bboxhi[0], r.bininv[0], r.nbin[0]);
  int dj = ref_coord2bin(g1, r.bboxlo[1], r.bboxhi[1], r.bininv[1], r.nbin[1]) -
           ref_coord2bin(xi1, r.bboxlo[1], r.bboxhi[1], r.bininv[1], r.nbin[1]);
  int dk = ref_coord2bin(g2, r.bboxlo[2], r.bboxhi[2], r.bininv[2], r.nbin[2]) -
           ref_coord2bin(xi2, r.bboxlo[2], r.bboxhi[2], r.bininv[2], r.nbin[2]);
  return dk > 0 || (dk == 0 && (dj > 0 || (dj == 0 && di > 0)));
}

// rsq of a - b without fused multiply-adds: the reference's x86-64 rounding
__device__ __forceinline__ double rsq_exact(double a0, double a1, double a2, double b0, double b1, double b2) {
#pragma clang fp contract(off)
  double dx = a0 - b0;
  double dy = a1 - b1;
  double dz = a2 - b2;
  return dx * dx + dy * dy + dz * dz;
}

// The FP64 keep test of one candidate with the image shift (shx, shy, shz):
// rsq on the ghost coordinate as CommBrick::borders forms it, and for an
// image pair within a hair of the cutoff, the reference's side rule.
__device__ __forceinline__ bool keep_exact(const NeighArgs &a, const double4 &xi, const double4 &xj, double shx,
                                           double shy, double shz, int code, double c) {
  const double g0 = xj.x + shx, g1 = xj.y + shy, g2 = xj.z + shz;
  const double r = rsq_exact(xi.x, xi.y, xi.z, g0, g1, g2);
  bool keep = r <= c;
  // the mirrored rsq of an image pair differs from r by a few ulps, so the
  // reference's side can only matter within a hair of c
  if (code != MDX_NOSHIFT && fabs(r - c) <= 1e-10 * c) {
    const bool keep_b = rsq_exact(xj.x, xj.y, xj.z, xi.x - shx, xi.y - shy, xi.z - shz) <= c;
    if (keep_b != keep && !ref_owner_side(a.ref, xi.x, xi.y, xi.z, g0, g1, g2)) keep = keep_b;
  }
  return keep;
}

// Fixed-point candidate test (mixed precision): rsq from the wrapping
// integer differences of the 32-bit position mirror, in FP32.  Its error
// (quantisation <= 5e-7 relative, enforced where the mirror is enabled, plus
// FP32 rounding < 7e-7) is well inside kBuildBand, so a candidate outside
// [c(1-band), c(1+band)] is decided exactly as the FP64 test decides it; the
// rare one inside is re-tested by keep_exact on the FP64 positions.  The
// wrapping difference is the minimum image, which is the candidate's own
// image whenever that image can be within the cutoff, provided no atom
// appears in the stencil under two images (2S+1 <= nbin in every periodic
// dimension, checked on the host).
constexpr double kBuildBand = 2e-6;

// bins of the build's grid per neighbour cutoff (the reference's grid has 2)
#ifndef MDX_BIN_DIV
#define MDX_BIN_DIV 2
#endif
constexpr int kBinDiv = MDX_BIN_DIV;
struct QStep {
  float s[3];
};

#ifndef MDX_BUILD_Q_WAVES
#define MDX_BUILD_Q_WAVES 5
#endif
#ifndef MDX_BUILD_LANE_U
#define MDX_BUILD_LANE_U 4
#endif
constexpr int kLaneU = MDX_BUILD_LANE_U;  // candidates loaded together per lane

#ifndef MDX_BUILD_PRUNE
#define MDX_BUILD_PRUNE 1
#endif
// 0: whole stencil, 1: planes, rows and row ends, 2: planes and rows
constexpr int kPrune = MDX_BUILD_PRUNE;

#ifndef MDX_BUILD_MIN_WAVES
#define MDX_BUILD_MIN_WAVES 1
#endif

// The (ntypes+1)^2 cutoff table in LDS: a kernel-argument array indexed by
// per-lane types is re-read from memory for every candidate.  Called by every
// thread of the block before any thread leaves.
__device__ __forceinline__ void stage_cutneighsq(const NeighArgs &a, double *s) {
  const int nt2 = a.ntypes1 * a.ntypes1;
  for (int q = threadIdx.x; q < nt2; q += blockDim.x) s[q] = a.cutneighsq[q];
  __syncthreads();
}

// the largest squared neighbour cutoff of each type (row maxima of the table)
__device__ __forceinline__ void stage_cmax(const NeighArgs &a, double *s) {
  for (int t = threadIdx.x; t < a.ntypes1; t += blockDim.x) {
    double m = 0.0;
    for (int u = 0; u < a.ntypes1; u++) m = fmax(m, a.cutneighsq[t * a.ntypes1 + u]);
    s[t] = m;
  }
}

// the FP32 decision bounds c(1 -/+ kBuildBand) of the fixed-point test
__device__ __forceinline__ void stage_band(const NeighArgs &a, float *lo, float *hi) {
  const int nt2 = a.ntypes1 * a.ntypes1;
  for (int q = threadIdx.x; q < nt2; q += blockDim.x) {
    lo[q] = static_cast<float>(a.cutneighsq[q] * (1.0 - kBuildBand));
    hi[q] = static_cast<float>(a.cutneighsq[q] * (1.0 + kBuildBand));
  }
}

// Domain::pbc + bin index + histogram; one thread per atom.  slot[i] is the
// atom's arrival order in its bin (not deterministic; k_bin_rank orders
// each bin by atom index).  Atoms arrive nearly bin-sorted (the previous
// build's order), so the lanes of a wavefront with equal bins form runs: one
// returning atomic per run reserves the run's slots, and no two lanes of a
// wavefront hit one counter.
__global__ void k_wrap_bin(NeighArgs a, double4 *__restrict__ x, int *__restrict__ image,
                           unsigned *__restrict__ key, int *__restrict__ slot, int *__restrict__ count,
                           int *__restrict__ dflags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  unsigned b = 0xffffffffu;  // past the end: its own run, no atomic
  if (i < a.n) {
  double4 p = x[i];
  double c[3] = {p.x, p.y, p.z};
  int img = image[i];
  int im[3] = {img & MDX_IMGMASK, (img >> 10) & MDX_IMGMASK, img >> 20};
  int ib[3];
  bool lost = false, moved = false;
  for (int d = 0; d < 3; d++) {
    double lo = a.lo[d], hi = a.lo[d] + a.prd[d];
    if (a.periodic[d]) {
      if (c[d] < lo) {
        c[d] += a.prd[d];
        im[d] = (im[d] - 1) & MDX_IMGMASK;
        moved = true;
      }
      if (c[d] >= hi) {
        c[d] -= a.prd[d];
        c[d] = fmax(c[d], lo);
        im[d] = (im[d] + 1) & MDX_IMGMASK;
        moved = true;
      }
    } else if (c[d] < lo || c[d] >= hi) {
      lost = true;
    }
    int b = static_cast<int>((c[d] - lo) * a.bininv[d]);
    b = min(max(b, 0), a.nbin[d] - 1);
    ib[d] = b;
  }
  if (lost) atomicOr(&dflags[4], 1);
  // most atoms stay inside the box between builds: store only the wrapped ones
  if (moved) {
    p.x = c[0];
    p.y = c[1];
    p.z = c[2];
    x[i] = p;
    image[i] = (im[2] << 20) | (im[1] << 10) | im[0];
  }
  b = (unsigned)bin_index(a, ib[0], ib[1], ib[2]);
  key[i] = b;
  }
  // runs of equal bins across the wavefront's lanes (all 64 lanes take part)
  const unsigned prev = __shfl_up(b, 1);
  const bool head = lane == 0 || prev != b;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
  const int start = 63 - __clzll(heads & upto);
  const unsigned long long above = heads & ~upto;
  const int len = (above ? __ffsll((long long)above) - 1 : 64) - lane;  // run length at a head
  int base = 0;
  if (head && b != 0xffffffffu) base = atomicAdd(&count[b], len);
  base = __shfl(base, start);
  if (i < a.n) slot[i] = base + (lane - start);
}

// Stable counting sort by bin, in two passes after the scan of the bin
// counts: each atom lands in its bin's range at its arrival slot, then each
// entry of a bin takes the rank of its atom index among the bin's entries.
// The result is the order a stable sort of (bin, current index) gives:
// bins ascending, current index ascending within a bin.
// Also clears, for the next build, the bin counters (the scan has consumed
// them) and the other half of the double-buffered build flags (null: none),
// so that a build issues no memsets.
__global__ void k_bin_scatter(int n, const unsigned *__restrict__ key, const int *__restrict__ slot,
                              const int *__restrict__ bin_start, int *__restrict__ tmp, int *__restrict__ count,
                              int ncount, int *__restrict__ next_flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = i; k < ncount; k += gridDim.x * blockDim.x) count[k] = 0;
  if (next_flags && i < 16) next_flags[i] = 0;
  if (i >= n) return;
  tmp[bin_start[key[i]] + slot[i]] = i;
}

// The rank scan is linear in the bin's occupancy, so the pass is quadratic in
// it.  Every two atoms of one bin are neighbours (the bin diagonal,
// sqrt(3)/2 of the neighbour cutoff, is inside it), so a bin holding more
// than `one` + 1 atoms gives each of them more than `one` neighbours: the
// reference stops there with "Neighbor list overflow" (neigh_half_bin.cpp:
// 414-415, the first atom of the bin's chain).  Such a bin keeps its arrival
// order (still a permutation, so the passes after it stay in bounds) and
// raises the overflow flag, which the build reports; the scan is thus
// bounded by `one` + 1 entries per atom.
__global__ void k_bin_rank(int n, const unsigned *__restrict__ key, const int *__restrict__ bin_start,
                           const int *__restrict__ tmp, int *__restrict__ src, unsigned *__restrict__ key_s,
                           int max_occupancy, int *__restrict__ dflags) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = tmp[p];
  const unsigned b = key[i];
  const int s = bin_start[b], e = bin_start[b + 1];
  int r = p - s;
  if (e - s <= max_occupancy) {
    r = 0;
    for (int k = s; k < e; k++) r += tmp[k] < i;
  } else if (p == s) {
    atomicOr(&dflags[5], 1);
  }
  src[s + r] = i;
  key_s[s + r] = b;
}

__global__ void k_permute(int n, const int *__restrict__ src, const double4 *__restrict__ x,
                          const double4 *__restrict__ v, const int *__restrict__ tag,
                          const int *__restrict__ orig, const int *__restrict__ image, double4 *__restrict__ xo,
                          double4 *__restrict__ vo, int *__restrict__ tago, int *__restrict__ origo,
                          int *__restrict__ imageo, int *__restrict__ inv, double4 *__restrict__ xhold,
                          int4 *__restrict__ xq, QScale qs) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int o = src[k];
  double4 p = x[o];
  xo[k] = p;
  xhold[k] = p;
  if (xq) xq[k] = quantize(qs, p);
  vo[k] = v[o];
  tago[k] = tag[o];
  int oi = orig[o];
  origo[k] = oi;
  imageo[k] = image[o];
  inv[oi] = k;
}

// One lane per atom.  A lane walks ~300 candidates, so
// one gather in flight at a time leaves it latency bound: candidates are
// taken four at a time (four gathers in flight), and accepted entries are
// stored four at a time as one 16-B store into the row.  Candidate order,
// keep test and tie rule do not depend on Q or ONE, so every variant gives
// the same rows entry for entry.  Q: candidates are read from the fixed-point
// mirror (16 B instead of 32) and tested in FP32 (see kBuildBand); that
// variant is held to 5 wavefronts per SIMD (96 VGPRs), which measured 9%
// faster on 2 M atoms than the 4 its register count would otherwise give.
//
// Per-atom stencil pruning skips only candidates the keep test rejects, so
// the rows do not change: a z-plane, a row or the end bins of a row whose
// closest point lies beyond the largest cutoff of type ti hold no neighbour
// of i (about a third of the 5x5x5 stencil for an atom inside its bin).
// Distances are taken in bin units from t = (x - lo) * bininv, the
// coordinate the binning used, in FP32: its rounding (below 1e-5 bin for
// up to 10^4 bins per dimension) is dwarfed by the margins (1e-3 bin on
// every gap and reach, 1e-4 relative on the squared cutoff).
//
// ONE: a single atom type, so the cutoff and its band are loop-invariant
// values instead of per-candidate table reads.  Accepted entries go through
// a four-entry shift register (q0 oldest) that is stored as one 16-B write
// whenever it holds a group of four.
template <bool Q, bool ONE>
__global__ __launch_bounds__(256, Q ? MDX_BUILD_Q_WAVES : MDX_BUILD_MIN_WAVES) void k_build_lane(NeighArgs a, const double4 *__restrict__ x,
                                                   const int4 *__restrict__ xq, QStep qs,
                                                   const unsigned *__restrict__ bkey,
                                                   const int *__restrict__ bin_start, int *__restrict__ nbr,
                                                   int *__restrict__ numneigh, int *__restrict__ dflags,
                                                   unsigned long long *__restrict__ nsum) {
  __shared__ double s_cut[MDX_NT1 * MDX_NT1];
  __shared__ float s_lo[Q ? MDX_NT1 * MDX_NT1 : 1], s_hi[Q ? MDX_NT1 * MDX_NT1 : 1];
  __shared__ double s_cmax[MDX_NT1];
  if constexpr (Q) stage_band(a, s_lo, s_hi);
  stage_cmax(a, s_cmax);
  stage_cutneighsq(a, s_cut);
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double4 xi = x[i];
  const int ti = static_cast<int>(xi.w);
  int4 qi = make_int4(0, 0, 0, 0);
  if constexpr (Q) qi = xq[i];
  const unsigned b = bkey[i];
  const int nbx = a.nbin[0], nby = a.nbin[1], nbz = a.nbin[2];
  int ix, iy, iz;
  bin_coords(a, (int)b, ix, iy, iz);
  // this atom's column of the chunk-major list: chunk c at col[c * npad]
  int4 *col = reinterpret_cast<int4 *>(nbr) + i;
  int cnt = 0;
  int q0 = 0, q1 = 0, q2 = 0, q3 = 0;  // the last four accepted entries, q3 the newest
  const int t11 = a.ntypes1 + 1;       // (1, 1) in the type tables
  const float c_lo = (ONE && Q) ? s_lo[Q ? t11 : 0] : 0.0f, c_hi = (ONE && Q) ? s_hi[Q ? t11 : 0] : 0.0f;
  const double c_cut = ONE ? s_cut[t11] : 0.0;
  const float tx = static_cast<float>((xi.x - a.lo[0]) * a.bininv[0]),
              ty = static_cast<float>((xi.y - a.lo[1]) * a.bininv[1]),
              tz = static_cast<float>((xi.z - a.lo[2]) * a.bininv[2]);
  const float bsy = static_cast<float>(1.0 / a.bininv[1]), bsz = static_cast<float>(1.0 / a.bininv[2]);
  const float reach = static_cast<float>(s_cmax[ti] * (1.0 + 1e-4));
  // FP32 bin coordinates stay within the margins up to 8192 bins a side
  const int prune = (a.nbin[0] <= 8192 && a.nbin[1] <= 8192 && a.nbin[2] <= 8192) ? kPrune : 0;
  for (int dz = -a.S[2]; dz <= a.S[2]; dz++) {
    // gap between atom i and the plane of bins iz+dz, in bins
    const float gz = (dz > 0 ? (iz + dz) - tz : dz < 0 ? tz - (iz + dz + 1) : 0.0f) - 1e-3f;
    const float dz2 = gz > 0.0f ? gz * gz * bsz * bsz : 0.0f;
    if (prune && dz2 > reach) continue;
    int jz = iz + dz, sz = 0;
    if (jz < 0) {
      if (!a.periodic[2]) continue;
      jz += nbz;
      sz = -1;
    } else if (jz >= nbz) {
      if (!a.periodic[2]) continue;
      jz -= nbz;
      sz = 1;
    }
    for (int dy = -a.S[1]; dy <= a.S[1]; dy++) {
      const float gy = (dy > 0 ? (iy + dy) - ty : dy < 0 ? ty - (iy + dy + 1) : 0.0f) - 1e-3f;
      const float dyz2 = dz2 + (gy > 0.0f ? gy * gy * bsy * bsy : 0.0f);
      if (prune && dyz2 > reach) continue;
      int jy = iy + dy, sy = 0;
      if (jy < 0) {
        if (!a.periodic[1]) continue;
        jy += nby;
        sy = -1;
      } else if (jy >= nby) {
        if (!a.periodic[1]) continue;
        jy -= nby;
        sy = 1;
      }
      const int rowbase = bin_row(a, jy, jz) * nbx;
      // bins of this row within reach in x: gap max(0, jx - tx, tx - jx - 1) <= wb
      int x0 = ix - a.S[0], x1 = ix + a.S[0];
      if (prune == 1) {
        const float wb = sqrtf(reach - dyz2) * static_cast<float>(a.bininv[0]) + 1e-3f;
        x0 = max(x0, static_cast<int>(ceilf(tx - 1.0f - wb)));
        x1 = min(x1, static_cast<int>(floorf(tx + wb)));
      }
      const double shy = sy * a.prd[1], shz = sz * a.prd[2];
      for (int s = 0; s < 3; s++) {
        int lo, hi, sx;
        if (s == 0) {
          if (!(x0 < 0 && a.periodic[0])) continue;
          lo = max(x0 + nbx, 0);
          hi = nbx - 1;
          sx = -1;
        } else if (s == 1) {
          lo = max(x0, 0);
          hi = min(x1, nbx - 1);
          sx = 0;
        } else {
          if (!(x1 >= nbx && a.periodic[0])) continue;
          lo = 0;
          hi = min(x1 - nbx, nbx - 1);
          sx = 1;
        }
        if (lo > hi) continue;
        const double shx = sx * a.prd[0];
        const int code = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
        const int j0 = bin_start[rowbase + lo];
        const int j1 = bin_start[rowbase + hi + 1];
        for (int base = j0; base < j1; base += kLaneU) {
          double4 xj[Q ? 1 : kLaneU];
          int4 qj[Q ? kLaneU : 1];
#pragma unroll
          for (int u = 0; u < kLaneU; u++) {  // in range; unused tail ignored
            if constexpr (Q) qj[u] = xq[min(base + u, j1 - 1)];
            else xj[u] = x[min(base + u, j1 - 1)];
          }
#pragma unroll
          for (int u = 0; u < kLaneU; u++) {
            // no early exit: every loaded candidate is used, so the four
            // loads stay in flight together
            const int j = base + u;
            const bool valid = j < j1 && !(code == MDX_NOSHIFT && j == i);
            bool keep;
            if constexpr (Q) {
              const float dx = static_cast<float>(static_cast<int>((unsigned)qi.x - (unsigned)qj[u].x)) * qs.s[0];
              const float dy = static_cast<float>(static_cast<int>((unsigned)qi.y - (unsigned)qj[u].y)) * qs.s[1];
              const float dz = static_cast<float>(static_cast<int>((unsigned)qi.z - (unsigned)qj[u].z)) * qs.s[2];
              const float rq = dx * dx + dy * dy + dz * dz;
              if constexpr (ONE) {
                keep = valid && rq <= c_lo;
                if (valid && !keep && rq <= c_hi) keep = keep_exact(a, xi, x[j], shx, shy, shz, code, c_cut);
              } else {
                const int t = ti * a.ntypes1 + qj[u].w;
                keep = valid && rq <= s_lo[t];
                if (valid && !keep && rq <= s_hi[t]) keep = keep_exact(a, xi, x[j], shx, shy, shz, code, s_cut[t]);
              }
            } else {
              const double c = ONE ? c_cut : s_cut[ti * a.ntypes1 + static_cast<int>(xj[u].w)];
              keep = valid && keep_exact(a, xi, xj[u], shx, shy, shz, code, c);
            }
            if (keep) {
              q0 = q1;
              q1 = q2;
              q2 = q3;
              q3 = j | (code << MDX_IDX_BITS);
              cnt++;
              if ((cnt & 3) == 0 && cnt <= a.maxneigh) col[(size_t)((cnt - 4) >> 2) * a.npad] = make_int4(q0, q1, q2, q3);
            }
          }
        }
      }
    }
  }
  // entries after the last full group of four: the newest `tail` of the
  // register, into the first components of chunk t0/4
  const int tail = cnt & 3, t0 = cnt - tail;
  const int r0 = tail == 1 ? q3 : tail == 2 ? q2 : q1, r1 = tail == 2 ? q3 : q2;
  int *last = reinterpret_cast<int *>(col + (size_t)(t0 >> 2) * a.npad);
  if (tail > 0 && t0 < a.maxneigh) last[0] = r0;
  if (tail > 1 && t0 + 1 < a.maxneigh) last[1] = r1;
  if (tail > 2 && t0 + 2 < a.maxneigh) last[2] = q3;
  numneigh[i] = cnt;
  atomicMax(&dflags[0], cnt);
  atomicAdd(nsum, (unsigned long long)cnt);
}

// The (-1, 0, +1) image shifts of a periodic image code (see MDX_NOSHIFT)
__device__ __forceinline__ void code_shift(const NeighArgs &a, int code, double &shx, double &shy, double &shz) {
  const int cz = code / 9, r = code - 9 * cz, cy = r / 3, cx = r - 3 * cy;
  shx = (cx - 1) * a.prd[0];
  shy = (cy - 1) * a.prd[1];
  shz = (cz - 1) * a.prd[2];
}

#ifndef MDX_BUILD_FLAT
#define MDX_BUILD_FLAT 1
#endif
#ifndef MDX_FLAT_WAVES
#define MDX_FLAT_WAVES 5
#endif
#ifndef MDX_FLAT_MAXS
#define MDX_FLAT_MAXS MDX_BIN_DIV
#endif
constexpr int kFlatBlock = 256;
constexpr int kFlatMaxS = MDX_FLAT_MAXS;    // largest stencil extent of the flat build
constexpr int kFlatRows = 2 * kFlatMaxS + 1;  // rows of a stencil plane
constexpr int kFlatSeg = 2 * kFlatRows + 2;  // queued segments per lane: one plane, an end mark, a spare slot

// The list build as one candidate stream per lane and stencil plane.
// k_build_lane runs each stencil row for as long as the lane with the most
// candidates in it (bin occupancies vary from lane to lane), padded to
// whole groups of four: a wavefront executes ~2.5x the candidates of its
// mean lane (counted on a perturbed fcc lattice at the lj/cut benchmark's
// density and cutoff).  Here each plane of the stencil is taken in two steps.
// First the bin ranges of the plane's rows (at most two pieces per row: the
// main one and one wrapped across a periodic x boundary) are formed and
// all their bin_start reads issued together, and the non-empty segments
// (a contiguous atom range with its image code) are queued in LDS.  Then
// the queue is drained as one stream, four candidates at a time, a lane
// moving to its next segment as soon as one ends, so the wavefront runs as
// long as its busiest lane's plane total (over a 5-plane stencil, ~1.35x
// the candidates of its mean lane).  Candidate order, keep test and tie rule
// are k_build_lane's: the rows are entry for entry the same.  Queue slot k
// of lane t is [k * 256 + t] (consecutive lanes, consecutive banks); a
// segment is held as two list entries (first, end), j | code <<
// MDX_IDX_BITS, so the candidate's entry is the running value itself and
// the self pair is one compare.  Measured at 2.048 M lj/cut atoms: 1.15 ms
// per build against 1.33 (fixed-point candidates), 1.41 against 1.66 (FP64
// candidates).
// launch-uniform values of the flat build, formed on the host so that they
// stay in scalar registers: FP32 bin sizes and inverse, the pruning switch,
// and for a single type the cutoff, its FP32 band and the pruning reach
struct FlatConst {
  float bsy, bsz, binx;
  float c_lo, c_hi, reach;
  double c_cut;
  int prune;
};
template <bool Q, bool ONE>
__global__ __launch_bounds__(kFlatBlock, MDX_FLAT_WAVES) void k_build_flat(NeighArgs a, const double4 *__restrict__ x,
                                                   const int4 *__restrict__ xq, QStep qs,
                                                   const unsigned *__restrict__ bkey,
                                                   const int *__restrict__ bin_start, int *__restrict__ nbr,
                                                   int *__restrict__ numneigh, int *__restrict__ dflags,
                                                   unsigned long long *__restrict__ nsum, FlatConst fc) {
  __shared__ double s_cut[MDX_NT1 * MDX_NT1];
  __shared__ float s_lo[Q ? MDX_NT1 * MDX_NT1 : 1], s_hi[Q ? MDX_NT1 * MDX_NT1 : 1];
  __shared__ double s_cmax[MDX_NT1];
  __shared__ unsigned s_q0[kFlatSeg * kFlatBlock], s_q1[kFlatSeg * kFlatBlock];
  if constexpr (Q) stage_band(a, s_lo, s_hi);
  stage_cmax(a, s_cmax);
  stage_cutneighsq(a, s_cut);
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  unsigned *const lq0 = s_q0 + threadIdx.x, *const lq1 = s_q1 + threadIdx.x;
  const double4 xi = x[i];
  const int ti = static_cast<int>(xi.w);
  int4 qi = make_int4(0, 0, 0, 0);
  if constexpr (Q) qi = xq[i];
  const unsigned b = bkey[i];
  const int nbx = a.nbin[0], nby = a.nbin[1], nbz = a.nbin[2];
  int ix, iy, iz;
  bin_coords(a, (int)b, ix, iy, iz);
  int4 *col = reinterpret_cast<int4 *>(nbr) + i;
  int cnt = 0;
  int q0 = 0, q1 = 0, q2 = 0, q3 = 0;
  const int t11 = a.ntypes1 + 1;
  const float c_lo = fc.c_lo, c_hi = fc.c_hi;
  const double c_cut = fc.c_cut;
  const float tx = static_cast<float>((xi.x - a.lo[0]) * a.bininv[0]),
              ty = static_cast<float>((xi.y - a.lo[1]) * a.bininv[1]),
              tz = static_cast<float>((xi.z - a.lo[2]) * a.bininv[2]);
  const float bsy = fc.bsy, bsz = fc.bsz;
  const float reach = ONE ? fc.reach : static_cast<float>(s_cmax[ti] * (1.0 + 1e-4));
  const int prune = fc.prune;
  const unsigned self = static_cast<unsigned>(i) | (MDX_NOSHIFT << MDX_IDX_BITS);
  int nq = 0;
  for (int dz = -a.S[2]; dz <= a.S[2]; dz++) {
    const float gz = (dz > 0 ? (iz + dz) - tz : dz < 0 ? tz - (iz + dz + 1) : 0.0f) - 1e-3f;
    const float dz2 = gz > 0.0f ? gz * gz * bsz * bsz : 0.0f;
    bool zok = !(prune && dz2 > reach);
    int jz = iz + dz, sz = 0;
    if (jz < 0) {
      zok = zok && a.periodic[2];
      jz += nbz;
      sz = -1;
    } else if (jz >= nbz) {
      zok = zok && a.periodic[2];
      jz -= nbz;
      sz = 1;
    }
    // the plane's rows: bin ranges of each row's (at most two) pieces in
    // scan order, then all their bin_start reads at once, then the queue
    unsigned blo[kFlatRows][2], bhi[kFlatRows][2], tag[kFlatRows][2];
#pragma unroll
    for (int r = 0; r < kFlatRows; r++) {
      const int dy = r - kFlatRows / 2;
      blo[r][0] = bhi[r][0] = blo[r][1] = bhi[r][1] = 0u;  // empty
      tag[r][0] = tag[r][1] = 0u;
      if (!zok || dy < -a.S[1] || dy > a.S[1]) continue;
      const float gy = (dy > 0 ? (iy + dy) - ty : dy < 0 ? ty - (iy + dy + 1) : 0.0f) - 1e-3f;
      const float dyz2 = dz2 + (gy > 0.0f ? gy * gy * bsy * bsy : 0.0f);
      if (prune && dyz2 > reach) continue;
      int jy = iy + dy, sy = 0;
      if (jy < 0) {
        if (!a.periodic[1]) continue;
        jy += nby;
        sy = -1;
      } else if (jy >= nby) {
        if (!a.periodic[1]) continue;
        jy -= nby;
        sy = 1;
      }
      const unsigned rowbase = static_cast<unsigned>(bin_row(a, jy, jz) * nbx);
      int x0 = ix - a.S[0], x1 = ix + a.S[0];
      if (prune == 1) {
        const float wb = sqrtf(reach - dyz2) * fc.binx + 1e-3f;
        x0 = max(x0, static_cast<int>(ceilf(tx - 1.0f - wb)));
        x1 = min(x1, static_cast<int>(floorf(tx + wb)));
      }
      const int cyz = 3 * (sy + 1) + 9 * (sz + 1);
      // main piece [max(x0, 0), min(x1, nbx - 1)]; a wrapped piece below 0
      // comes before it, one above nbx - 1 after it (2S <= nbx: not both)
      const int mlo = max(x0, 0), mhi = min(x1, nbx - 1);
      const bool wlo = x0 < 0 && a.periodic[0], whi = x1 >= nbx && a.periodic[0];
      // main piece (empty when lo > hi: both bounds at rowbase)
      const unsigned m0 = mlo <= mhi ? rowbase + mlo : rowbase, m1 = mlo <= mhi ? rowbase + mhi + 1 : rowbase;
      const unsigned mtag = static_cast<unsigned>(1 + cyz) << MDX_IDX_BITS;
      if (wlo) {
        blo[r][0] = rowbase + (x0 + nbx);
        bhi[r][0] = rowbase + nbx;
        tag[r][0] = static_cast<unsigned>(cyz) << MDX_IDX_BITS;  // sx = -1
        blo[r][1] = m0;
        bhi[r][1] = m1;
        tag[r][1] = mtag;
      } else {
        blo[r][0] = m0;
        bhi[r][0] = m1;
        tag[r][0] = mtag;
        if (whi) {
          blo[r][1] = rowbase;
          bhi[r][1] = rowbase + (x1 - nbx) + 1;
          tag[r][1] = static_cast<unsigned>(2 + cyz) << MDX_IDX_BITS;  // sx = +1
        }
      }
    }
    int jlo[kFlatRows][2], jhi[kFlatRows][2];
#pragma unroll
    for (int r = 0; r < kFlatRows; r++)
#pragma unroll
      for (int p = 0; p < 2; p++) {
        // 32-bit byte offsets from the array base: one address register each
        jlo[r][p] = *reinterpret_cast<const int *>(reinterpret_cast<const char *>(bin_start) + blo[r][p] * 4u);
        jhi[r][p] = *reinterpret_cast<const int *>(reinterpret_cast<const char *>(bin_start) + bhi[r][p] * 4u);
      }
#pragma unroll
    for (int r = 0; r < kFlatRows; r++)
#pragma unroll
      for (int p = 0; p < 2; p++)
        if (jlo[r][p] < jhi[r][p]) {
          lq0[nq * kFlatBlock] = static_cast<unsigned>(jlo[r][p]) | tag[r][p];
          lq1[nq * kFlatBlock] = static_cast<unsigned>(jhi[r][p]) | tag[r][p];
          nq++;
        }
    {
      // the plane's candidates as one stream per lane
      // end mark (0, 0) after the queued segments (a real segment ends at an
      // entry > 0): a lane that reaches it stops advancing, so it reads at
      // most one slot past it
      lq0[nq * kFlatBlock] = 0u;
      lq1[nq * kFlatBlock] = 0u;
      unsigned e = lq0[0], e1 = lq1[0];                              // current segment
      unsigned ne = lq0[kFlatBlock], ne1 = lq1[kFlatBlock];          // next (unused past the mark)
      const unsigned *n0 = lq0 + 2 * kFlatBlock, *n1 = lq1 + 2 * kFlatBlock;
      for (;;) {
        if (e == e1) break;  // the end mark (segments are never empty)
        unsigned eu[kLaneU];
#pragma unroll
        for (int u = 0; u < kLaneU; u++) {
          eu[u] = e != e1 ? e : self;  // past the end: the lane's own atom, rejected below
          if (e != e1) e++;
          if (e == e1 && e1 != 0u) {  // segment done: take the prefetched one, prefetch the next
            e = ne;
            e1 = ne1;
            ne = *n0;
            ne1 = *n1;
            n0 += kFlatBlock;
            n1 += kFlatBlock;
          }
        }
        double4 xj[Q ? 1 : kLaneU];
        int4 qj[Q ? kLaneU : 1];
#pragma unroll
        for (int u = 0; u < kLaneU; u++) {
          const int j = eu[u] & MDX_IDX_MASK;
          if constexpr (Q) qj[u] = xq[j];
          else xj[u] = x[j];
        }
#pragma unroll
        for (int u = 0; u < kLaneU; u++) {
          const bool valid = eu[u] != self;
          const int code = static_cast<int>(static_cast<unsigned>(eu[u]) >> MDX_IDX_BITS);
          bool keep;
          if constexpr (Q) {
            const float dx = static_cast<float>(static_cast<int>((unsigned)qi.x - (unsigned)qj[u].x)) * qs.s[0];
            const float dy = static_cast<float>(static_cast<int>((unsigned)qi.y - (unsigned)qj[u].y)) * qs.s[1];
            const float dz = static_cast<float>(static_cast<int>((unsigned)qi.z - (unsigned)qj[u].z)) * qs.s[2];
            const float rq = dx * dx + dy * dy + dz * dz;
            const int t = ONE ? t11 : ti * a.ntypes1 + qj[u].w;
            const float lo = ONE ? c_lo : s_lo[t], hi = ONE ? c_hi : s_hi[t];
            keep = valid && rq <= lo;
            if (valid && !keep && rq <= hi) {
              double shx, shy, shz;
              code_shift(a, code, shx, shy, shz);
              keep = keep_exact(a, xi, x[eu[u] & MDX_IDX_MASK], shx, shy, shz, code, ONE ? c_cut : s_cut[t]);
            }
          } else {
            double shx, shy, shz;
            code_shift(a, code, shx, shy, shz);
            const double c = ONE ? c_cut : s_cut[ti * a.ntypes1 + static_cast<int>(xj[u].w)];
            keep = valid && keep_exact(a, xi, xj[u], shx, shy, shz, code, c);
          }
          if (keep) {
            q0 = q1;
            q1 = q2;
            q2 = q3;
            q3 = static_cast<int>(eu[u]);
            cnt++;
            if ((cnt & 3) == 0 && cnt <= a.maxneigh) col[(size_t)((cnt - 4) >> 2) * a.npad] = make_int4(q0, q1, q2, q3);
          }
        }
      }
      nq = 0;
    }
  }
  const int tail = cnt & 3, t0 = cnt - tail;
  const int r0 = tail == 1 ? q3 : tail == 2 ? q2 : q1, r1 = tail == 2 ? q3 : q2;
  int *last = reinterpret_cast<int *>(col + (size_t)(t0 >> 2) * a.npad);
  if (tail > 0 && t0 < a.maxneigh) last[0] = r0;
  if (tail > 1 && t0 + 1 < a.maxneigh) last[1] = r1;
  if (tail > 2 && t0 + 2 < a.maxneigh) last[2] = q3;
  numneigh[i] = cnt;
  atomicMax(&dflags[0], cnt);
  atomicAdd(nsum, (unsigned long long)cnt);
}

// The fixed-point build needs the mirror (written by k_permute just before
// the build, in either precision, where its resolution fits the band; mixed
// precision also keeps it current between builds for its gathers) and a
// stencil that never holds one atom under two periodic images.
bool q_build_ok(const mdx_ctx *c) {
  if (!(c->qpos || c->qbuild) || !c->xq) return false;
  for (int d = 0; d < 3; d++)
    if (c->na.periodic[d] && 2 * c->na.S[d] + 1 > c->na.nbin[d]) return false;
  return true;
}

void launch_build(mdx_ctx *c) {
  const int B = 256;
  const int G = xcd_grid(((long)c->n + B - 1) / B);
  QStep qs{{(float)c->qs.step[0], (float)c->qs.step[1], (float)c->qs.step[2]}};
  unsigned long long *nsum = reinterpret_cast<unsigned long long *>(c->dflags + 8);
  const bool q = q_build_ok(c), one = c->ntypes == 1;
  auto go = [&](auto kernel, const int4 *xq) {
    hipLaunchKernelGGL(kernel, dim3(G), dim3(B), 0, c->stream, c->na, c->x, xq, qs, c->bkey_s, c->bin_start, c->nbr,
                       c->numneigh, c->dflags, nsum);
  };
  // the flat build takes stencils of at most 5x5 rows with at most two
  // pieces per row (no row wraps around the box at both ends: a row spans
  // 2S+1 bins).  Larger stencils come from bins thinner than half the
  // cutoff, in a box dimension below it (one bin): the per-row build.
  const NeighArgs &na = c->na;
  const bool flat = MDX_BUILD_FLAT && na.S[0] <= kFlatMaxS && na.S[1] <= kFlatMaxS && na.S[2] <= kFlatMaxS &&
                    (!na.periodic[0] || 2 * na.S[0] <= na.nbin[0]);
  if (flat) {
    FlatConst fc{};
    fc.bsy = static_cast<float>(1.0 / na.bininv[1]);
    fc.bsz = static_cast<float>(1.0 / na.bininv[2]);
    fc.binx = static_cast<float>(na.bininv[0]);
    fc.prune = (na.nbin[0] <= 8192 && na.nbin[1] <= 8192 && na.nbin[2] <= 8192) ? kPrune : 0;
    if (one) {  // type 1: the (1,1) cutoff, its band, and the row maximum of the table (stage_cmax)
      const int t11 = na.ntypes1 + 1;
      double m = 0.0;
      for (int u = 0; u < na.ntypes1; u++) m = std::fmax(m, na.cutneighsq[1 * na.ntypes1 + u]);
      fc.c_cut = na.cutneighsq[t11];
      fc.c_lo = static_cast<float>(na.cutneighsq[t11] * (1.0 - kBuildBand));
      fc.c_hi = static_cast<float>(na.cutneighsq[t11] * (1.0 + kBuildBand));
      fc.reach = static_cast<float>(m * (1.0 + 1e-4));
    }
    auto goflat = [&](auto kernel, const int4 *xq) {
      hipLaunchKernelGGL(kernel, dim3(G), dim3(B), 0, c->stream, c->na, c->x, xq, qs, c->bkey_s, c->bin_start,
                         c->nbr, c->numneigh, c->dflags, nsum, fc);
    };
    if (q && one) goflat(k_build_flat<true, true>, c->xq);
    else if (q) goflat(k_build_flat<true, false>, c->xq);
    else if (one) goflat(k_build_flat<false, true>, nullptr);
    else goflat(k_build_flat<false, false>, nullptr);
    return;
  }
  if (q && one) go(k_build_lane<true, true>, c->xq);
  else if (q) go(k_build_lane<true, false>, c->xq);
  else if (one) go(k_build_lane<false, true>, nullptr);
  else go(k_build_lane<false, false>, nullptr);
}

}  // namespace

int neigh_setup_grid(mdx_ctx *c) {
  NeighArgs &a = c->na;
  const double cut = c->cutneighmax;
  for (int d = 0; d < 3; d++) {
    if (c->periodic[d] && c->prd[d] < 2.0 * cut) {
      set_error("periodic box length below 2x the neighbour cutoff (one periodic image per pair required)");
      return MDX_ERR_ARG;
    }
  }
  // Neighbor::setup_bins: binsize = 0.5*cutneighmax, nbin = floor(bbox/binsize)
  // (the reference's grid, kept for the tie rule); the build's own grid is
  // cutneighmax / kBinDiv
  const double ref_bs = cut > 0.0 ? 0.5 * cut : c->prd[0];
  const double own_bs = cut > 0.0 ? cut / kBinDiv : c->prd[0];
  for (int d = 0; d < 3; d++) {
    int nr = static_cast<int>(c->prd[d] / ref_bs);
    if (nr == 0) nr = 1;
    int nb = static_cast<int>(c->prd[d] / own_bs);
    if (nb == 0) nb = 1;
    const double bs = c->prd[d] / nb;
    a.lo[d] = c->lo[d];
    a.prd[d] = c->prd[d];
    a.nbin[d] = nb;
    a.bininv[d] = 1.0 / bs;
    a.S[d] = stencil_extent(cut, bs, a.bininv[d]);
    a.periodic[d] = c->periodic[d];
    a.ref.bboxlo[d] = c->lo[d];
    a.ref.bboxhi[d] = c->hi[d];
    a.ref.bininv[d] = 1.0 / (c->prd[d] / nr);  // bininv = 1/binsize, rounded as the reference forms it
    a.ref.nbin[d] = nr;
  }
  long nb = (long)a.nbin[0] * a.nbin[1] * a.nbin[2];
  if (nb > (1L << 30)) {
    set_error("too many neighbour bins");
    return MDX_ERR_ARG;
  }
  c->nbins = (int)nb;
  if (c->nbins + 1 > c->cap_bins) {
    if (c->fixed_storage) {
      set_error("the box needs more neighbour bins than the fixed storage holds");
      return MDX_ERR_ARG;
    }
    (void)hipFree(c->bin_count);
    (void)hipFree(c->bin_start);
    c->cap_bins = c->nbins + 1;
    if (hipMalloc(&c->bin_count, sizeof(int) * c->cap_bins) != hipSuccess ||
        hipMalloc(&c->bin_start, sizeof(int) * c->cap_bins) != hipSuccess) {
      set_error("Insufficient memory on accelerator (bins)");
      return MDX_ERR_NOMEM;
    }
    (void)hipFree(c->scan_tmp);
    c->scan_tmp = nullptr;
    c->scan_tmp_bytes = 0;
  }
  a.ntypes1 = c->ntypes + 1;
  c->bins_dirty = true;  // a new grid starts from cleared counters
  return MDX_OK;
}

// (re)allocates the list with `cap` entries per atom
int alloc_list(mdx_ctx *c, int cap) {
  // a multiple of 4 entries: whole chunks of four (k_build_lane's stores)
  cap = (std::max(cap, 16) + 3) / 4 * 4;
  if (c->fixed_storage) {  // the list lives in storage this context does not own
    if ((long)cap * c->npad > c->nbr_cap_elems) {
      set_error("Neighbor list overflow: more neighbours per atom than the fixed storage holds");
      return MDX_ERR_OVERFLOW;
    }
    c->maxneigh = cap;
    c->nbr_elems = (long)cap * c->npad;
    return MDX_OK;
  }
  (void)hipFree(c->nbr);
  c->nbr = nullptr;
  if (hipMalloc(&c->nbr, sizeof(int) * (size_t)cap * c->npad) != hipSuccess) {
    set_error("Insufficient memory on accelerator (neighbour list)");
    return MDX_ERR_NOMEM;
  }
  c->maxneigh = cap;
  c->nbr_elems = (long)cap * c->npad;
  return MDX_OK;
}

int launch_list(mdx_ctx *c, bool again);

// The launches of a build: pbc wrap + bins, the stable counting sort, the
// permutation and the list (capacity maxneigh), then the build flags copied
// into pinned memory behind them.
int neigh_build_launch(mdx_ctx *c) {
  NeighArgs &a = c->na;
  a.n = c->n;
  a.npad = c->npad;
  c->built = false;  // a failed build must never leave a stale list in use
  if (c->n == 0) {
    c->built = true;
    return MDX_OK;
  }
  prof_begin(c, MDX_K_BUILD);
  const int B = 256;
  const int G = (c->n + B - 1) / B;
  hipError_t e;
  // Build flags: with two halves, build k uses half k mod 2, which build k-1
  // cleared (every reader of build k-1's flags precedes build k on the
  // stream), and clears the other half for build k+1.  A build that did not
  // finish leaves its half dirty: then both are cleared here.  Bin counters
  // likewise: cleared by the previous build's scatter pass, or here.
  int *next_flags = nullptr;
  if (c->dflags_base) {
    c->dflags = c->dflags_base + 16 * (c->build_serial & 1);
    next_flags = c->dflags_base + 16 * ((c->build_serial + 1) & 1);
    if (c->flags_dirty && (e = hipMemsetAsync(c->dflags_base, 0, 128, c->stream)) != hipSuccess)
      return hip_fail(e, "memset flags");
  } else if ((e = hipMemsetAsync(c->dflags, 0, 64, c->stream)) != hipSuccess) {
    return hip_fail(e, "memset flags");
  }
  c->flags_dirty = true;  // until this build's flags check passes
  if (c->bins_dirty && (e = hipMemsetAsync(c->bin_count, 0, sizeof(int) * (c->nbins + 1), c->stream)) != hipSuccess)
    return hip_fail(e, "memset bins");
  hipLaunchKernelGGL(k_wrap_bin, dim3(G), dim3(B), 0, c->stream, a, c->x, c->image, c->bkey, c->bidx,
                     c->bin_count, c->dflags);
  c->bins_dirty = true;  // until the scatter pass clears them
  // stable sort of (bin, current index): bin_start from the counts, then
  // scatter by arrival slot and rank within each bin (deterministic order)
  size_t need = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, need, c->bin_count, c->bin_start, c->nbins + 1, c->stream);
  if (need > c->scan_tmp_bytes) {
    if (c->fixed_storage) {
      set_error("bin scan storage exceeds the fixed storage");
      return MDX_ERR_ARG;
    }
    (void)hipFree(c->scan_tmp);
    c->scan_tmp_bytes = need;
    if (hipMalloc(&c->scan_tmp, need) != hipSuccess) {
      set_error("Insufficient memory on accelerator (scan)");
      return MDX_ERR_NOMEM;
    }
  }
  e = hipcub::DeviceScan::ExclusiveSum(c->scan_tmp, need, c->bin_count, c->bin_start, c->nbins + 1, c->stream);
  if (e != hipSuccess) return hip_fail(e, "bin scan");
  hipLaunchKernelGGL(k_bin_scatter, dim3(G), dim3(B), 0, c->stream, c->n, c->bkey, c->bidx, c->bin_start, c->bidx_s,
                     c->bin_count, c->nbins + 1, next_flags);
  c->bins_dirty = false;
  hipLaunchKernelGGL(k_bin_rank, dim3(G), dim3(B), 0, c->stream, c->n, c->bkey, c->bin_start, c->bidx_s, c->bidx,
                     c->bkey_s, c->one + 1, c->dflags);
  hipLaunchKernelGGL(k_permute, dim3(G), dim3(B), 0, c->stream, c->n, c->bidx, c->x, c->v, c->tag, c->orig,
                     c->image, c->xb, c->vb, c->tagb, c->origb, c->imageb, c->inv, c->xhold,
                     (c->qpos || c->qbuild) ? c->xq : nullptr, c->qs);
  if (c->qpos) c->xq_stale = false;
  std::swap(c->x, c->xb);
  std::swap(c->v, c->vb);
  std::swap(c->tag, c->tagb);
  std::swap(c->orig, c->origb);
  std::swap(c->image, c->imageb);
  if (!c->nbr) {
    int rc = alloc_list(c, c->maxneigh);
    if (rc) return rc;
  }
  return launch_list(c, false);
}

// the list pass of a build (the atoms are sorted), with its flags copied
// into pinned memory behind it; `again`: a repeat after a regrow, whose
// list flags are cleared first
int launch_list(mdx_ctx *c, bool again) {
  NeighArgs &a = c->na;
  hipError_t e;
  a.maxneigh = c->maxneigh;
  if (again) {
    if ((e = hipMemsetAsync(c->dflags, 0, 3 * sizeof(int), c->stream)) != hipSuccess)
      return hip_fail(e, "memset flags");
    if ((e = hipMemsetAsync(c->dflags + 8, 0, 8, c->stream)) != hipSuccess) return hip_fail(e, "memset flags");
  }
  launch_build(c);
  prof_end(c);
  if ((e = hipMemcpyAsync(c->bflags_host, c->dflags, 64, hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
    return hip_fail(e, "read build flags");
  if ((e = hipEventRecord(c->build_ev, c->stream)) != hipSuccess) return hip_fail(e, "build event");
  return MDX_OK;
}

// Waits for the flags of the last list pass and checks them: errors as the
// reference reports them, else *regrow when the largest row exceeded the
// capacity (the capacity is grown here, the caller lists again).
int neigh_build_finish(mdx_ctx *c, bool *regrow) {
  *regrow = false;
  if (c->n == 0) return MDX_OK;
  hipError_t e = hipEventSynchronize(c->build_ev);
  if (e != hipSuccess) return hip_fail(e, "neighbour build");
  int flags[16];
  memcpy(flags, c->bflags_host, sizeof(flags));
  if (flags[4]) {
    set_error("Lost atoms: an atom left a non-periodic box");
    return MDX_ERR_ARG;
  }
  const int mx = flags[0];
  if (mx > c->one || flags[5]) {  // flags[5]: a bin holds more than one + 1 atoms (k_bin_rank)
    set_error("Neighbor list overflow, boost neigh_modify one");
    return MDX_ERR_OVERFLOW;
  }
  unsigned long long tot;
  memcpy(&tot, &flags[8], sizeof(tot));
  c->total_neighbors = (long)tot;
  c->max_neighbors = mx;
  if (mx > c->maxneigh) {
    // regrow: the accelerator library's max_nbors resize (lal_neighbor.cpp:582-595)
    int cap = ((mx + mx / 4 + 15) / 16) * 16;
    cap = std::min(cap, c->one);
    int rc = alloc_list(c, cap);
    if (rc) return rc;
    *regrow = true;
    return MDX_OK;
  }
  c->built = true;
  c->flags_dirty = false;
  c->build_serial++;
  c->ago = 0;
  return MDX_OK;
}

// list build; regrow the per-atom capacity and list again on overflow
int neigh_build(mdx_ctx *c) {
  int rc = neigh_build_launch(c);
  for (int attempt = 0; !rc && attempt < 4; attempt++) {
    bool regrow = false;
    if ((rc = neigh_build_finish(c, &regrow)) || !regrow) break;
    prof_begin(c, MDX_K_BUILD);
    rc = launch_list(c, true);
  }
  if (!rc && !c->built) {
    set_error("neighbour list capacity could not be grown to fit");
    return MDX_ERR_OVERFLOW;
  }
  return rc;
}

}  // namespace mdx
