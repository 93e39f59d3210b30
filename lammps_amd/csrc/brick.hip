// brick.hip - LDS-staged ("bin brick") neighbour build and pair kernels for gfx950.
//
// Why: with atoms bin-sorted, an atom's ~78 neighbours still lie in 25
// different bin rows, so a gather-per-neighbour kernel touches ~50 cache
// lines per atom and is bound by L1/L2 line throughput, not by HBM.  Here
// one workgroup owns a brick of R0 x R1 x R2 bins.  It stages the positions
// of every atom in the brick's halo (R+2S bins per dimension, S = stencil
// extent) into LDS once, with the periodic image shift already applied
// (x_j + s*L, rounded exactly as CommBrick::borders forms a ghost,
// src/atom_vec_atomic.cpp:290-297), then all pair work reads LDS.
//
// LDS layout: halo rows (y,z) in order, each row's window bins in x order,
// each bin's atoms in sorted order.  For an atom in local bin wx of a brick
// row, the candidates of stencil row (dy,dz) are ONE contiguous LDS range
// [off[row'][wx-S0], off[row'][wx+S0+1]), wrapped bins included, in the
// same order as the reference's stencil scan.  The neighbour list therefore
// stores 16-bit LDS slot numbers (half the bytes of global indices), which
// the pair kernels resolve against the same staging.
//
// The list build (Neighbor::build, src/neigh_half_bin.cpp:285-419) keeps j
// iff rsq <= cutneighsq with rsq in FP64, no contraction, on the staged
// image coordinate; for a periodic pair within a hair of the cutoff it also
// evaluates the mirrored orientation and takes the reference's side
// (half_bin_newton evaluates an owned-ghost pair once).  A wavefront ballot
// and a popcount prefix keep the sequential scan order for any TPA.
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kNT2 = MDX_NT1 * MDX_NT1;
constexpr int kMaxEntries = 4096;  // halo rows x window bins per brick
constexpr int kUnroll = 4;         // list entries loaded ahead per lane
static_assert(kUnroll == 4, "k_lj_brick walks chunks of four entries, as k_lj does");

__device__ __forceinline__ double rsq_exact(double a0, double a1, double a2, double b0, double b1, double b2) {
#pragma clang fp contract(off)
  const double dx = a0 - b0;
  const double dy = a1 - b1;
  const double dz = a2 - b2;
  return dx * dx + dy * dy + dz * dz;
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

// upper-half-stencil test of the reference (src/neigh_stencil.cpp:143-158)
__device__ bool ref_owner_side(const RefBins &r, double xi0, double xi1, double xi2, double g0, double g1,
                               double g2) {
  const int di = ref_coord2bin(g0, r.bboxlo[0], r.bboxhi[0], r.bininv[0], r.nbin[0]) -
                 ref_coord2bin(xi0, r.bboxlo[0], r.bboxhi[0], r.bininv[0], r.nbin[0]);
  const int dj = ref_coord2bin(g1, r.bboxlo[1], r.bboxhi[1], r.bininv[1], r.nbin[1]) -
                 ref_coord2bin(xi1, r.bboxlo[1], r.bboxhi[1], r.bininv[1], r.nbin[1]);
  const int dk = ref_coord2bin(g2, r.bboxlo[2], r.bboxhi[2], r.bininv[2], r.nbin[2]) -
                 ref_coord2bin(xi2, r.bboxlo[2], r.bboxhi[2], r.bininv[2], r.nbin[2]);
  return dk > 0 || (dk == 0 && (dj > 0 || (dj == 0 && di > 0)));
}

// Per-workgroup view of one brick after staging.
struct Staged {
  int nx, ny, nz;  // brick extent in bins (clipped at the grid edge)
  int wx, wy, wz;  // halo extent: n + 2S
  int total;       // staged atoms
  int ni;          // atoms owned by the brick
};

// Block-wide exclusive scan of s_cnt[0..E) into s_off[0..E] (s_off[E] = total).
__device__ void block_exclusive_scan(const int *s_cnt, int *s_off, int E, int *s_tmp) {
  const int T = blockDim.x, t = threadIdx.x;
  const int per = (E + T - 1) / T;
  const int b0 = min(t * per, E), b1 = min(b0 + per, E);
  int local = 0;
  for (int e = b0; e < b1; e++) local += s_cnt[e];
  s_tmp[t] = local;
  __syncthreads();
  // Hillis-Steele over T partial sums
  for (int off = 1; off < T; off <<= 1) {
    const int v = t >= off ? s_tmp[t - off] : 0;
    __syncthreads();
    s_tmp[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? s_tmp[t - 1] : 0;
  for (int e = b0; e < b1; e++) {
    s_off[e] = run;
    run += s_cnt[e];
  }
  if (t == T - 1) s_off[E] = s_tmp[T - 1];
  __syncthreads();
}

// Maps halo entry e (row-major: rz, ry, wxi) of brick `brick` to a bin and
// its periodic image (sx, sy, sz); returns false when the entry lies outside
// a non-periodic box or past the clipped brick window.
__device__ __host__ inline bool brick_entry(const NeighArgs &a, const BrickArgs &br, int brick, int e, int &bin,
                                            int &sx, int &sy, int &sz) {
  const int bxn = br.nbrick[0], byn = br.nbrick[1];
  const int bx = brick % bxn, by = (brick / bxn) % byn, bz = brick / (bxn * byn);
  const int x0 = bx * br.R[0], y0 = by * br.R[1], z0 = bz * br.R[2];
  const int nx = min(br.R[0], a.nbin[0] - x0), ny = min(br.R[1], a.nbin[1] - y0), nz = min(br.R[2], a.nbin[2] - z0);
  const int wx = nx + 2 * a.S[0], wy = ny + 2 * a.S[1], wz = nz + 2 * a.S[2];
  const int Wx = br.R[0] + 2 * a.S[0], Wy = br.R[1] + 2 * a.S[1];
  const int xi = e % Wx, yi = (e / Wx) % Wy, zi = e / (Wx * Wy);
  if (xi >= wx || yi >= wy || zi >= wz) return false;
  int c[3] = {x0 - a.S[0] + xi, y0 - a.S[1] + yi, z0 - a.S[2] + zi};
  int s[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) {
    if (c[d] < 0) {
      if (!a.periodic[d]) return false;
      c[d] += a.nbin[d];
      s[d] = -1;
    } else if (c[d] >= a.nbin[d]) {
      if (!a.periodic[d]) return false;
      c[d] -= a.nbin[d];
      s[d] = 1;
    }
  }
  bin = bin_index(a, c[0], c[1], c[2]);
  sx = s[0];
  sy = s[1];
  sz = s[2];
  return true;
}

// Stages brick `brick`: fills s_off (entry offsets), s_x (image positions),
// s_g (global index), and optionally s_code (image code) and s_fp.
template <bool CODE, bool FP>
__device__ Staged stage_brick(const NeighArgs &a, const BrickArgs &br, int brick, const int *__restrict__ bin_start,
                              const double4 *__restrict__ x, const double *__restrict__ fp, int *s_cnt, int *s_off,
                              int *s_tmp, double4 *s_x, int *s_g, unsigned char *s_code, double *s_fp,
                              int *__restrict__ overflow) {
  Staged st;
  const int bxn = br.nbrick[0], byn = br.nbrick[1];
  const int bx = brick % bxn, by = (brick / bxn) % byn, bz = brick / (bxn * byn);
  st.nx = min(br.R[0], a.nbin[0] - bx * br.R[0]);
  st.ny = min(br.R[1], a.nbin[1] - by * br.R[1]);
  st.nz = min(br.R[2], a.nbin[2] - bz * br.R[2]);
  st.wx = st.nx + 2 * a.S[0];
  st.wy = st.ny + 2 * a.S[1];
  st.wz = st.nz + 2 * a.S[2];
  const int Wx = br.R[0] + 2 * a.S[0], Wy = br.R[1] + 2 * a.S[1], Wz = br.R[2] + 2 * a.S[2];
  const int E = Wx * Wy * Wz;
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int bin, sx, sy, sz;
    s_cnt[e] = brick_entry(a, br, brick, e, bin, sx, sy, sz) ? bin_start[bin + 1] - bin_start[bin] : 0;
  }
  __syncthreads();
  block_exclusive_scan(s_cnt, s_off, E, s_tmp);
  st.total = s_off[E];
  const bool fits = st.total <= br.max_slots;
  if (!fits && threadIdx.x == 0) atomicMax(overflow, st.total);
  if (fits) {
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
      int bin, sx, sy, sz;
      if (!brick_entry(a, br, brick, e, bin, sx, sy, sz)) continue;
      const int j0 = bin_start[bin], cnt = s_cnt[e], base = s_off[e];
      const double shx = sx * a.prd[0], shy = sy * a.prd[1], shz = sz * a.prd[2];
      for (int k = 0; k < cnt; k++) {
        const double4 p = x[j0 + k];
        s_x[base + k] = make_double4(p.x + shx, p.y + shy, p.z + shz, p.w);
        s_g[base + k] = j0 + k;
        if (CODE) s_code[base + k] = (unsigned char)((sx + 1) + 3 * (sy + 1) + 9 * (sz + 1));
        if (FP) s_fp[base + k] = fp[j0 + k];
      }
    }
  }
  // atoms owned by the brick: its own (non-halo) rows and bins
  st.ni = 0;
  if (fits) {
    for (int lz = 0; lz < st.nz; lz++)
      for (int ly = 0; ly < st.ny; ly++) {
        const int row = ((lz + a.S[2]) * Wy + (ly + a.S[1])) * Wx;
        st.ni += s_off[row + a.S[0] + st.nx] - s_off[row + a.S[0]];
      }
  }
  __syncthreads();
  st.total = fits ? st.total : -1;
  return st;
}

// v-th owned atom of the brick -> its LDS slot and local bin coordinates
__device__ __forceinline__ int owned_slot(const NeighArgs &a, const BrickArgs &br, const Staged &st,
                                          const int *s_off, int v, int &ly, int &lz, int &wxi) {
  const int Wx = br.R[0] + 2 * a.S[0], Wy = br.R[1] + 2 * a.S[1];
  for (lz = 0; lz < st.nz; lz++)
    for (ly = 0; ly < st.ny; ly++) {
      const int row = ((lz + a.S[2]) * Wy + (ly + a.S[1])) * Wx;
      const int lo = s_off[row + a.S[0]], hi = s_off[row + a.S[0] + st.nx];
      if (v < hi - lo) {
        const int slot = lo + v;
        // local window bin of the slot (binary search over the row's bins)
        int b0 = a.S[0], b1 = a.S[0] + st.nx - 1;
        while (b0 < b1) {
          const int mid = (b0 + b1 + 1) >> 1;
          if (s_off[row + mid] <= slot) b0 = mid;
          else b1 = mid - 1;
        }
        wxi = b0;
        return slot;
      }
      v -= hi - lo;
    }
  return -1;
}

struct BrickShared {
  int *cnt, *off, *tmp;
  double4 *x;
  int *g;
  unsigned char *code;
  double *fp;
};

// dynamic LDS carve-up: [x: max_slots double4][fp][g: int][cnt][off][tmp][code]
__device__ BrickShared carve(unsigned char *base, const BrickArgs &br, bool with_fp, bool with_code, int E) {
  BrickShared s;
  size_t p = 0;
  s.x = reinterpret_cast<double4 *>(base + p);
  p += sizeof(double4) * br.max_slots;
  s.fp = reinterpret_cast<double *>(base + p);
  if (with_fp) p += sizeof(double) * br.max_slots;
  s.g = reinterpret_cast<int *>(base + p);
  p += sizeof(int) * br.max_slots;
  s.cnt = reinterpret_cast<int *>(base + p);
  p += sizeof(int) * E;
  s.off = reinterpret_cast<int *>(base + p);
  p += sizeof(int) * (E + 1);
  s.tmp = reinterpret_cast<int *>(base + p);
  p += sizeof(int) * 1024;
  s.code = base + p;
  (void)with_code;
  return s;
}

// ------------------------------------------------------------------ build
template <int TPA>
__global__ __launch_bounds__(256) void k_build_brick(NeighArgs a, BrickArgs br, const double4 *__restrict__ x,
                                                    const int *__restrict__ bin_start,
                                                    unsigned short *__restrict__ nbr, int *__restrict__ numneigh,
                                                    int *__restrict__ dflags, unsigned long long *__restrict__ nsum) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int Wx = br.R[0] + 2 * a.S[0], Wy = br.R[1] + 2 * a.S[1], Wz = br.R[2] + 2 * a.S[2];
  const int E = Wx * Wy * Wz;
  BrickShared s = carve(lds, br, false, true, E);
  const Staged st = stage_brick<true, false>(a, br, blockIdx.x, bin_start, x, nullptr, s.cnt, s.off, s.tmp, s.x,
                                             s.g, s.code, nullptr, &dflags[5]);
  if (st.total < 0) return;
  const int lane = threadIdx.x % TPA;
  const int group = threadIdx.x / TPA, ngroups = blockDim.x / TPA;
  const int gshift = (threadIdx.x & 63) - lane;
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long gmask = (TPA == 64) ? ~0ull : ((1ull << TPA) - 1ull);
  int local_max = 0;
  unsigned long long local_sum = 0;
  for (int v = group; v < st.ni; v += ngroups) {
    int ly, lz, wxi;
    const int si = owned_slot(a, br, st, s.off, v, ly, lz, wxi);
    const double4 xi = s.x[si];
    const int gi = s.g[si];
    const int ti = static_cast<int>(xi.w);
    unsigned short *row = nbr + (size_t)gi * br.cap;
    int cnt = 0;
    for (int dz = -a.S[2]; dz <= a.S[2]; dz++)
      for (int dy = -a.S[1]; dy <= a.S[1]; dy++) {
        const int r = ((lz + a.S[2] + dz) * Wy + (ly + a.S[1] + dy)) * Wx;
        const int lo = s.off[r + wxi - a.S[0]], hi = s.off[r + wxi + a.S[0] + 1];
        for (int base = lo; base < hi; base += TPA) {
          const int sj = base + lane;
          bool keep = false;
          if (sj < hi && sj != si) {
            const double4 g = s.x[sj];
            const double c = a.cutneighsq[ti * a.ntypes1 + static_cast<int>(g.w)];
            const double rsq = rsq_exact(xi.x, xi.y, xi.z, g.x, g.y, g.z);
            keep = rsq <= c;
            const int code = s.code[sj];
            if (code != MDX_NOSHIFT && fabs(rsq - c) <= 1e-12 * c) {
              // mirrored orientation x_j - (x_i - s*L) on the unshifted x_j
              const double4 xj = x[s.g[sj]];
              const double shx = ((code % 3) - 1) * a.prd[0], shy = (((code / 3) % 3) - 1) * a.prd[1],
                           shz = ((code / 9) - 1) * a.prd[2];
              const bool keep_b = rsq_exact(xj.x, xj.y, xj.z, xi.x - shx, xi.y - shy, xi.z - shz) <= c;
              if (keep_b != keep && !ref_owner_side(a.ref, xi.x, xi.y, xi.z, g.x, g.y, g.z)) keep = keep_b;
            }
          }
          const unsigned long long m = (__ballot(keep) >> gshift) & gmask;
          if (keep) {
            const int pos = cnt + __popcll(m & below);
            if (pos < br.cap) row[pos] = (unsigned short)sj;
          }
          cnt += __popcll(m);
        }
      }
    if (lane == 0) {
      numneigh[gi] = cnt;
      local_max = max(local_max, cnt);
      local_sum += (unsigned long long)cnt;
    }
  }
  if (lane == 0 && local_sum) {
    atomicMax(&dflags[0], local_max);
    atomicAdd(nsum, local_sum);
  }
}

// number of atoms each brick stages (sizes the LDS of the brick kernels)
__global__ void k_brick_count(NeighArgs a, BrickArgs br, const int *__restrict__ bin_start, int *__restrict__ out) {
  const int brick = blockIdx.x * blockDim.x + threadIdx.x;
  if (brick >= br.nbricks) return;
  const int Wx = br.R[0] + 2 * a.S[0], Wy = br.R[1] + 2 * a.S[1], Wz = br.R[2] + 2 * a.S[2];
  int total = 0;
  for (int e = 0; e < Wx * Wy * Wz; e++) {
    int bin, sx, sy, sz;
    if (brick_entry(a, br, brick, e, bin, sx, sy, sz)) total += bin_start[bin + 1] - bin_start[bin];
  }
  atomicMax(out, total);
}

// ------------------------------------------------------------------ LJ
template <int TPA>
__device__ __forceinline__ double gsum(double v) {
#pragma unroll
  for (int off = TPA / 2; off > 0; off >>= 1) v += __shfl_down(v, off, TPA);
  return v;
}

template <int NV>
__device__ __forceinline__ void block_reduce_accum(double (&v)[NV], double *__restrict__ out) {
  __shared__ double s[NV][16];
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) s[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    const int nw = (blockDim.x + 63) >> 6;
    double acc = 0.0;
    for (int q = 0; q < nw; q++) acc += s[threadIdx.x][q];
    out[(size_t)blockIdx.x * NV + threadIdx.x] = acc;
  }
}

template <bool MIXED, bool EV, int TPA>
__global__ __launch_bounds__(256) void k_lj_brick(NeighArgs a, BrickArgs br, const double4 *__restrict__ x,
                                                 const int *__restrict__ bin_start,
                                                 const unsigned short *__restrict__ nbr,
                                                 const int *__restrict__ numneigh, const LJTab *__restrict__ tab,
                                                 double4 *__restrict__ f, double *__restrict__ vatom, int npad,
                                                 double *__restrict__ partials, int *__restrict__ dflags) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ double s_cutsq[kNT2], s_lj1[kNT2], s_lj2[kNT2], s_lj3[kNT2], s_lj4[kNT2], s_off[kNT2];
  __shared__ float f_lj1[kNT2], f_lj2[kNT2], f_lj3[kNT2], f_lj4[kNT2], f_off[kNT2];
  const int nt1 = a.ntypes1, nt2 = nt1 * nt1;
  for (int q = threadIdx.x; q < nt2; q += blockDim.x) {
    s_cutsq[q] = tab->cutsq[q];
    s_lj1[q] = tab->lj1[q];
    s_lj2[q] = tab->lj2[q];
    s_lj3[q] = tab->lj3[q];
    s_lj4[q] = tab->lj4[q];
    s_off[q] = tab->offset[q];
    f_lj1[q] = (float)tab->lj1[q];
    f_lj2[q] = (float)tab->lj2[q];
    f_lj3[q] = (float)tab->lj3[q];
    f_lj4[q] = (float)tab->lj4[q];
    f_off[q] = (float)tab->offset[q];
  }
  const int Wx = br.R[0] + 2 * a.S[0], Wy = br.R[1] + 2 * a.S[1], Wz = br.R[2] + 2 * a.S[2];
  const int E = Wx * Wy * Wz;
  BrickShared s = carve(lds, br, false, false, E);
  const Staged st = stage_brick<false, false>(a, br, blockIdx.x, bin_start, x, nullptr, s.cnt, s.off, s.tmp, s.x,
                                              s.g, nullptr, nullptr, &dflags[5]);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (st.total >= 0) {
    const int lane = threadIdx.x % TPA;
    const int group = threadIdx.x / TPA, ngroups = blockDim.x / TPA;
    const int niter = (st.ni + ngroups - 1) / ngroups;  // uniform trip count: shuffles need whole groups
    for (int it = 0; it < niter; it++) {
      const int v = group + it * ngroups;
      double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
      double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      int gi = -1;
      if (v < st.ni) {
        int ly, lz, wxi;
        const int si = owned_slot(a, br, st, s.off, v, ly, lz, wxi);
        const double4 xi = s.x[si];
        gi = s.g[si];
        const int row = static_cast<int>(xi.w) * nt1;
        const int jn = numneigh[gi];
        const unsigned short *nb = nbr + (size_t)gi * br.cap;
        // chunks of four consecutive entries per lane, k_lj's order
        for (int k0 = 4 * lane; k0 < jn; k0 += kUnroll * TPA) {
          int sl[kUnroll];
#pragma unroll
          for (int u = 0; u < kUnroll; u++) sl[u] = (k0 + u < jn) ? (int)nb[k0 + u] : -1;
#pragma unroll
          for (int u = 0; u < kUnroll; u++) {
          if (sl[u] < 0) continue;
          const double4 xj = s.x[sl[u]];
          const double delx = xi.x - xj.x, dely = xi.y - xj.y, delz = xi.z - xj.z;
          const double rsq = delx * delx + dely * dely + delz * delz;
          const int t = row + static_cast<int>(xj.w);
          if (rsq < s_cutsq[t]) {
            double fpair, evdwl = 0.0;
            if (MIXED) {
              const float r2inv = __builtin_amdgcn_rcpf((float)rsq);
              const float r6inv = r2inv * r2inv * r2inv;
              fpair = (double)(r6inv * (f_lj1[t] * r6inv - f_lj2[t]) * r2inv);
              if (EV) evdwl = (double)(r6inv * (f_lj3[t] * r6inv - f_lj4[t]) - f_off[t]);
            } else {
              const double r2inv = 1.0 / rsq;
              const double r6inv = r2inv * r2inv * r2inv;
              fpair = r6inv * (s_lj1[t] * r6inv - s_lj2[t]) * r2inv;
              if (EV) evdwl = r6inv * (s_lj3[t] * r6inv - s_lj4[t]) - s_off[t];
            }
            fx += delx * fpair;
            fy += dely * fpair;
            fz += delz * fpair;
            if (EV) {
              e += 0.5 * evdwl;
              const double hf = 0.5 * fpair;
              w[0] += delx * delx * hf;
              w[1] += dely * dely * hf;
              w[2] += delz * delz * hf;
              w[3] += delx * dely * hf;
              w[4] += delx * delz * hf;
              w[5] += dely * delz * hf;
            }
          }
          }
        }
      }
      fx = gsum<TPA>(fx);
      fy = gsum<TPA>(fy);
      fz = gsum<TPA>(fz);
      if (EV) {
        e = gsum<TPA>(e);
#pragma unroll
        for (int q = 0; q < 6; q++) w[q] = gsum<TPA>(w[q]);
      }
      if (gi >= 0 && lane == 0) {
        f[gi] = make_double4(fx, fy, fz, e);
        if (EV) {
#pragma unroll
          for (int q = 0; q < 6; q++) vatom[(size_t)q * npad + gi] = w[q];
          acc[0] += e;
#pragma unroll
          for (int q = 0; q < 6; q++) acc[1 + q] += w[q];
        }
      }
    }
  }
  if (EV) block_reduce_accum<7>(acc, partials);
}

// ------------------------------------------------------------------ EAM
template <bool MIXED, bool EV, int TPA>
__global__ __launch_bounds__(256) void k_eam_rho_brick(NeighArgs a, BrickArgs br, EAMArgs ea,
                                                      const double4 *__restrict__ x,
                                                      const int *__restrict__ bin_start,
                                                      const unsigned short *__restrict__ nbr,
                                                      const int *__restrict__ numneigh, double *__restrict__ rho_out,
                                                      double *__restrict__ fp_out, double *__restrict__ eemb,
                                                      int *__restrict__ dflags) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int Wx = br.R[0] + 2 * a.S[0], Wy = br.R[1] + 2 * a.S[1], Wz = br.R[2] + 2 * a.S[2];
  const int E = Wx * Wy * Wz;
  BrickShared s = carve(lds, br, false, false, E);
  const Staged st = stage_brick<false, false>(a, br, blockIdx.x, bin_start, x, nullptr, s.cnt, s.off, s.tmp, s.x,
                                              s.g, nullptr, nullptr, &dflags[5]);
  if (st.total < 0) return;
  const int lane = threadIdx.x % TPA;
  const int group = threadIdx.x / TPA, ngroups = blockDim.x / TPA;
  const int niter = (st.ni + ngroups - 1) / ngroups;
  for (int it = 0; it < niter; it++) {
    const int v = group + it * ngroups;
    double rho = 0.0;
    int gi = -1, ti = 0;
    if (v < st.ni) {
      int ly, lz, wxi;
      const int si = owned_slot(a, br, st, s.off, v, ly, lz, wxi);
      const double4 xi = s.x[si];
      gi = s.g[si];
      ti = static_cast<int>(xi.w);
      const int jn = numneigh[gi];
      const unsigned short *nb = nbr + (size_t)gi * br.cap;
      for (int k = lane; k < jn; k += TPA) {
        const double4 xj = s.x[nb[k]];
        const double delx = xi.x - xj.x, dely = xi.y - xj.y, delz = xi.z - xj.z;
        const double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < ea.cutforcesq) {
          rho += eam_rho_term<MIXED>(ea, ea.type2rhor[static_cast<int>(xj.w) * ea.ntypes1 + ti], rsq);
        }
      }
    }
    rho = gsum<TPA>(rho);
    if (gi >= 0 && lane == 0) {
      // fp pass (pair_eam.cpp:213-228), FP64 in both modes: one per atom
      double p = rho * ea.rdrho + 1.0;
      int m = static_cast<int>(p);
      m = max(1, min(m, ea.nrho - 1));
      p -= m;
      p = fmin(p, 1.0);
      const double *c = ea.frho + ((size_t)ea.type2frho[ti] * (ea.nrho + 1) + m) * 7;
      const double fpi = (c[0] * p + c[1]) * p + c[2];
      rho_out[gi] = rho;
      fp_out[gi] = fpi;
      if (EV) {
        double phi = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
        if (rho > ea.rhomax) phi += fpi * (rho - ea.rhomax);
        eemb[gi] = phi;
      }
    }
  }
}

template <bool MIXED, bool EV, int TPA>
__global__ __launch_bounds__(256) void k_eam_force_brick(NeighArgs a, BrickArgs br, EAMArgs ea,
                                                        const double4 *__restrict__ x,
                                                        const int *__restrict__ bin_start,
                                                        const unsigned short *__restrict__ nbr,
                                                        const int *__restrict__ numneigh,
                                                        const double *__restrict__ fp, const double *__restrict__ eemb,
                                                        double4 *__restrict__ f, double *__restrict__ vatom, int npad,
                                                        double *__restrict__ partials, int *__restrict__ dflags) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int Wx = br.R[0] + 2 * a.S[0], Wy = br.R[1] + 2 * a.S[1], Wz = br.R[2] + 2 * a.S[2];
  const int E = Wx * Wy * Wz;
  BrickShared s = carve(lds, br, true, false, E);
  const Staged st = stage_brick<false, true>(a, br, blockIdx.x, bin_start, x, fp, s.cnt, s.off, s.tmp, s.x, s.g,
                                             nullptr, s.fp, &dflags[5]);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (st.total >= 0) {
    const int lane = threadIdx.x % TPA;
    const int group = threadIdx.x / TPA, ngroups = blockDim.x / TPA;
    const int niter = (st.ni + ngroups - 1) / ngroups;
    for (int it = 0; it < niter; it++) {
      const int v = group + it * ngroups;
      double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
      double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      int gi = -1;
      if (v < st.ni) {
        int ly, lz, wxi;
        const int si = owned_slot(a, br, st, s.off, v, ly, lz, wxi);
        const double4 xi = s.x[si];
        gi = s.g[si];
        const int ti = static_cast<int>(xi.w);
        const double fpi = s.fp[si];
        const int jn = numneigh[gi];
        const unsigned short *nb = nbr + (size_t)gi * br.cap;
        for (int k = lane; k < jn; k += TPA) {
          const int sj = nb[k];
          const double4 xj = s.x[sj];
          const double delx = xi.x - xj.x, dely = xi.y - xj.y, delz = xi.z - xj.z;
          const double rsq = delx * delx + dely * dely + delz * delz;
          if (rsq < ea.cutforcesq) {
            const int tj = static_cast<int>(xj.w);
            const int tij = ti * ea.ntypes1 + tj, tji = tj * ea.ntypes1 + ti;
            double fpair, phi;
            eam_pair_term<MIXED>(ea, ea.type2rhor[tij], ea.type2rhor[tji], ea.type2z2r[tij], rsq, fpi, s.fp[sj], fpair,
                                 phi);
            fx += delx * fpair;
            fy += dely * fpair;
            fz += delz * fpair;
            if (EV) {
              e += 0.5 * phi;
              const double hf = 0.5 * fpair;
              w[0] += delx * delx * hf;
              w[1] += dely * dely * hf;
              w[2] += delz * delz * hf;
              w[3] += delx * dely * hf;
              w[4] += delx * delz * hf;
              w[5] += dely * delz * hf;
            }
          }
        }
        if (EV && lane == 0) e += eemb[gi];  // embedding term joins lane 0's partial sum
      }
      fx = gsum<TPA>(fx);
      fy = gsum<TPA>(fy);
      fz = gsum<TPA>(fz);
      if (EV) {
        e = gsum<TPA>(e);
#pragma unroll
        for (int q = 0; q < 6; q++) w[q] = gsum<TPA>(w[q]);
      }
      if (gi >= 0 && lane == 0) {
        f[gi] = make_double4(fx, fy, fz, e);
        if (EV) {
#pragma unroll
          for (int q = 0; q < 6; q++) vatom[(size_t)q * npad + gi] = w[q];
          acc[0] += e;
#pragma unroll
          for (int q = 0; q < 6; q++) acc[1 + q] += w[q];
        }
      }
    }
  }
  if (EV) block_reduce_accum<7>(acc, partials);
}

size_t lds_bytes(const NeighArgs &a, const BrickArgs &br, bool with_fp) {
  const int E = (br.R[0] + 2 * a.S[0]) * (br.R[1] + 2 * a.S[1]) * (br.R[2] + 2 * a.S[2]);
  size_t b = sizeof(double4) * br.max_slots + sizeof(int) * br.max_slots + sizeof(int) * (2 * E + 1) +
             sizeof(int) * 1024 + br.max_slots;
  if (with_fp) b += sizeof(double) * br.max_slots;
  return (b + 15) & ~size_t(15);
}

}  // namespace

// Sizes the bricks for the current bins: picks the brick shape (user request
// or default), counts the staged atoms of every brick and shrinks the brick
// until the staging fits the LDS budget.
int brick_prepare(mdx_ctx *c) {
  NeighArgs &a = c->na;
  BrickArgs &br = c->br;
  int R[3] = {c->brick_req[0], c->brick_req[1], c->brick_req[2]};
  const size_t budget = c->lds_budget;
  for (int attempt = 0; attempt < 12; attempt++) {
    int E = 1;
    for (int d = 0; d < 3; d++) {
      R[d] = std::max(1, std::min(R[d], a.nbin[d]));
      br.R[d] = R[d];
      br.nbrick[d] = (a.nbin[d] + R[d] - 1) / R[d];
      E *= R[d] + 2 * a.S[d];
    }
    br.nbricks = br.nbrick[0] * br.nbrick[1] * br.nbrick[2];
    if (E > kMaxEntries) {
      int dmax = R[0] >= R[1] && R[0] >= R[2] ? 0 : (R[1] >= R[2] ? 1 : 2);
      if (R[dmax] == 1) {
        set_error("neighbour stencil too wide for LDS staging");
        return MDX_ERR_ARG;
      }
      R[dmax] = (R[dmax] + 1) / 2;
      continue;
    }
    hipError_t e = hipMemsetAsync(c->dflags + 6, 0, sizeof(int), c->stream);
    if (e != hipSuccess) return hip_fail(e, "brick count");
    hipLaunchKernelGGL(k_brick_count, dim3((br.nbricks + 255) / 256), dim3(256), 0, c->stream, a, br,
                       c->bin_start, c->dflags + 6);
    int mx = 0;
    if ((e = hipMemcpyAsync(&mx, c->dflags + 6, sizeof(int), hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess)
      return hip_fail(e, "brick count");
    // headroom for atoms moving between bins is not needed: bins are fixed until the next build
    br.max_slots = std::max(mx, 1);
    if (lds_bytes(a, br, true) <= budget) return MDX_OK;
    int dmax = R[0] >= R[1] && R[0] >= R[2] ? 0 : (R[1] >= R[2] ? 1 : 2);
    if (R[dmax] == 1) {
      set_error("bins too crowded for LDS staging (density too high for the brick kernels)");
      return MDX_ERR_ARG;
    }
    R[dmax] = (R[dmax] + 1) / 2;
  }
  set_error("could not size the LDS bricks");
  return MDX_ERR_ARG;
}

template <int TPA>
static void launch_build_brick(mdx_ctx *c) {
  const size_t sh = lds_bytes(c->na, c->br, false);
  hipLaunchKernelGGL((k_build_brick<TPA>), dim3(c->br.nbricks), dim3(256), sh, c->stream, c->na, c->br, c->x,
                     c->bin_start, c->nbr16, c->numneigh, c->dflags,
                     reinterpret_cast<unsigned long long *>(c->dflags + 8));
}

int brick_build_list(mdx_ctx *c) {
  c->br.cap = c->maxneigh;
  switch (c->launch.build_tpa) {
    case 1: launch_build_brick<1>(c); break;
    case 4: launch_build_brick<4>(c); break;
    case 8: launch_build_brick<8>(c); break;
    case 16: launch_build_brick<16>(c); break;
    case 32: launch_build_brick<32>(c); break;
    default:
      set_error("build threads per atom must be one of 1, 4, 8, 16, 32");
      return MDX_ERR_ARG;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MDX_OK : hip_fail(e, "brick build launch");
}

template <bool MIXED, bool EV, int TPA>
static void launch_pair_brick(mdx_ctx *c) {
  const int B = c->launch.block;
  if (c->pair_style == MDX_PAIR_LJ) {
    const size_t sh = lds_bytes(c->na, c->br, false);
    prof_begin(c, MDX_K_LJ);
    hipLaunchKernelGGL((k_lj_brick<MIXED, EV, TPA>), dim3(c->br.nbricks), dim3(B), sh, c->stream, c->na, c->br,
                       c->x, c->bin_start, c->nbr16, c->numneigh, c->lj_dev, c->f, c->vatom, c->npad, c->partials,
                       c->dflags);
    prof_end(c);
  } else {
    size_t sh = lds_bytes(c->na, c->br, false);
    prof_begin(c, MDX_K_EAM_RHO);
    hipLaunchKernelGGL((k_eam_rho_brick<MIXED, EV, TPA>), dim3(c->br.nbricks), dim3(B), sh, c->stream, c->na,
                       c->br, c->eam, c->x, c->bin_start, c->nbr16, c->numneigh, c->rho, c->fp, c->eemb,
                       c->dflags);
    prof_end(c);
    sh = lds_bytes(c->na, c->br, true);
    prof_begin(c, MDX_K_EAM_FORCE);
    hipLaunchKernelGGL((k_eam_force_brick<MIXED, EV, TPA>), dim3(c->br.nbricks), dim3(B), sh, c->stream, c->na,
                       c->br, c->eam, c->x, c->bin_start, c->nbr16, c->numneigh, c->fp, c->eemb, c->f, c->vatom,
                       c->npad, c->partials, c->dflags);
    prof_end(c);
  }
  c->last_grid = c->br.nbricks;
}

template <bool MIXED, bool EV>
static int dispatch_brick(mdx_ctx *c) {
  switch (c->launch.tpa) {
    case 1: launch_pair_brick<MIXED, EV, 1>(c); break;
    case 2: launch_pair_brick<MIXED, EV, 2>(c); break;
    case 4: launch_pair_brick<MIXED, EV, 4>(c); break;
    case 8: launch_pair_brick<MIXED, EV, 8>(c); break;
    case 16: launch_pair_brick<MIXED, EV, 16>(c); break;
    case 32: launch_pair_brick<MIXED, EV, 32>(c); break;
    default:
      set_error("threads per atom must be one of 1, 2, 4, 8, 16, 32");
      return MDX_ERR_ARG;
  }
  return MDX_OK;
}

int brick_pair(mdx_ctx *c, int ev) {
  const bool mixed = c->precision == MDX_MIXED;
  if (mixed) return ev ? dispatch_brick<true, true>(c) : dispatch_brick<true, false>(c);
  return ev ? dispatch_brick<false, true>(c) : dispatch_brick<false, false>(c);
}

// host decode of one brick's staging (same order as stage_brick): global
// index of every LDS slot
void brick_slot_map(const NeighArgs &a, const BrickArgs &br, int brick, const int *bin_start,
                    std::vector<int> &gidx) {
  gidx.clear();
  const int E = (br.R[0] + 2 * a.S[0]) * (br.R[1] + 2 * a.S[1]) * (br.R[2] + 2 * a.S[2]);
  for (int e = 0; e < E; e++) {
    int bin, sx, sy, sz;
    if (!brick_entry(a, br, brick, e, bin, sx, sy, sz)) continue;
    for (int j = bin_start[bin]; j < bin_start[bin + 1]; j++) gidx.push_back(j);
  }
}

int brick_of_bin(const BrickArgs &br, const NeighArgs &a, int bin) {
  int ix, iy, iz;
  bin_coords(a, bin, ix, iy, iz);
  return ((iz / br.R[2]) * br.nbrick[1] + iy / br.R[1]) * br.nbrick[0] + ix / br.R[0];
}

}  // namespace mdx
