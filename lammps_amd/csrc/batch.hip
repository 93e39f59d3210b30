// batch.hip - forces and force-matching adjoints over many independent
// configurations ("replicas") in single persistent launches (mdx_batch.h).
//
// Layout: all replicas' atoms in one array, sorted by a global bin key
// (replica bin offset + local bin), so replicas stay contiguous and in
// order, and within a replica atoms are bin-sorted.  Each replica has its
// own orthogonal box, periodicity and bin grid (binsize >= cutneigh/2,
// stencil +-2 bins).  Lists are full, row-major, one row per atom; an entry
// holds the neighbour index (bits 0-26) and its periodic image code
// 9(sx+1) + 3(sy+1) + (sz+1) (bits 27-31).  The ghost coordinate is
// x_j + s*prd rounded once, as CommBrick::borders forms it
// (atom_vec_atomic.cpp:290-297).
//
// Every pair kernel is persistent: block b takes a contiguous chunk of
// atoms (chunks ordered so each XCD gets a contiguous eighth) and loops over
// it in tiles of block/TPA atoms, so one launch covers all replicas
// whatever their sizes, and the grid is a tuning parameter.
//
// Adjoints (reference expressions: pair_lj_cut.cpp:104-112,
// pair_eam.cpp:160-303).  With G_i = dL/dF_i and full-list rows:
//   lj:  dL/deps_ab   += (G_i.del) r^-8 (48 s^12 r^-6 - 24 s^6)
//        dL/dsigma_ab += (G_i.del) r^-8 (576 eps s^11 r^-6 - 144 eps s^5)
//   eam: b = -(G_i.del)/r;  z2r'(r) gets b/r, z2r(r) gets -b/r^2,
//        rhor'(r) gets b (fp_i + fp_j),
//        u_i = sum_j -((G_i - G_j).del)/r rhor'(r)  (= dL/dfp_i, both rows),
//        frho'(rho_i) gets u_i, and v_i = u_i F''(rho_i) reaches rhor(r)
//        through rho_i (second pass).
// The adjoint kernels run FP64 arithmetic whatever the forward precision.
// Gradients are accumulated per spline interval and coefficient (c3..c6:
// the value weights p^3, p^2, p, 1 and derivative weights 3p^2, 2p, 1, 0
// over the knot spacing) in LDS with ds_add_f64, flushed per block and
// summed over blocks in a fixed order; the host maps coefficient gradients
// to knot gradients with the transpose of the spline construction.
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "../../include/mdx_batch.h"
#include "mdx_internal.h"

namespace mdx {
namespace {

constexpr int kS = 2;  // stencil half-width in bins (binsize >= cutneigh / 2)
constexpr int kNT = MDX_NT1 * MDX_NT1;
constexpr int kTabLDS = 6144;  // doubles of spline table staged in LDS at most (48 KB)

struct RepGeom {
  double lo[3], prd[3], bininv[3];
  // 32-bit fixed-point positions of the mixed-precision force passes:
  // q = round((x - lo) * qscale) (mod 2^32 periodic, where the wrapping
  // difference is the minimum-image separation), qstep = prd * 2^-32
  // (periodic) or prd * 2^-30, exact
  double qscale[3], qstep[3];
  int nbin[3], periodic[3];
  int bin_off;
};

// a pair within kQBand of the cutoff (relative, on rsq) from the fixed-point
// separation is re-tested on the FP64 positions, so the cutoff decision is
// the FP64 one
constexpr double kQBand = 2e-6;

__device__ __forceinline__ int4 quantize_rep(const RepGeom &g, const double4 &p) {
  return make_int4((int)(unsigned)(long long)rint((p.x - g.lo[0]) * g.qscale[0]),
                   (int)(unsigned)(long long)rint((p.y - g.lo[1]) * g.qscale[1]),
                   (int)(unsigned)(long long)rint((p.z - g.lo[2]) * g.qscale[2]), (int)p.w);
}

__device__ __forceinline__ double qsep(int qi, int qj, double step) {
  return (double)(int)((unsigned)qi - (unsigned)qj) * step;
}



// lj/cut tables (ntypes+1)^2: cutsq, lj1..lj4 (init_one), and the
// gradient factors A1 = 48 s^12, A2 = 24 s^6, B1 = 576 eps s^11, B2 = 144 eps s^5
enum { T_CUTSQ = 0, T_LJ1, T_LJ2, T_LJ3, T_LJ4, T_A1, T_A2, T_B1, T_B2, T_N };

struct EamTab {
  int nrho, nr;
  double rdrho, rdr, rhomax, cutsq;
};

struct Tune {
  int block = 256, tpa = 8, grid = 0;
};

__device__ __forceinline__ int chunk_begin(int n, int apb, int *end) {
  const int G = gridDim.x;
  const int lb = (G % kXCD == 0) ? xcd_block() : blockIdx.x;
  const int tiles = (n + apb - 1) / apb;
  const int per = (tiles + G - 1) / G;
  const int b0 = lb * per * apb;
  *end = min(n, b0 + per * apb);
  return b0;
}

template <int TPA, typename T>
__device__ __forceinline__ T gsum(T v) {
#pragma unroll
  for (int off = TPA / 2; off > 0; off >>= 1) v += __shfl_down(v, off, TPA);
  return v;
}

__device__ __forceinline__ void shift_of(int code, double &sx, double &sy, double &sz) {
  sx = (double)(code / 9 - 1);
  sy = (double)((code / 3) % 3 - 1);
  sz = (double)(code % 3 - 1);
}

// Separation of the pair (i, list entry) and whether it lies inside cutsq:
// from the FP64 positions with the entry's image shift, or (MIXED) from the
// fixed-point mirror with the FP64 re-test within kQBand of the cutoff.
template <bool MIXED>
__device__ __forceinline__ bool b_pair(const double4 &xi, const int4 &qi, const double4 *__restrict__ x,
                                       const int4 *__restrict__ xq, const RepGeom &g, int ent, double cutsq,
                                       double &dx, double &dy, double &dz, double &rsq) {
  const int j = ent & MDX_IDX_MASK;
  if (MIXED) {
    const int4 qj = xq[j];
    dx = qsep(qi.x, qj.x, g.qstep[0]);
    dy = qsep(qi.y, qj.y, g.qstep[1]);
    dz = qsep(qi.z, qj.z, g.qstep[2]);
    rsq = dx * dx + dy * dy + dz * dz;
    if (rsq < cutsq * (1.0 - kQBand)) return true;
    if (!(rsq < cutsq * (1.0 + kQBand))) return false;
  }
  double sx, sy, sz;
  shift_of((int)((unsigned)ent >> MDX_IDX_BITS), sx, sy, sz);
  const double4 xj = x[j];
  const double ex = xi.x - (xj.x + sx * g.prd[0]), ey = xi.y - (xj.y + sy * g.prd[1]),
               ez = xi.z - (xj.z + sz * g.prd[2]);
  const double esq = ex * ex + ey * ey + ez * ez;
  if (!MIXED) {
    dx = ex;
    dy = ey;
    dz = ez;
    rsq = esq;
  }
  return esq < cutsq;
}

__device__ __forceinline__ int local_bin(const RepGeom &g, double x, int d) {
  int c = (int)floor((x - g.lo[d]) * g.bininv[d]);
  return min(max(c, 0), g.nbin[d] - 1);
}

// ---------------------------------------------------------------- build

// caller order: wrap (Domain::pbc), bin key, histogram
__global__ void k_b_load(int n, const double *__restrict__ xin, const int *__restrict__ type,
                         const int *__restrict__ rep_c, const RepGeom *__restrict__ geom, double4 *__restrict__ xw,
                         unsigned *__restrict__ key, int *__restrict__ idx, int *__restrict__ bin_count,
                         int *__restrict__ dflags) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const RepGeom &g = geom[rep_c[c]];
  double v[3];
  for (int d = 0; d < 3; d++) {
    double u = xin[3 * (size_t)c + d];
    const double hi = g.lo[d] + g.prd[d];
    if (g.periodic[d]) {
      if (u < g.lo[d] || u >= hi) {
        u -= floor((u - g.lo[d]) / g.prd[d]) * g.prd[d];
        if (u >= hi) u -= g.prd[d];
        u = fmax(u, g.lo[d]);
      }
    } else if (!(u >= g.lo[d] && u <= hi)) {
      atomicOr(&dflags[4], 1);  // outside a non-periodic box
    }
    v[d] = u;
  }
  xw[c] = make_double4(v[0], v[1], v[2], (double)type[c]);
  const int bx = local_bin(g, v[0], 0), by = local_bin(g, v[1], 1), bz = local_bin(g, v[2], 2);
  const unsigned k = (unsigned)(g.bin_off + (bz * g.nbin[1] + by) * g.nbin[0] + bx);
  key[c] = k;
  idx[c] = c;
  atomicAdd(&bin_count[k], 1);
}

__global__ void k_b_permute(int n, const int *__restrict__ idx_s, const double4 *__restrict__ xw,
                            const int *__restrict__ rep_c, const RepGeom *__restrict__ geom, double4 *__restrict__ x,
                            int4 *__restrict__ xq, int *__restrict__ rep, int *__restrict__ orig,
                            int *__restrict__ inv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int c = idx_s[k];
  const double4 p = xw[c];
  x[k] = p;
  xq[k] = quantize_rep(geom[rep_c[c]], p);
  rep[k] = rep_c[c];
  orig[k] = c;
  inv[c] = k;
}

template <int TPA>
__global__ __launch_bounds__(1024) void k_b_build(int n, const double4 *__restrict__ x, const int *__restrict__ rep,
                                                  const RepGeom *__restrict__ geom,
                                                  const int *__restrict__ bin_start, double cutneighsq, int cap,
                                                  int *__restrict__ nbr, int *__restrict__ numneigh,
                                                  int *__restrict__ dflags, unsigned long long *__restrict__ total) {
  const int apb = blockDim.x / TPA;
  int end;
  const int b0 = chunk_begin(n, apb, &end);
  const int lane = threadIdx.x % TPA;
  const int gshift = (threadIdx.x % 64) / TPA * TPA;
  const unsigned long long below = (lane == 0) ? 0ull : ((1ull << lane) - 1ull) << gshift;
  const unsigned long long gmask = ((TPA == 64) ? ~0ull : ((1ull << TPA) - 1ull));
  for (int base = b0; base < end; base += apb) {
    const int i = base + threadIdx.x / TPA;
    if (i >= end) continue;
    const double4 xi = x[i];
    const RepGeom &g = geom[rep[i]];
    const int c0 = local_bin(g, xi.x, 0), c1 = local_bin(g, xi.y, 1), c2 = local_bin(g, xi.z, 2);
    int *row = nbr + (size_t)i * cap;
    int cnt = 0;
    for (int dz = -kS; dz <= kS; dz++) {
      int uz = c2 + dz, sz = 0;
      if (g.periodic[2]) {
        if (uz < 0) { uz += g.nbin[2]; sz = -1; }
        else if (uz >= g.nbin[2]) { uz -= g.nbin[2]; sz = 1; }
      } else if (uz < 0 || uz >= g.nbin[2]) continue;
      for (int dy = -kS; dy <= kS; dy++) {
        int uy = c1 + dy, sy = 0;
        if (g.periodic[1]) {
          if (uy < 0) { uy += g.nbin[1]; sy = -1; }
          else if (uy >= g.nbin[1]) { uy -= g.nbin[1]; sy = 1; }
        } else if (uy < 0 || uy >= g.nbin[1]) continue;
        for (int dx = -kS; dx <= kS; dx++) {
          int ux = c0 + dx, sx = 0;
          if (g.periodic[0]) {
            if (ux < 0) { ux += g.nbin[0]; sx = -1; }
            else if (ux >= g.nbin[0]) { ux -= g.nbin[0]; sx = 1; }
          } else if (ux < 0 || ux >= g.nbin[0]) continue;
          const int bidx = g.bin_off + (uz * g.nbin[1] + uy) * g.nbin[0] + ux;
          const int js = bin_start[bidx], je = bin_start[bidx + 1];
          const int code = 9 * (sx + 1) + 3 * (sy + 1) + (sz + 1);
          const double gx = sx * g.prd[0], gy = sy * g.prd[1], gz = sz * g.prd[2];
          for (int k0 = js; k0 < je; k0 += TPA) {
            const int j = k0 + lane;
            bool keep = false;
            if (j < je) {
              const double4 xj = x[j];
              const double dxv = xi.x - (xj.x + gx), dyv = xi.y - (xj.y + gy), dzv = xi.z - (xj.z + gz);
              const double rsq = dxv * dxv + dyv * dyv + dzv * dzv;
              keep = rsq < cutneighsq && !(j == i && code == 13);
            }
            const unsigned long long m = (__ballot(keep) >> gshift) & gmask;
            if (keep) {
              const int pos = cnt + __popcll((m << gshift) & below);
              if (pos < cap) row[pos] = j | (code << MDX_IDX_BITS);
            }
            cnt += __popcll(m);
          }
        }
      }
    }
    if (lane == 0) {
      numneigh[i] = cnt;
      atomicMax(&dflags[0], cnt);
      atomicAdd(total, (unsigned long long)cnt);
    }
  }
}

// ---------------------------------------------------------------- lj/cut

// MIXED: neighbours are read from the fixed-point mirror (16 B instead of
// 32, no image decode: the wrapping difference is the minimum-image
// separation), separations rebuilt in FP64, a pair within kQBand of the
// cutoff re-tested on the FP64 positions with its image shift; FP32 pair
// arithmetic, FP64 accumulation.
template <bool MIXED, int TPA>
__global__ __launch_bounds__(1024) void k_b_lj(int n, int nt1, const double4 *__restrict__ x,
                                               const int4 *__restrict__ xq, const int *__restrict__ rep,
                                               const RepGeom *__restrict__ geom, const int *__restrict__ nbr,
                                               const int *__restrict__ numneigh, int cap,
                                               const double *__restrict__ tab, double4 *__restrict__ f) {
  __shared__ double s_tab[T_N * kNT];
  __shared__ double s_qlo[MIXED ? kNT : 1], s_qhi[MIXED ? kNT : 1];
  __shared__ float s_f[MIXED ? 4 * kNT : 1];
  for (int k = threadIdx.x; k < T_N * kNT; k += blockDim.x) s_tab[k] = tab[k];
  if (MIXED)
    for (int k = threadIdx.x; k < kNT; k += blockDim.x) {
      s_qlo[k] = tab[T_CUTSQ * kNT + k] * (1.0 - kQBand);
      s_qhi[k] = tab[T_CUTSQ * kNT + k] * (1.0 + kQBand);
      s_f[k] = (float)tab[T_LJ1 * kNT + k];
      s_f[kNT + k] = (float)tab[T_LJ2 * kNT + k];
      s_f[2 * kNT + k] = (float)tab[T_LJ3 * kNT + k];
      s_f[3 * kNT + k] = (float)tab[T_LJ4 * kNT + k];
    }
  __syncthreads();
  const int apb = blockDim.x / TPA;
  int end;
  const int b0 = chunk_begin(n, apb, &end);
  const int lane = threadIdx.x % TPA;
  for (int base = b0; base < end; base += apb) {
    const int i = base + threadIdx.x / TPA;
    if (i >= end) continue;
    const double4 xi = x[i];
    const int ti = (int)xi.w;
    const RepGeom &g = geom[rep[i]];
    const double px = g.prd[0], py = g.prd[1], pz = g.prd[2];
    const int *row = nbr + (size_t)i * cap;
    const int jn = numneigh[i];
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
    if constexpr (MIXED) {
      // chunks of four consecutive entries per lane (one 16-B index load,
      // rows start 16-B aligned), then four 16-B gathers in flight
      const int4 qi = xq[i];
      const double q0 = g.qstep[0], q1 = g.qstep[1], q2 = g.qstep[2];
      for (int k0 = 4 * lane; k0 < jn; k0 += 4 * TPA) {
        const int4 e4 = *reinterpret_cast<const int4 *>(row + k0);
        const int ev[4] = {e4.x, e4.y, e4.z, e4.w};
        int ent[4];
        int4 qj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          ent[u] = k0 + u < jn ? ev[u] : i;  // the row's tail past jn is not written
          qj[u] = xq[ent[u] & MDX_IDX_MASK];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (k0 + u >= jn) continue;
          const double dx = qsep(qi.x, qj[u].x, q0), dy = qsep(qi.y, qj[u].y, q1), dz = qsep(qi.z, qj[u].z, q2);
          const double rsq = dx * dx + dy * dy + dz * dz;
          const int t = ti * nt1 + qj[u].w;
          bool inside = rsq < s_qlo[t];
          if (!inside && rsq < s_qhi[t]) {
            double sx, sy, sz;
            shift_of((unsigned)ent[u] >> MDX_IDX_BITS, sx, sy, sz);
            const double4 xj = x[ent[u] & MDX_IDX_MASK];
            const double ex = xi.x - (xj.x + sx * px), ey = xi.y - (xj.y + sy * py), ez = xi.z - (xj.z + sz * pz);
            inside = ex * ex + ey * ey + ez * ez < s_tab[T_CUTSQ * kNT + t];
          }
          if (inside) {
            const float r2inv = __builtin_amdgcn_rcpf((float)rsq);
            const float r6inv = r2inv * r2inv * r2inv;
            const double fpair = (double)(r6inv * (s_f[t] * r6inv - s_f[kNT + t]) * r2inv);
            fx += dx * fpair;
            fy += dy * fpair;
            fz += dz * fpair;
            e += 0.5 * (double)(r6inv * (s_f[2 * kNT + t] * r6inv - s_f[3 * kNT + t]));
          }
        }
      }
    } else {
      for (int k0 = 4 * lane; k0 < jn; k0 += 4 * TPA) {
        const int4 e4 = *reinterpret_cast<const int4 *>(row + k0);
        const int ev[4] = {e4.x, e4.y, e4.z, e4.w};
        int ent[4];
        double4 xj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          ent[u] = k0 + u < jn ? ev[u] : i;
          xj[u] = x[ent[u] & MDX_IDX_MASK];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (k0 + u >= jn) continue;
          double sx, sy, sz;
          shift_of((unsigned)ent[u] >> MDX_IDX_BITS, sx, sy, sz);
          const double dx = xi.x - (xj[u].x + sx * px), dy = xi.y - (xj[u].y + sy * py),
                       dz = xi.z - (xj[u].z + sz * pz);
          const double rsq = dx * dx + dy * dy + dz * dz;
          const int t = ti * nt1 + (int)xj[u].w;
          if (rsq < s_tab[T_CUTSQ * kNT + t]) {
            const double r2inv = 1.0 / rsq;
            const double r6inv = r2inv * r2inv * r2inv;
            const double forcelj = r6inv * (s_tab[T_LJ1 * kNT + t] * r6inv - s_tab[T_LJ2 * kNT + t]);
            const double fpair = forcelj * r2inv;
            fx += dx * fpair;
            fy += dy * fpair;
            fz += dz * fpair;
            e += 0.5 * (r6inv * (s_tab[T_LJ3 * kNT + t] * r6inv - s_tab[T_LJ4 * kNT + t]));
          }
        }
      }
    }
    fx = gsum<TPA>(fx);
    fy = gsum<TPA>(fy);
    fz = gsum<TPA>(fz);
    e = gsum<TPA>(e);
    if (lane == 0) f[i] = make_double4(fx, fy, fz, e);
  }
}

// per-block partial rows: [eps grad (nt1^2) | sigma grad (nt1^2)]
template <bool MIXED, int TPA>
__global__ __launch_bounds__(1024) void k_b_lj_vjp(int n, int nt1, const double4 *__restrict__ x,
                                                   const int *__restrict__ rep, const RepGeom *__restrict__ geom,
                                                   const int *__restrict__ nbr, const int *__restrict__ numneigh,
                                                   int cap, const double *__restrict__ tab,
                                                   const double4 *__restrict__ gin, double *__restrict__ partials) {
  __shared__ double s_tab[T_N * kNT];
  __shared__ double s_acc[2 * kNT];
  for (int k = threadIdx.x; k < T_N * kNT; k += blockDim.x) s_tab[k] = tab[k];
  for (int k = threadIdx.x; k < 2 * kNT; k += blockDim.x) s_acc[k] = 0.0;
  __syncthreads();
  const int apb = blockDim.x / TPA;
  int end;
  const int b0 = chunk_begin(n, apb, &end);
  const int lane = threadIdx.x % TPA;
  for (int base = b0; base < end; base += apb) {
    const int i = base + threadIdx.x / TPA;
    if (i >= end) continue;
    const double4 xi = x[i];
    const double4 gi = gin[i];
    const int ti = (int)xi.w;
    const RepGeom &g = geom[rep[i]];
    const double px = g.prd[0], py = g.prd[1], pz = g.prd[2];
    const int *row = nbr + (size_t)i * cap;
    const int jn = numneigh[i];
    double ae[MDX_NT1], as[MDX_NT1];
#pragma unroll
    for (int t = 0; t < MDX_NT1; t++) ae[t] = as[t] = 0.0;
    for (int k = lane; k < jn; k += TPA) {
      const int ent = row[k];
      const int j = ent & MDX_IDX_MASK;
      double sx, sy, sz;
      shift_of((unsigned)ent >> MDX_IDX_BITS, sx, sy, sz);
      const double4 xj = x[j];
      const double dx = xi.x - (xj.x + sx * px), dy = xi.y - (xj.y + sy * py), dz = xi.z - (xj.z + sz * pz);
      const double rsq = dx * dx + dy * dy + dz * dz;
      const int tj = (int)xj.w;
      const int t = ti * nt1 + tj;
      if (rsq < s_tab[T_CUTSQ * kNT + t]) {
        const double a = gi.x * dx + gi.y * dy + gi.z * dz;
        double de, ds;
        if (MIXED) {
          const float r2inv = 1.0f / (float)rsq;
          const float r6inv = r2inv * r2inv * r2inv;
          const float r8 = r2inv * r6inv;
          de = (double)(r8 * ((float)s_tab[T_A1 * kNT + t] * r6inv - (float)s_tab[T_A2 * kNT + t]));
          ds = (double)(r8 * ((float)s_tab[T_B1 * kNT + t] * r6inv - (float)s_tab[T_B2 * kNT + t]));
        } else {
          const double r2inv = 1.0 / rsq;
          const double r6inv = r2inv * r2inv * r2inv;
          const double r8 = r2inv * r6inv;
          de = r8 * (s_tab[T_A1 * kNT + t] * r6inv - s_tab[T_A2 * kNT + t]);
          ds = r8 * (s_tab[T_B1 * kNT + t] * r6inv - s_tab[T_B2 * kNT + t]);
        }
#pragma unroll
        for (int tt = 1; tt < MDX_NT1; tt++)
          if (tt == tj) {
            ae[tt] += a * de;
            as[tt] += a * ds;
          }
      }
    }
#pragma unroll
    for (int tt = 1; tt < MDX_NT1; tt++) {
      if (tt < nt1) {
        const double se = gsum<TPA>(ae[tt]), ss = gsum<TPA>(as[tt]);
        if (lane == 0) {
          atomicAdd(&s_acc[ti * nt1 + tt], se);
          atomicAdd(&s_acc[kNT + ti * nt1 + tt], ss);
        }
      }
    }
  }
  __syncthreads();
  double *out = partials + (size_t)blockIdx.x * 2 * kNT;
  for (int k = threadIdx.x; k < 2 * kNT; k += blockDim.x) out[k] = s_acc[k];
}

// ---------------------------------------------------------------- eam

// spline tables (n+1)*7 in LDS when they fit, else read from global memory
template <typename T>
__device__ __forceinline__ const T *stage(T *lds, const T *g, int count, bool use_lds) {
  if (!use_lds) return g;
  for (int k = threadIdx.x; k < count; k += blockDim.x) lds[k] = g[k];
  return lds;
}

template <typename T>
__device__ __forceinline__ void r_spline(const EamTab &et, T r, int &m, T &p) {
  T pp = r * (T)et.rdr + (T)1;
  m = (int)pp;
  m = min(m, et.nr - 1);
  pp -= (T)m;
  p = pp < (T)1 ? pp : (T)1;
}

template <bool MIXED, int TPA>
__global__ __launch_bounds__(1024) void k_b_eam_rho(int n, const double4 *__restrict__ x,
                                                    const int4 *__restrict__ xq, const int *__restrict__ rep,
                                                    const RepGeom *__restrict__ geom, const int *__restrict__ nbr,
                                                    const int *__restrict__ numneigh, int cap, EamTab et,
                                                    const double *__restrict__ rhor_d, const float *__restrict__ rhor_f,
                                                    const double *__restrict__ frho_d, int use_lds,
                                                    double *__restrict__ rho, double *__restrict__ fp,
                                                    double *__restrict__ eemb) {
  extern __shared__ double s_dyn[];
  using T = typename std::conditional<MIXED, float, double>::type;
  const T *rhor;
  if (MIXED) rhor = (const T *)stage((float *)s_dyn, rhor_f, (et.nr + 1) * 7, use_lds);
  else rhor = (const T *)stage((double *)s_dyn, rhor_d, (et.nr + 1) * 7, use_lds);
  __syncthreads();
  const int apb = blockDim.x / TPA;
  int end;
  const int b0 = chunk_begin(n, apb, &end);
  const int lane = threadIdx.x % TPA;
  for (int base = b0; base < end; base += apb) {
    const int i = base + threadIdx.x / TPA;
    if (i >= end) continue;
    const double4 xi = x[i];
    const int4 qi = MIXED ? xq[i] : make_int4(0, 0, 0, 0);
    const RepGeom &g = geom[rep[i]];
    const int *row = nbr + (size_t)i * cap;
    const int jn = numneigh[i];
    double acc = 0.0;
    for (int k = lane; k < jn; k += TPA) {
      double dx, dy, dz, rsq;
      if (b_pair<MIXED>(xi, qi, x, xq, g, row[k], et.cutsq, dx, dy, dz, rsq)) {
        int m;
        T p;
        if constexpr (MIXED) (void)eam_index_mixed(rsq, et.rdr, et.nr, m, p);
        else r_spline<T>(et, (T)sqrt(rsq), m, p);
        const T *c = rhor + m * 7;
        acc += (double)(((c[3] * p + c[4]) * p + c[5]) * p + c[6]);
      }
    }
    acc = gsum<TPA>(acc);
    if (lane == 0) {
      // fp pass (pair_eam.cpp:213-228), FP64
      double p = acc * et.rdrho + 1.0;
      int m = (int)p;
      m = max(1, min(m, et.nrho - 1));
      p -= m;
      p = fmin(p, 1.0);
      const double *c = frho_d + m * 7;
      const double fpi = (c[0] * p + c[1]) * p + c[2];
      double phi = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
      if (acc > et.rhomax) phi += fpi * (acc - et.rhomax);
      rho[i] = acc;
      fp[i] = fpi;
      eemb[i] = phi;
    }
  }
}

template <bool MIXED, int TPA>
__global__ __launch_bounds__(1024) void k_b_eam_force(int n, const double4 *__restrict__ x,
                                                      const int4 *__restrict__ xq, const int *__restrict__ rep, const RepGeom *__restrict__ geom,
                                                      const int *__restrict__ nbr, const int *__restrict__ numneigh,
                                                      int cap, EamTab et, const double *__restrict__ rhor_d,
                                                      const double *__restrict__ z2r_d, const float *__restrict__ rhor_f,
                                                      const float *__restrict__ z2r_f, int use_lds,
                                                      const double *__restrict__ fp, const double *__restrict__ eemb,
                                                      double4 *__restrict__ f) {
  extern __shared__ double s_dyn[];
  using T = typename std::conditional<MIXED, float, double>::type;
  const int tn = (et.nr + 1) * 7;
  const T *rhor, *z2r;
  if (MIXED) {
    rhor = (const T *)stage((float *)s_dyn, rhor_f, tn, use_lds);
    z2r = (const T *)stage((float *)s_dyn + tn, z2r_f, tn, use_lds);
  } else {
    rhor = (const T *)stage((double *)s_dyn, rhor_d, tn, use_lds);
    z2r = (const T *)stage((double *)s_dyn + tn, z2r_d, tn, use_lds);
  }
  __syncthreads();
  const int apb = blockDim.x / TPA;
  int end;
  const int b0 = chunk_begin(n, apb, &end);
  const int lane = threadIdx.x % TPA;
  for (int base = b0; base < end; base += apb) {
    const int i = base + threadIdx.x / TPA;
    if (i >= end) continue;
    const double4 xi = x[i];
    const int4 qi = MIXED ? xq[i] : make_int4(0, 0, 0, 0);
    const double fpi = fp[i];
    const RepGeom &g = geom[rep[i]];
    const int *row = nbr + (size_t)i * cap;
    const int jn = numneigh[i];
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
    for (int k = lane; k < jn; k += TPA) {
      const int ent = row[k];
      double dx, dy, dz, rsq;
      if (b_pair<MIXED>(xi, qi, x, xq, g, ent, et.cutsq, dx, dy, dz, rsq)) {
        const double fpj = fp[ent & MDX_IDX_MASK];
        int m;
        T p, r;
        if constexpr (MIXED) r = eam_index_mixed(rsq, et.rdr, et.nr, m, p);
        else {
          r = (T)sqrt(rsq);
          r_spline<T>(et, r, m, p);
        }
        const T *c = rhor + m * 7;
        const T rhop = (c[0] * p + c[1]) * p + c[2];
        c = z2r + m * 7;
        const T z2p = (c[0] * p + c[1]) * p + c[2];
        const T z2 = ((c[3] * p + c[4]) * p + c[5]) * p + c[6];
        const T recip = MIXED ? (T)__builtin_amdgcn_rcpf((float)r) : (T)1 / r;
        const T phi = z2 * recip;
        const T phip = z2p * recip - phi * recip;
        const T psip = (T)fpi * rhop + (T)fpj * rhop + phip;
        const double fpair = (double)(-psip * recip);
        fx += dx * fpair;
        fy += dy * fpair;
        fz += dz * fpair;
        e += 0.5 * (double)phi;
      }
    }
    fx = gsum<TPA>(fx);
    fy = gsum<TPA>(fy);
    fz = gsum<TPA>(fz);
    e = gsum<TPA>(e);
    if (lane == 0) f[i] = make_double4(fx, fy, fz, e + eemb[i]);
  }
}

// partial row layout: [frho c3..c6 (nrho+1)*4 | rhor (nr+1)*4 | z2r (nr+1)*4]
template <bool MIXED, int TPA>
__global__ __launch_bounds__(1024) void k_b_eam_adj_a(int n, const double4 *__restrict__ x,
                                                      const int *__restrict__ rep, const RepGeom *__restrict__ geom,
                                                      const int *__restrict__ nbr, const int *__restrict__ numneigh,
                                                      int cap, EamTab et, const double *__restrict__ rhor_d,
                                                      const float *__restrict__ rhor_f,
                                                      const double *__restrict__ frho_d, const double *__restrict__ rho,
                                                      const double *__restrict__ fp, const double4 *__restrict__ gin,
                                                      double *__restrict__ vout, double *__restrict__ partials) {
  extern __shared__ double s_dyn[];
  using T = typename std::conditional<MIXED, float, double>::type;
  const int P = 4 * (et.nrho + 1) + 8 * (et.nr + 1);
  double *s_gf = s_dyn, *s_gr = s_dyn + 4 * (et.nrho + 1), *s_gz = s_gr + 4 * (et.nr + 1);
  for (int k = threadIdx.x; k < P; k += blockDim.x) s_dyn[k] = 0.0;
  const int tn = (et.nr + 1) * 7;
  const T *rhor;
  if (MIXED) rhor = (const T *)stage((float *)(s_dyn + P), rhor_f, tn, false);
  else rhor = (const T *)stage((double *)(s_dyn + P), rhor_d, tn, false);
  __syncthreads();
  const int apb = blockDim.x / TPA;
  int end;
  const int b0 = chunk_begin(n, apb, &end);
  const int lane = threadIdx.x % TPA;
  const double rdr = et.rdr;
  for (int base = b0; base < end; base += apb) {
    const int i = base + threadIdx.x / TPA;
    if (i >= end) continue;
    const double4 xi = x[i];
    const double4 gi = gin[i];
    const double fpi = fp[i];
    const RepGeom &g = geom[rep[i]];
    const double px = g.prd[0], py = g.prd[1], pz = g.prd[2];
    const int *row = nbr + (size_t)i * cap;
    const int jn = numneigh[i];
    double u = 0.0;
    for (int k = lane; k < jn; k += TPA) {
      const int ent = row[k];
      const int j = ent & MDX_IDX_MASK;
      double sx, sy, sz;
      shift_of((unsigned)ent >> MDX_IDX_BITS, sx, sy, sz);
      const double4 xj = x[j];
      const double dx = xi.x - (xj.x + sx * px), dy = xi.y - (xj.y + sy * py), dz = xi.z - (xj.z + sz * pz);
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq < et.cutsq) {
        const double4 gj = gin[j];
        const double fpj = fp[j];
        int m;
        T p;
        const T r = MIXED ? (T)sqrtf((float)rsq) : (T)sqrt(rsq);
        r_spline<T>(et, r, m, p);
        const T *c = rhor + m * 7;
        const double rhop = (double)((c[0] * p + c[1]) * p + c[2]);
        const double rr = (double)r;
        const double b = -(gi.x * dx + gi.y * dy + gi.z * dz) / rr;
        const double bs = -((gi.x - gj.x) * dx + (gi.y - gj.y) * dy + (gi.z - gj.z) * dz) / rr;
        const double pd = (double)p;
        const double dw0 = 3.0 * pd * pd * rdr, dw1 = 2.0 * pd * rdr, dw2 = rdr;
        const double vw0 = pd * pd * pd, vw1 = pd * pd, vw2 = pd;
        const double gz2p = b / rr, gz2 = -b / (rr * rr), grp = b * (fpi + fpj);
        double *gz = s_gz + 4 * m, *gr = s_gr + 4 * m;
        atomicAdd(gz + 0, gz2p * dw0 + gz2 * vw0);
        atomicAdd(gz + 1, gz2p * dw1 + gz2 * vw1);
        atomicAdd(gz + 2, gz2p * dw2 + gz2 * vw2);
        atomicAdd(gz + 3, gz2);
        atomicAdd(gr + 0, grp * dw0);
        atomicAdd(gr + 1, grp * dw1);
        atomicAdd(gr + 2, grp * dw2);
        u += bs * rhop;
      }
    }
    u = gsum<TPA>(u);
    if (lane == 0) {
      // fp = F'(rho_i): dL/dF'(rho_i) = u, dL/drho_i = u F''(rho_i)
      const double rhoi = rho[i];
      double p = rhoi * et.rdrho + 1.0;
      int m = (int)p;
      m = max(1, min(m, et.nrho - 1));
      p -= m;
      const bool clamped = p > 1.0;
      p = fmin(p, 1.0);
      const double rdrho = et.rdrho;
      double *gf = s_gf + 4 * m;
      atomicAdd(gf + 0, u * 3.0 * p * p * rdrho);
      atomicAdd(gf + 1, u * 2.0 * p * rdrho);
      atomicAdd(gf + 2, u * rdrho);
      const double *c = frho_d + m * 7;
      const double fpp = clamped ? 0.0 : (2.0 * c[0] * p + c[1]) * rdrho;
      vout[i] = u * fpp;
    }
  }
  __syncthreads();
  double *out = partials + (size_t)blockIdx.x * P;
  for (int k = threadIdx.x; k < P; k += blockDim.x) out[k] = s_dyn[k];
}

template <bool MIXED, int TPA>
__global__ __launch_bounds__(1024) void k_b_eam_adj_b(int n, const double4 *__restrict__ x,
                                                      const int *__restrict__ rep, const RepGeom *__restrict__ geom,
                                                      const int *__restrict__ nbr, const int *__restrict__ numneigh,
                                                      int cap, EamTab et, const double *__restrict__ vin,
                                                      double *__restrict__ partials) {
  extern __shared__ double s_dyn[];
  using T = typename std::conditional<MIXED, float, double>::type;
  const int P = 4 * (et.nrho + 1) + 8 * (et.nr + 1);
  double *s_gr = s_dyn + 4 * (et.nrho + 1);
  for (int k = threadIdx.x; k < P; k += blockDim.x) s_dyn[k] = 0.0;
  __syncthreads();
  const int apb = blockDim.x / TPA;
  int end;
  const int b0 = chunk_begin(n, apb, &end);
  const int lane = threadIdx.x % TPA;
  for (int base = b0; base < end; base += apb) {
    const int i = base + threadIdx.x / TPA;
    if (i >= end) continue;
    const double vi = vin[i];
    if (vi == 0.0) continue;
    const double4 xi = x[i];
    const RepGeom &g = geom[rep[i]];
    const double px = g.prd[0], py = g.prd[1], pz = g.prd[2];
    const int *row = nbr + (size_t)i * cap;
    const int jn = numneigh[i];
    for (int k = lane; k < jn; k += TPA) {
      const int ent = row[k];
      const int j = ent & MDX_IDX_MASK;
      double sx, sy, sz;
      shift_of((unsigned)ent >> MDX_IDX_BITS, sx, sy, sz);
      const double4 xj = x[j];
      const double dx = xi.x - (xj.x + sx * px), dy = xi.y - (xj.y + sy * py), dz = xi.z - (xj.z + sz * pz);
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq < et.cutsq) {
        int m;
        T p;
        r_spline<T>(et, MIXED ? (T)sqrtf((float)rsq) : (T)sqrt(rsq), m, p);
        const double pd = (double)p;
        double *gr = s_gr + 4 * m;
        atomicAdd(gr + 0, vi * pd * pd * pd);
        atomicAdd(gr + 1, vi * pd * pd);
        atomicAdd(gr + 2, vi * pd);
        atomicAdd(gr + 3, vi);
      }
    }
  }
  __syncthreads();
  double *out = partials + (size_t)blockIdx.x * P;
  for (int k = threadIdx.x; k < P; k += blockDim.x) out[k] = s_dyn[k];
}

// ---------------------------------------------------------------- utilities

// Column sums of the per-block partial rows, in two fixed-order passes
// (deterministic): k_b_colsum_chunks sums chunk ry of the rows (rows
// [rows*ry/nchunk, rows*(ry+1)/nchunk)) for 64 columns per block, four row
// lanes per column combined in LDS; k_b_colsum then sums the nchunk chunk
// rows.  One thread per column over every row leaves the device mostly idle
// (a few hundred columns, thousands of rows).
constexpr int kColChunks = 64;
__global__ __launch_bounds__(256) void k_b_colsum_chunks(const double *__restrict__ part, int rows, int cols,
                                                         double *__restrict__ mid) {
  __shared__ double s[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), sub = threadIdx.x >> 6;
  const int ry = blockIdx.y, nchunk = gridDim.y;
  const int r0 = (int)((long)rows * ry / nchunk), r1 = (int)((long)rows * (ry + 1) / nchunk);
  double acc = 0.0;
  if (c < cols)
    for (int r = r0 + sub; r < r1; r += 4) acc += part[(size_t)r * cols + c];
  s[sub][threadIdx.x & 63] = acc;
  __syncthreads();
  if (sub == 0 && c < cols) {
    const int l = threadIdx.x & 63;
    mid[(size_t)ry * cols + c] = ((s[0][l] + s[1][l]) + s[2][l]) + s[3][l];
  }
}

// column sums over rows in a fixed order (deterministic)
__global__ void k_b_colsum(const double *__restrict__ part, int rows, int cols, double *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  double s = 0.0;
  for (int r = 0; r < rows; r++) s += part[(size_t)r * cols + c];
  out[c] = s;
}

// per-replica energy: one block per replica, fixed-order tree
__global__ void k_b_rep_energy(const double4 *__restrict__ f, const int *__restrict__ off, double *__restrict__ out) {
  __shared__ double s[256];
  const int r = blockIdx.x;
  double acc = 0.0;
  for (int i = off[r] + threadIdx.x; i < off[r + 1]; i += 256) acc += f[i].w;
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[r] = s[0];
}

__global__ void k_b_gather_f(int n, const int *__restrict__ inv, const double4 *__restrict__ f,
                             double *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double4 v = f[inv[c]];
  out[3 * (size_t)c] = v.x;
  out[3 * (size_t)c + 1] = v.y;
  out[3 * (size_t)c + 2] = v.z;
}

__global__ void k_b_gather_s(int n, const int *__restrict__ inv, const double *__restrict__ a, double *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) out[c] = a[inv[c]];
}

__global__ void k_b_scatter_g(int n, const int *__restrict__ inv, const double *__restrict__ gin,
                              double4 *__restrict__ g) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  g[inv[c]] = make_double4(gin[3 * (size_t)c], gin[3 * (size_t)c + 1], gin[3 * (size_t)c + 2], 0.0);
}

}  // namespace
}  // namespace mdx

using namespace mdx;

struct mdx_batch {
  int device = 0;
  hipStream_t stream = nullptr;      // where the batch's work goes: own_stream or the caller's
  hipStream_t own_stream = nullptr;  // created with the batch
  int nrep = 0, ntypes = 0, n = 0, precision = MDX_FP64;
  double skin = 0.0;
  std::vector<int> natoms, off;
  std::vector<RepGeom> geom_h;
  RepGeom *geom = nullptr;
  int *off_d = nullptr;
  int total_bins = 0, bins_cap = 0;
  double cutneigh = 0.0;  // of the current bins/list
  bool positions = false, built = false;
  // atoms
  double *xin = nullptr;  // caller order staging [N][3]
  int *type_c = nullptr, *rep_c = nullptr;
  double4 *xw = nullptr, *x = nullptr, *f = nullptr, *g = nullptr;
  int4 *xq = nullptr;  // fixed-point copy of x (mixed-precision force passes)
  int *rep = nullptr, *orig = nullptr, *inv = nullptr;
  unsigned *key = nullptr, *key_s = nullptr;
  int *idx = nullptr, *idx_s = nullptr;
  int *bin_count = nullptr, *bin_start = nullptr;
  void *sort_tmp = nullptr, *scan_tmp = nullptr;
  size_t sort_bytes = 0, scan_bytes = 0;
  int *nbr = nullptr, *numneigh = nullptr;
  int cap = 0;
  int *dflags = nullptr;  // [0] max count, [4] outside box, [8..9] u64 total
  long total_neighbors = 0;
  int max_neighbors = 0;
  double *staging = nullptr;  // [N][3] output staging
  double *rho = nullptr, *fp = nullptr, *eemb = nullptr, *vadj = nullptr;
  double *partials = nullptr;
  size_t partials_cap = 0;  // doubles
  double *colsum = nullptr;
  size_t colsum_cap = 0;
  // pair
  int pair = MDX_PAIR_NONE;
  double *ljtab = nullptr;  // T_N * kNT
  double lj_cutmax = 0.0;
  std::vector<double> lj_tab;  // the table last uploaded (a repeated init uploads nothing)
  EamTab et{};
  double *frho_d = nullptr, *rhor_d = nullptr, *z2r_d = nullptr;
  float *rhor_f = nullptr, *z2r_f = nullptr;
  std::vector<double> eam_knots[3];
  double eam_drho = 0.0, eam_dr = 0.0, eam_cut = 0.0;
  int eam_rho_prec = -1;  // precision of the rho/fp on the device for the current positions and tables (-1: stale)
  // tuning / timing
  Tune tune[MDX_BK_NKIND];
  bool autotune = false;
  bool tuned[MDX_BK_NKIND] = {false};
  double kms[MDX_BK_NKIND] = {0};
  long klaunch[MDX_BK_NKIND] = {0};
  int num_cu = 256;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int lds_max = 65536;
};

namespace {

int bfail(hipError_t e, const char *what) { return hip_fail(e, what); }

template <typename P>
int dalloc(P **p, size_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, count * sizeof(P));
  return e == hipSuccess ? MDX_OK : bfail(e, "batch allocation");
}

bool is_device_ptr(const void *p, int device) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice && a.device == device;
}

int check_b(mdx_batch *b) {
  if (!b) {
    set_error("null batch");
    return MDX_ERR_ARG;
  }
  hipError_t e = hipSetDevice(b->device);
  return e == hipSuccess ? MDX_OK : bfail(e, "hipSetDevice");
}

// run one launch of kind `kind` under timing; `launch(tune)` enqueues it
template <typename F>
int timed(mdx_batch *b, int kind, const Tune &t, F &&launch, float *ms_out = nullptr, bool count = true) {
  (void)hipEventRecord(b->ev0, b->stream);
  launch(t);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail(e, "batch kernel launch");
  (void)hipEventRecord(b->ev1, b->stream);
  e = hipEventSynchronize(b->ev1);
  if (e != hipSuccess) return bfail(e, "batch kernel");
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
  if (ms_out) *ms_out = ms;
  if (count) {
    b->kms[kind] += ms;
    b->klaunch[kind]++;
  }
  return MDX_OK;
}

// kernels that flush per-block partial rows cap the grid at kPartialBlocksPerCU per CU
constexpr int kPartialBlocksPerCU = 8;
int persistent_grid(mdx_batch *b, const Tune &t, bool partial_rows = false) {
  const int apb = t.block / t.tpa;
  const long tiles = (b->n + apb - 1) / apb;
  long G = t.grid > 0 ? t.grid : tiles;
  G = std::max<long>(1, std::min<long>(G, tiles));
  if (partial_rows) G = std::min<long>(G, (long)kPartialBlocksPerCU * b->num_cu);
  if (G >= kXCD) G = (G / kXCD) * kXCD;
  return (int)G;
}

// candidates for the autotuner
std::vector<Tune> candidates(mdx_batch *b) {
  std::vector<Tune> v;
  for (int tpa : {2, 4, 8, 16})
    for (int block : {128, 256, 512})
      for (int per_cu : {2, 4, 8, 0}) {
        Tune t;
        t.tpa = tpa;
        t.block = block;
        t.grid = per_cu ? per_cu * b->num_cu : 0;
        v.push_back(t);
      }
  return v;
}

// runs kind with the current tune, or tunes first when asked; the last run
// uses the chosen shape so its results are the call's results
template <typename F>
int run_kind(mdx_batch *b, int kind, F &&launch) {
  if (b->autotune && !b->tuned[kind]) {
    float best = 1e30f;
    Tune bt = b->tune[kind];
    for (const Tune &t : candidates(b)) {
      float ms = 0.0f;
      for (int rep = 0; rep < 2; rep++) {  // second timing counts (warm caches)
        int rc = timed(b, kind, t, launch, &ms, false);
        if (rc) return rc;
      }
      if (ms < best) {
        best = ms;
        bt = t;
      }
    }
    b->tune[kind] = bt;
    b->tuned[kind] = true;
  }
  return timed(b, kind, b->tune[kind], launch);
}

#define TPA_DISPATCH(tpa, CALL)                      \
  switch (tpa) {                                     \
    case 1: { constexpr int TP = 1; CALL; } break;   \
    case 2: { constexpr int TP = 2; CALL; } break;   \
    case 4: { constexpr int TP = 4; CALL; } break;   \
    case 8: { constexpr int TP = 8; CALL; } break;   \
    case 16: { constexpr int TP = 16; CALL; } break; \
    case 32: { constexpr int TP = 32; CALL; } break; \
    default: { constexpr int TP = 64; CALL; } break; \
  }

int ensure_partials(mdx_batch *b, size_t doubles) {
  if (doubles <= b->partials_cap) return MDX_OK;
  b->partials_cap = doubles;
  return dalloc(&b->partials, doubles);
}

int colsum(mdx_batch *b, int rows, int cols, std::vector<double> &out) {
  const size_t need = (size_t)(kColChunks + 1) * cols;  // chunk sums, then the result
  if (need > b->colsum_cap) {
    int rc = dalloc(&b->colsum, need);
    if (rc) return rc;
    b->colsum_cap = need;
  }
  const int nchunk = std::max(1, std::min(kColChunks, rows));
  double *mid = b->colsum + cols;
  hipLaunchKernelGGL(k_b_colsum_chunks, dim3((cols + 63) / 64, nchunk), dim3(256), 0, b->stream, b->partials, rows,
                     cols, mid);
  hipLaunchKernelGGL(k_b_colsum, dim3((cols + 255) / 256), dim3(256), 0, b->stream, mid, nchunk, cols, b->colsum);
  out.assign(cols, 0.0);
  hipError_t e = hipMemcpyAsync(out.data(), b->colsum, sizeof(double) * cols, hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  return e == hipSuccess ? MDX_OK : bfail(e, "gradient reduction");
}

double pair_cutmax(mdx_batch *b) {
  if (b->pair == MDX_PAIR_LJ) return b->lj_cutmax;
  if (b->pair == MDX_PAIR_EAM) return b->eam_cut;
  return 0.0;
}

// bins for cutneigh, then wrap/sort/permute/list
int build(mdx_batch *b) {
  if (!b->positions) {
    set_error("positions not set (mdx_batch_set_positions)");
    return MDX_ERR_ARG;
  }
  const double cutneigh = pair_cutmax(b) + b->skin;
  if (!(cutneigh > 0)) {
    set_error("no pair style set");
    return MDX_ERR_ARG;
  }
  int rc;
  hipError_t e;
  if (cutneigh != b->cutneigh || !b->bin_start) {
    const double binsize = 0.5 * cutneigh;
    long total = 0;
    for (int r = 0; r < b->nrep; r++) {
      RepGeom &g = b->geom_h[r];
      for (int d = 0; d < 3; d++) {
        if (g.periodic[d] && g.prd[d] < 2.0 * cutneigh) {
          set_error("a periodic replica box is shorter than twice the neighbour cutoff");
          return MDX_ERR_ARG;
        }
        int nb = (int)std::floor(g.prd[d] / binsize);
        nb = std::max(1, std::min(nb, 1024));
        g.nbin[d] = nb;
        g.bininv[d] = nb / g.prd[d];
      }
      g.bin_off = (int)total;
      total += (long)g.nbin[0] * g.nbin[1] * g.nbin[2];
      if (total >= (1l << 30)) {
        set_error("too many bins over all replicas");
        return MDX_ERR_ARG;
      }
    }
    b->total_bins = (int)total;
    if ((e = hipMemcpy(b->geom, b->geom_h.data(), sizeof(RepGeom) * b->nrep, hipMemcpyHostToDevice)) != hipSuccess)
      return bfail(e, "replica geometry upload");
    if (b->total_bins + 1 > b->bins_cap) {
      if ((rc = dalloc(&b->bin_count, b->total_bins + 1)) || (rc = dalloc(&b->bin_start, b->total_bins + 1)))
        return rc;
      b->bins_cap = b->total_bins + 1;
    }
    b->cutneigh = cutneigh;
    int eb = 1;
    while ((1l << eb) <= b->total_bins) eb++;
    size_t need = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, need, b->key, b->key_s, b->idx, b->idx_s, b->n, 0, eb,
                                             b->stream);
    if (need > b->sort_bytes) {
      (void)hipFree(b->sort_tmp);
      b->sort_tmp = nullptr;
      if ((e = hipMalloc(&b->sort_tmp, need)) != hipSuccess) return bfail(e, "sort scratch");
      b->sort_bytes = need;
    }
    need = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, need, b->bin_count, b->bin_start, b->total_bins + 1,
                                           b->stream);
    if (need > b->scan_bytes) {
      (void)hipFree(b->scan_tmp);
      b->scan_tmp = nullptr;
      if ((e = hipMalloc(&b->scan_tmp, need)) != hipSuccess) return bfail(e, "scan scratch");
      b->scan_bytes = need;
    }
  }
  int eb = 1;
  while ((1l << eb) <= b->total_bins) eb++;
  const int n = b->n;
  if ((e = hipMemsetAsync(b->bin_count, 0, sizeof(int) * (b->total_bins + 1), b->stream)) != hipSuccess ||
      (e = hipMemsetAsync(b->dflags, 0, sizeof(int) * 16, b->stream)) != hipSuccess)
    return bfail(e, "build flags");
  hipLaunchKernelGGL(k_b_load, dim3((n + 255) / 256), dim3(256), 0, b->stream, n, b->xin, b->type_c, b->rep_c,
                     b->geom, b->xw, b->key, b->idx, b->bin_count, b->dflags);
  size_t sb = b->sort_bytes;
  if ((e = hipcub::DeviceRadixSort::SortPairs(b->sort_tmp, sb, b->key, b->key_s, b->idx, b->idx_s, n, 0, eb,
                                              b->stream)) != hipSuccess)
    return bfail(e, "bin sort");
  size_t scb = b->scan_bytes;
  if ((e = hipcub::DeviceScan::ExclusiveSum(b->scan_tmp, scb, b->bin_count, b->bin_start, b->total_bins + 1,
                                            b->stream)) != hipSuccess)
    return bfail(e, "bin scan");
  hipLaunchKernelGGL(k_b_permute, dim3((n + 255) / 256), dim3(256), 0, b->stream, n, b->idx_s, b->xw, b->rep_c,
                     b->geom, b->x, b->xq, b->rep, b->orig, b->inv);
  if (b->cap == 0) {
    b->cap = 96;
    if ((rc = dalloc(&b->nbr, (size_t)b->cap * n))) return rc;
  }
  const double cutneighsq = cutneigh * cutneigh;
  for (int attempt = 0; attempt < 4; attempt++) {
    if ((e = hipMemsetAsync(b->dflags, 0, sizeof(int), b->stream)) != hipSuccess ||
        (e = hipMemsetAsync(b->dflags + 8, 0, sizeof(unsigned long long), b->stream)) != hipSuccess)
      return bfail(e, "build flags");
    auto launch = [&](const Tune &t) {
      const int G = persistent_grid(b, t);
      TPA_DISPATCH(t.tpa, hipLaunchKernelGGL((k_b_build<TP>), dim3(G), dim3(t.block), 0, b->stream, n, b->x, b->rep,
                                             b->geom, b->bin_start, cutneighsq, b->cap, b->nbr, b->numneigh,
                                             b->dflags, (unsigned long long *)(b->dflags + 8)));
    };
    if ((rc = timed(b, MDX_BK_BUILD, b->tune[MDX_BK_BUILD], launch))) return rc;
    int fl[16];
    if ((e = hipMemcpy(fl, b->dflags, sizeof(fl), hipMemcpyDeviceToHost)) != hipSuccess)
      return bfail(e, "build flags");
    if (fl[4]) {
      set_error("an atom lies outside its replica's box in a non-periodic dimension");
      return MDX_ERR_ARG;
    }
    if (fl[0] <= b->cap) {
      unsigned long long tot;
      std::memcpy(&tot, fl + 8, sizeof(tot));
      b->total_neighbors = (long)tot;
      b->max_neighbors = fl[0];
      b->built = true;
      b->eam_rho_prec = -1;
      return MDX_OK;
    }
    b->cap = (fl[0] + 16 + 15) / 16 * 16;
    if ((rc = dalloc(&b->nbr, (size_t)b->cap * n))) return rc;
  }
  set_error("neighbour list capacity did not converge");
  return MDX_ERR_OVERFLOW;
}

int ensure_built(mdx_batch *b) {
  const double cutneigh = pair_cutmax(b) + b->skin;
  if (b->built && cutneigh == b->cutneigh) return MDX_OK;
  return build(b);
}

// copy [N][3] output from the sorted double4 array to a host or device pointer
int write_vec3(mdx_batch *b, const double4 *src, double *out) {
  const int n = b->n;
  const bool dev = is_device_ptr(out, b->device);
  double *dst = dev ? out : b->staging;
  hipLaunchKernelGGL(k_b_gather_f, dim3((n + 255) / 256), dim3(256), 0, b->stream, n, b->inv, src, dst);
  hipError_t e = hipSuccess;
  if (!dev) e = hipMemcpyAsync(out, b->staging, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  return e == hipSuccess ? MDX_OK : bfail(e, "force output");
}

int read_g(mdx_batch *b, const double *gf) {
  const int n = b->n;
  const double *src = gf;
  hipError_t e;
  if (!is_device_ptr(gf, b->device)) {
    if ((e = hipMemcpyAsync(b->staging, gf, sizeof(double) * 3 * n, hipMemcpyHostToDevice, b->stream)) !=
        hipSuccess)
      return bfail(e, "dL/dF upload");
    src = b->staging;
  }
  hipLaunchKernelGGL(k_b_scatter_g, dim3((n + 255) / 256), dim3(256), 0, b->stream, n, b->inv, src, b->g);
  return MDX_OK;
}

// transpose of the knot -> (c3..c6) map of the spline construction
// (pair_eam.cpp:723-751): g_knot[k] = sum_m,t gc[m][t] * d coeff_t(m) / d knot_k
void spline_coeffs(int n, const double *f /* [n] */, double *c /* (n+1)*4: c3..c6 */) {
  std::vector<double> c5(n + 1, 0.0);
  auto F = [&](int m) { return f[m - 1]; };
  c5[1] = F(2) - F(1);
  c5[2] = 0.5 * (F(3) - F(1));
  c5[n - 1] = 0.5 * (F(n) - F(n - 2));
  c5[n] = F(n) - F(n - 1);
  for (int m = 3; m <= n - 2; m++) c5[m] = ((F(m - 2) - F(m + 2)) + 8.0 * (F(m + 1) - F(m - 1))) / 12.0;
  for (int t = 0; t < 4 * (n + 1); t++) c[t] = 0.0;
  for (int m = 1; m <= n; m++) {
    c[4 * m + 2] = c5[m];
    c[4 * m + 3] = F(m);
  }
  for (int m = 1; m <= n - 1; m++) {
    c[4 * m + 1] = 3.0 * (F(m + 1) - F(m)) - 2.0 * c5[m] - c5[m + 1];
    c[4 * m + 0] = c5[m] + c5[m + 1] - 2.0 * (F(m + 1) - F(m));
  }
}

void knot_gradient(int n, const double *gc, double *g) {
  std::vector<double> e(n, 0.0), c(4 * (n + 1));
  for (int k = 0; k < n; k++) {
    e[k] = 1.0;
    spline_coeffs(n, e.data(), c.data());
    double acc = 0.0;
    for (int t = 4; t < 4 * (n + 1); t++) acc += gc[t] * c[t];
    g[k] = acc;
    e[k] = 0.0;
  }
}

// full 7-coefficient table from knots (host copy uploaded to the device)
void spline7(int n, double delta, const double *f, std::vector<double> &sp) {
  std::vector<double> c(4 * (n + 1));
  spline_coeffs(n, f, c.data());
  sp.assign(7 * (n + 1), 0.0);
  for (int m = 1; m <= n; m++) {
    sp[7 * m + 3] = c[4 * m + 0];
    sp[7 * m + 4] = c[4 * m + 1];
    sp[7 * m + 5] = c[4 * m + 2];
    sp[7 * m + 6] = c[4 * m + 3];
    sp[7 * m + 2] = sp[7 * m + 5] / delta;
    sp[7 * m + 1] = 2.0 * sp[7 * m + 4] / delta;
    sp[7 * m + 0] = 3.0 * sp[7 * m + 3] / delta;
  }
}

int upload(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? MDX_OK : bfail(e, "table upload");
}

}  // namespace

extern "C" {

int mdx_batch_create(int device, int nrep, const int *natoms, const double *lo, const double *hi,
                     const int *periodic, int ntypes, const int *type, mdx_batch **out) {
  if (!out || nrep <= 0 || !natoms || !lo || !hi || !periodic || !type || ntypes < 1 || ntypes > MDX_MAXT) {
    set_error("mdx_batch_create: invalid arguments (nrep > 0, 1 <= ntypes <= 8)");
    return MDX_ERR_ARG;
  }
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    (void)hipGetLastError();
    set_error("no HIP device available");
    return MDX_ERR_HIP;
  }
  long total = 0;
  for (int r = 0; r < nrep; r++) {
    if (natoms[r] < 0) {
      set_error("negative replica size");
      return MDX_ERR_ARG;
    }
    total += natoms[r];
    for (int d = 0; d < 3; d++)
      if (!(hi[3 * r + d] > lo[3 * r + d])) {
        set_error("empty replica box");
        return MDX_ERR_ARG;
      }
  }
  if (total <= 0 || total >= (1l << MDX_IDX_BITS)) {
    set_error("total atoms must be in [1, 2^27)");
    return MDX_ERR_ARG;
  }
  for (long i = 0; i < total; i++)
    if (type[i] < 1 || type[i] > ntypes) {
      set_error("atom type out of range");
      return MDX_ERR_ARG;
    }
  mdx_batch *b = new mdx_batch();
  b->device = device;
  b->nrep = nrep;
  b->ntypes = ntypes;
  b->n = (int)total;
  b->natoms.assign(natoms, natoms + nrep);
  b->off.assign(nrep + 1, 0);
  for (int r = 0; r < nrep; r++) b->off[r + 1] = b->off[r] + natoms[r];
  b->geom_h.assign(nrep, RepGeom{});
  for (int r = 0; r < nrep; r++)
    for (int d = 0; d < 3; d++) {
      b->geom_h[r].lo[d] = lo[3 * r + d];
      b->geom_h[r].prd[d] = hi[3 * r + d] - lo[3 * r + d];
      b->geom_h[r].periodic[d] = periodic[3 * r + d] ? 1 : 0;
      b->geom_h[r].nbin[d] = 1;
      b->geom_h[r].bininv[d] = 1.0 / b->geom_h[r].prd[d];
      const double span = periodic[3 * r + d] ? 4294967296.0 : 1073741824.0;
      b->geom_h[r].qscale[d] = span / b->geom_h[r].prd[d];
      b->geom_h[r].qstep[d] = b->geom_h[r].prd[d] / span;
    }
  hipError_t e;
  int rc;
  if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreate(&b->own_stream)) != hipSuccess ||
      (e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess) {
    rc = bfail(e, "batch stream");
    mdx_batch_destroy(b);
    return rc;
  }
  b->stream = b->own_stream;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    b->num_cu = prop.multiProcessorCount;
    b->lds_max = (int)prop.sharedMemPerBlock;
  }
  const int n = b->n;
  std::vector<int> rep_c(n);
  for (int r = 0; r < nrep; r++)
    for (int i = b->off[r]; i < b->off[r + 1]; i++) rep_c[i] = r;
  if ((rc = dalloc(&b->geom, nrep)) || (rc = dalloc(&b->off_d, nrep + 1)) || (rc = dalloc(&b->xin, 3 * (size_t)n)) ||
      (rc = dalloc(&b->staging, std::max<size_t>(3 * (size_t)n, nrep))) || (rc = dalloc(&b->type_c, n)) || (rc = dalloc(&b->rep_c, n)) ||
      (rc = dalloc(&b->xw, n)) || (rc = dalloc(&b->x, n)) || (rc = dalloc(&b->xq, n)) || (rc = dalloc(&b->f, n)) ||
      (rc = dalloc(&b->g, n)) ||
      (rc = dalloc(&b->rep, n)) || (rc = dalloc(&b->orig, n)) || (rc = dalloc(&b->inv, n)) ||
      (rc = dalloc(&b->key, n)) || (rc = dalloc(&b->key_s, n)) || (rc = dalloc(&b->idx, n)) ||
      (rc = dalloc(&b->idx_s, n)) || (rc = dalloc(&b->numneigh, n)) || (rc = dalloc(&b->dflags, 16)) ||
      (rc = dalloc(&b->rho, n)) || (rc = dalloc(&b->fp, n)) || (rc = dalloc(&b->eemb, n)) ||
      (rc = dalloc(&b->vadj, n)) || (rc = upload(b->geom, b->geom_h.data(), sizeof(RepGeom) * nrep)) ||
      (rc = upload(b->off_d, b->off.data(), sizeof(int) * (nrep + 1))) ||
      (rc = upload(b->type_c, type, sizeof(int) * n)) || (rc = upload(b->rep_c, rep_c.data(), sizeof(int) * n))) {
    mdx_batch_destroy(b);
    return rc;
  }
  *out = b;
  return MDX_OK;
}

int mdx_batch_destroy(mdx_batch *b) {
  if (!b) return MDX_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  void *ptrs[] = {b->xq, b->geom,    b->off_d,    b->xin,      b->staging,  b->type_c,  b->rep_c,   b->xw,      b->x,
                  b->f,       b->g,        b->rep,      b->orig,     b->inv,     b->key,     b->key_s,   b->idx,
                  b->idx_s,   b->bin_count, b->bin_start, b->sort_tmp, b->scan_tmp, b->nbr,    b->numneigh, b->dflags,
                  b->rho,     b->fp,       b->eemb,     b->vadj,     b->partials, b->colsum,  b->ljtab,   b->frho_d,
                  b->rhor_d,  b->z2r_d,    b->rhor_f,   b->z2r_f};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->own_stream) {
    (void)hipStreamSynchronize(b->own_stream);
    (void)hipStreamDestroy(b->own_stream);
  }
  delete b;
  return MDX_OK;
}

int mdx_batch_set_stream(mdx_batch *b, int external, void *stream) {
  int rc = check_b(b);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(b->stream);  // work already queued finishes where it was queued
  if (e != hipSuccess) return bfail(e, "batch stream switch");
  b->stream = external ? static_cast<hipStream_t>(stream) : b->own_stream;
  return MDX_OK;
}

int mdx_batch_set_positions(mdx_batch *b, const double *x) {
  int rc = check_b(b);
  if (rc) return rc;
  if (!x) {
    set_error("null positions");
    return MDX_ERR_ARG;
  }
  hipError_t e = hipMemcpyAsync(b->xin, x, sizeof(double) * 3 * b->n,
                                is_device_ptr(x, b->device) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return bfail(e, "positions upload");
  b->positions = true;
  b->built = false;
  b->eam_rho_prec = -1;
  return MDX_OK;
}

int mdx_batch_set_precision(mdx_batch *b, int precision) {
  int rc = check_b(b);
  if (rc) return rc;
  if (precision != MDX_FP64 && precision != MDX_MIXED) {
    set_error("precision must be MDX_FP64 or MDX_MIXED");
    return MDX_ERR_ARG;
  }
  if (precision != b->precision) {
    b->precision = precision;
    b->eam_rho_prec = -1;
    for (int k = 0; k < MDX_BK_NKIND; k++) b->tuned[k] = false;
  }
  return MDX_OK;
}

int mdx_batch_set_skin(mdx_batch *b, double skin) {
  int rc = check_b(b);
  if (rc) return rc;
  if (!(skin >= 0.0)) {
    set_error("skin must be >= 0");
    return MDX_ERR_ARG;
  }
  b->skin = skin;
  return MDX_OK;
}

int mdx_batch_lj_init(mdx_batch *b, const double *eps, const double *sigma, const double *cut) {
  int rc = check_b(b);
  if (rc) return rc;
  if (!eps || !sigma || !cut) {
    set_error("null lj tables");
    return MDX_ERR_ARG;
  }
  const int nt1 = b->ntypes + 1;
  std::vector<double> tab(T_N * kNT, 0.0);
  double cutmax = 0.0;
  for (int i = 1; i < nt1; i++)
    for (int j = 1; j < nt1; j++) {
      const int s = i * nt1 + j;
      const double ep = eps[s], sg = sigma[s], c = cut[s];
      if (!(sg > 0.0) || !(c > 0.0) || !(ep >= 0.0)) {
        set_error("lj tables need eps >= 0, sigma > 0, cut > 0 for every type pair");
        return MDX_ERR_ARG;
      }
      const int t = i * nt1 + j;  // kernel index ti*nt1 + tj
      tab[T_CUTSQ * kNT + t] = c * c;
      tab[T_LJ1 * kNT + t] = 48.0 * ep * std::pow(sg, 12.0);
      tab[T_LJ2 * kNT + t] = 24.0 * ep * std::pow(sg, 6.0);
      tab[T_LJ3 * kNT + t] = 4.0 * ep * std::pow(sg, 12.0);
      tab[T_LJ4 * kNT + t] = 4.0 * ep * std::pow(sg, 6.0);
      tab[T_A1 * kNT + t] = 48.0 * std::pow(sg, 12.0);
      tab[T_A2 * kNT + t] = 24.0 * std::pow(sg, 6.0);
      tab[T_B1 * kNT + t] = 576.0 * ep * std::pow(sg, 11.0);
      tab[T_B2 * kNT + t] = 144.0 * ep * std::pow(sg, 5.0);
      cutmax = std::max(cutmax, c);
    }
  if (b->pair == MDX_PAIR_LJ && b->ljtab && tab == b->lj_tab) return MDX_OK;
  if (!b->ljtab && (rc = dalloc(&b->ljtab, T_N * kNT))) return rc;
  if ((rc = upload(b->ljtab, tab.data(), sizeof(double) * T_N * kNT))) return rc;
  b->lj_tab = tab;
  b->pair = MDX_PAIR_LJ;
  b->lj_cutmax = cutmax;
  return MDX_OK;
}

int mdx_batch_lj_forces(mdx_batch *b, double *f, double *energy) {
  int rc = check_b(b);
  if (rc) return rc;
  if (b->pair != MDX_PAIR_LJ) {
    set_error("lj/cut not initialised (mdx_batch_lj_init)");
    return MDX_ERR_ARG;
  }
  if ((rc = ensure_built(b))) return rc;
  const int nt1 = b->ntypes + 1;
  const bool mixed = b->precision == MDX_MIXED;
  auto launch = [&](const Tune &t) {
    const int G = persistent_grid(b, t);
    if (mixed)
      TPA_DISPATCH(t.tpa, hipLaunchKernelGGL((k_b_lj<true, TP>), dim3(G), dim3(t.block), 0, b->stream, b->n, nt1,
                                             b->x, b->xq, b->rep, b->geom, b->nbr, b->numneigh, b->cap, b->ljtab,
                                             b->f))
    else
      TPA_DISPATCH(t.tpa, hipLaunchKernelGGL((k_b_lj<false, TP>), dim3(G), dim3(t.block), 0, b->stream, b->n, nt1,
                                             b->x, b->xq, b->rep, b->geom, b->nbr, b->numneigh, b->cap, b->ljtab,
                                             b->f))
  };
  if ((rc = run_kind(b, MDX_BK_LJ, launch))) return rc;
  if (energy) {
    double *ed = b->staging;
    hipLaunchKernelGGL(k_b_rep_energy, dim3(b->nrep), dim3(256), 0, b->stream, b->f, b->off_d, ed);
    hipError_t e = hipMemcpyAsync(energy, ed, sizeof(double) * b->nrep, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bfail(e, "energy output");
  }
  if (f && (rc = write_vec3(b, b->f, f))) return rc;
  return MDX_OK;
}

int mdx_batch_lj_vjp(mdx_batch *b, const double *gf, double *geps, double *gsigma) {
  int rc = check_b(b);
  if (rc) return rc;
  if (b->pair != MDX_PAIR_LJ) {
    set_error("lj/cut not initialised (mdx_batch_lj_init)");
    return MDX_ERR_ARG;
  }
  if (!gf || !geps || !gsigma) {
    set_error("null argument");
    return MDX_ERR_ARG;
  }
  if ((rc = ensure_built(b)) || (rc = read_g(b, gf))) return rc;
  const int nt1 = b->ntypes + 1;
  // the adjoint's per-pair factors (48 s^12 r^-6 - 24 s^6 and the sigma
  // term) cancel near the potential minimum, where FP32 loses ~5e-6
  // relative: the adjoint runs FP64 arithmetic in both precision modes
  int grid_used = 0;
  auto launch = [&](const Tune &t) {
    const int G = persistent_grid(b, t, true);
    grid_used = G;
    TPA_DISPATCH(t.tpa, hipLaunchKernelGGL((k_b_lj_vjp<false, TP>), dim3(G), dim3(t.block), 0, b->stream, b->n, nt1,
                                           b->x, b->rep, b->geom, b->nbr, b->numneigh, b->cap, b->ljtab, b->g,
                                           b->partials))
  };
  // partial rows for the largest grid any candidate can use
  const long maxgrid = (long)kPartialBlocksPerCU * b->num_cu;
  if ((rc = ensure_partials(b, (size_t)maxgrid * 2 * kNT))) return rc;
  if ((rc = run_kind(b, MDX_BK_LJ_VJP, launch))) return rc;
  std::vector<double> sums;
  if ((rc = colsum(b, grid_used, 2 * kNT, sums))) return rc;
  for (int k = 0; k < nt1 * nt1; k++) geps[k] = gsigma[k] = 0.0;
  for (int a = 1; a < nt1; a++)
    for (int c = 1; c < nt1; c++) {
      const int lo = std::min(a, c), hi = std::max(a, c);
      geps[lo * nt1 + hi] += sums[a * nt1 + c];
      gsigma[lo * nt1 + hi] += sums[kNT + a * nt1 + c];
    }
  return MDX_OK;
}

int mdx_batch_eam_init(mdx_batch *b, int nrho, double drho, int nr, double dr, double cut, const double *frho,
                       const double *rhor, const double *z2r) {
  int rc = check_b(b);
  if (rc) return rc;
  if (nrho < 5 || nr < 5 || !(drho > 0) || !(dr > 0) || !(cut > 0) || !frho || !rhor || !z2r) {
    set_error("mdx_batch_eam_init: invalid tables (at least 5 knots, positive spacings and cutoff)");
    return MDX_ERR_ARG;
  }
  if (b->pair == MDX_PAIR_EAM && b->et.nrho == nrho && b->et.nr == nr && b->eam_drho == drho && b->eam_dr == dr &&
      b->eam_cut == cut && std::equal(frho, frho + nrho, b->eam_knots[0].begin()) &&
      std::equal(rhor, rhor + nr, b->eam_knots[1].begin()) && std::equal(z2r, z2r + nr, b->eam_knots[2].begin()))
    return MDX_OK;  // same tables: keep rho/fp of the last forward pass
  std::vector<double> sf, sr, sz;
  spline7(nrho, drho, frho, sf);
  spline7(nr, dr, rhor, sr);
  spline7(nr, dr, z2r, sz);
  std::vector<float> fr(sr.begin(), sr.end()), fz(sz.begin(), sz.end());
  if (b->et.nrho != nrho || !b->frho_d) {
    if ((rc = dalloc(&b->frho_d, sf.size()))) return rc;
  }
  if (b->et.nr != nr || !b->rhor_d) {
    if ((rc = dalloc(&b->rhor_d, sr.size())) || (rc = dalloc(&b->z2r_d, sz.size())) ||
        (rc = dalloc(&b->rhor_f, fr.size())) || (rc = dalloc(&b->z2r_f, fz.size())))
      return rc;
  }
  if ((rc = upload(b->frho_d, sf.data(), sizeof(double) * sf.size())) ||
      (rc = upload(b->rhor_d, sr.data(), sizeof(double) * sr.size())) ||
      (rc = upload(b->z2r_d, sz.data(), sizeof(double) * sz.size())) ||
      (rc = upload(b->rhor_f, fr.data(), sizeof(float) * fr.size())) ||
      (rc = upload(b->z2r_f, fz.data(), sizeof(float) * fz.size())))
    return rc;
  b->et.nrho = nrho;
  b->et.nr = nr;
  b->et.rdrho = 1.0 / drho;
  b->et.rdr = 1.0 / dr;
  b->et.rhomax = (nrho - 1) * drho;
  b->et.cutsq = cut * cut;
  b->eam_knots[0].assign(frho, frho + nrho);
  b->eam_knots[1].assign(rhor, rhor + nr);
  b->eam_knots[2].assign(z2r, z2r + nr);
  b->eam_drho = drho;
  b->eam_dr = dr;
  b->eam_cut = cut;
  b->pair = MDX_PAIR_EAM;
  b->eam_rho_prec = -1;
  return MDX_OK;
}

namespace {
// rho + fp pass in `precision`, then (with_force) the force pass
int eam_forward(mdx_batch *b, bool with_force, int precision) {
  int rc;
  if ((rc = ensure_built(b))) return rc;
  const bool mixed = precision == MDX_MIXED;
  const int tn = (b->et.nr + 1) * 7;
  const size_t esz = mixed ? sizeof(float) : sizeof(double);
  const bool lds1 = (size_t)tn * esz <= (size_t)kTabLDS * 8;
  const bool lds2 = (size_t)2 * tn * esz <= (size_t)kTabLDS * 8;
  const size_t sh1 = lds1 ? tn * esz : 0, sh2 = lds2 ? 2 * tn * esz : 0;
  auto rho_launch = [&](const Tune &t) {
    const int G = persistent_grid(b, t);
    if (mixed)
      TPA_DISPATCH(t.tpa, hipLaunchKernelGGL((k_b_eam_rho<true, TP>), dim3(G), dim3(t.block), sh1, b->stream, b->n,
                                             b->x, b->xq, b->rep, b->geom, b->nbr, b->numneigh, b->cap, b->et, b->rhor_d,
                                             b->rhor_f, b->frho_d, lds1 ? 1 : 0, b->rho, b->fp, b->eemb))
    else
      TPA_DISPATCH(t.tpa, hipLaunchKernelGGL((k_b_eam_rho<false, TP>), dim3(G), dim3(t.block), sh1, b->stream, b->n,
                                             b->x, b->xq, b->rep, b->geom, b->nbr, b->numneigh, b->cap, b->et, b->rhor_d,
                                             b->rhor_f, b->frho_d, lds1 ? 1 : 0, b->rho, b->fp, b->eemb))
  };
  if ((rc = run_kind(b, MDX_BK_EAM_RHO, rho_launch))) return rc;
  b->eam_rho_prec = precision;
  if (!with_force) return MDX_OK;
  auto force_launch = [&](const Tune &t) {
    const int G = persistent_grid(b, t);
    if (mixed)
      TPA_DISPATCH(t.tpa, hipLaunchKernelGGL((k_b_eam_force<true, TP>), dim3(G), dim3(t.block), sh2, b->stream, b->n,
                                             b->x, b->xq, b->rep, b->geom, b->nbr, b->numneigh, b->cap, b->et, b->rhor_d,
                                             b->z2r_d, b->rhor_f, b->z2r_f, lds2 ? 1 : 0, b->fp, b->eemb, b->f))
    else
      TPA_DISPATCH(t.tpa, hipLaunchKernelGGL((k_b_eam_force<false, TP>), dim3(G), dim3(t.block), sh2, b->stream,
                                             b->n, b->x, b->xq, b->rep, b->geom, b->nbr, b->numneigh, b->cap, b->et,
                                             b->rhor_d, b->z2r_d, b->rhor_f, b->z2r_f, lds2 ? 1 : 0, b->fp, b->eemb,
                                             b->f))
  };
  return run_kind(b, MDX_BK_EAM_FORCE, force_launch);
}
}  // namespace

int mdx_batch_eam_forces(mdx_batch *b, double *f, double *energy, double *rho, double *fp) {
  int rc = check_b(b);
  if (rc) return rc;
  if (b->pair != MDX_PAIR_EAM) {
    set_error("eam not initialised (mdx_batch_eam_init)");
    return MDX_ERR_ARG;
  }
  if ((rc = eam_forward(b, true, b->precision))) return rc;
  hipError_t e = hipSuccess;
  if (energy) {
    hipLaunchKernelGGL(k_b_rep_energy, dim3(b->nrep), dim3(256), 0, b->stream, b->f, b->off_d, b->staging);
    e = hipMemcpyAsync(energy, b->staging, sizeof(double) * b->nrep, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bfail(e, "energy output");
  }
  for (int k = 0; k < 2; k++) {
    double *out = k == 0 ? rho : fp;
    if (!out) continue;
    hipLaunchKernelGGL(k_b_gather_s, dim3((b->n + 255) / 256), dim3(256), 0, b->stream, b->n, b->inv,
                       k == 0 ? b->rho : b->fp, b->staging);
    e = hipMemcpyAsync(out, b->staging, sizeof(double) * b->n, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bfail(e, "rho/fp output");
  }
  if (f && (rc = write_vec3(b, b->f, f))) return rc;
  return MDX_OK;
}

int mdx_batch_eam_vjp(mdx_batch *b, const double *gf, double *g_frho, double *g_rhor, double *g_z2r) {
  int rc = check_b(b);
  if (rc) return rc;
  if (b->pair != MDX_PAIR_EAM) {
    set_error("eam not initialised (mdx_batch_eam_init)");
    return MDX_ERR_ARG;
  }
  if (!gf || !g_frho || !g_rhor || !g_z2r) {
    set_error("null argument");
    return MDX_ERR_ARG;
  }
  if ((rc = ensure_built(b))) return rc;
  // the adjoint uses FP64 rho/fp (a mixed forward's rho moves the frho
  // spline weights by ~1e-5 relative): rerun the rho pass in FP64 if needed
  if (b->eam_rho_prec != MDX_FP64 && (rc = eam_forward(b, false, MDX_FP64))) return rc;
  if ((rc = read_g(b, gf))) return rc;
  // FP64 adjoint arithmetic in both precision modes, as for lj/cut
  const int P = 4 * (b->et.nrho + 1) + 8 * (b->et.nr + 1);
  const size_t shA = sizeof(double) * P, shB = sizeof(double) * P;
  if ((long)shA > b->lds_max) {
    set_error("eam tables too large for the adjoint's LDS accumulators");
    return MDX_ERR_ARG;
  }
  const long maxgrid = (long)kPartialBlocksPerCU * b->num_cu;
  if ((rc = ensure_partials(b, (size_t)2 * maxgrid * P))) return rc;
  int gA = 0, gB = 0;
  auto launchA = [&](const Tune &t) {
    const int G = persistent_grid(b, t, true);
    gA = G;
    TPA_DISPATCH(t.tpa, hipLaunchKernelGGL((k_b_eam_adj_a<false, TP>), dim3(G), dim3(t.block), shA, b->stream, b->n,
                                           b->x, b->rep, b->geom, b->nbr, b->numneigh, b->cap, b->et, b->rhor_d,
                                           b->rhor_f, b->frho_d, b->rho, b->fp, b->g, b->vadj, b->partials))
  };
  if ((rc = run_kind(b, MDX_BK_EAM_ADJ_A, launchA))) return rc;
  double *partB = b->partials + (size_t)gA * P;
  auto launchB = [&](const Tune &t) {
    const int G = persistent_grid(b, t, true);
    gB = G;
    TPA_DISPATCH(t.tpa, hipLaunchKernelGGL((k_b_eam_adj_b<false, TP>), dim3(G), dim3(t.block), shB, b->stream, b->n,
                                           b->x, b->rep, b->geom, b->nbr, b->numneigh, b->cap, b->et, b->vadj, partB))
  };
  if ((rc = run_kind(b, MDX_BK_EAM_ADJ_B, launchB))) return rc;
  std::vector<double> gc;
  if ((rc = colsum(b, gA + gB, P, gc))) return rc;
  const int nrho = b->et.nrho, nr = b->et.nr;
  knot_gradient(nrho, gc.data(), g_frho);
  knot_gradient(nr, gc.data() + 4 * (nrho + 1), g_rhor);
  knot_gradient(nr, gc.data() + 4 * (nrho + 1) + 4 * (nr + 1), g_z2r);
  return MDX_OK;
}

int mdx_batch_set_launch(mdx_batch *b, int kind, int block, int tpa, int grid) {
  int rc = check_b(b);
  if (rc) return rc;
  if (kind < 0 || kind >= MDX_BK_NKIND || block < 64 || block > 1024 || block % 64 || tpa < 1 || tpa > 64 ||
      (tpa & (tpa - 1)) || grid < 0) {
    set_error("invalid launch shape (block 64..1024 multiple of 64, tpa power of two <= 64, grid >= 0)");
    return MDX_ERR_ARG;
  }
  b->tune[kind].block = block;
  b->tune[kind].tpa = tpa;
  b->tune[kind].grid = grid;
  b->tuned[kind] = true;
  return MDX_OK;
}

int mdx_batch_get_launch(mdx_batch *b, int kind, int *block, int *tpa, int *grid) {
  int rc = check_b(b);
  if (rc) return rc;
  if (kind < 0 || kind >= MDX_BK_NKIND) {
    set_error("invalid kernel kind");
    return MDX_ERR_ARG;
  }
  if (block) *block = b->tune[kind].block;
  if (tpa) *tpa = b->tune[kind].tpa;
  if (grid) *grid = b->tune[kind].grid;
  return MDX_OK;
}

int mdx_batch_autotune(mdx_batch *b, int on) {
  int rc = check_b(b);
  if (rc) return rc;
  b->autotune = on != 0;
  if (b->autotune)
    for (int k = 0; k < MDX_BK_NKIND; k++) b->tuned[k] = false;
  return MDX_OK;
}

int mdx_batch_kernel_time(mdx_batch *b, int kind, double *ms, long *launches) {
  int rc = check_b(b);
  if (rc) return rc;
  if (kind < 0 || kind >= MDX_BK_NKIND) {
    set_error("invalid kernel kind");
    return MDX_ERR_ARG;
  }
  if (ms) *ms = b->kms[kind];
  if (launches) *launches = b->klaunch[kind];
  return MDX_OK;
}

int mdx_batch_neigh_stats(mdx_batch *b, long *total, int *max_per_atom) {
  int rc = check_b(b);
  if (rc) return rc;
  if ((rc = ensure_built(b))) return rc;
  if (total) *total = b->total_neighbors;
  if (max_per_atom) *max_per_atom = b->max_neighbors;
  return MDX_OK;
}

}  // extern "C"
